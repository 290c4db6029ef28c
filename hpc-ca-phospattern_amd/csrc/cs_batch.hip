/*
 * cs_batch.hip -- ensembles of independent realisations on one MI355X.
 *
 * The reference runs one realisation per process: the seed is set once
 * (calcium_sparks.c:113-116) and the loop (calcium_sparks.c:190-228)
 * advances that one domain.  An ensemble here is many such domains of
 * uneven sizes and channel layouts that share one species model, stepped
 * together:
 *
 *   k_batch_stencil  one persistent launch over a packed list of column
 *                    tiles of every realisation: the one-step per-voxel
 *                    update of the single-domain pass (stencil_columns:
 *                    stencil, reactions, Dirichlet faces, fixed species),
 *                    tiles taken from eight per-XCD work queues so that
 *                    neighbouring tiles of a realisation share an L2;
 *   k_batch_tiles    the same update with one block per tile, the tiles
 *                    numbered in groups per XCD (the hardware schedules
 *                    the blocks; no claims, no queues);
 *   k_batch_bridge   one wavefront per realisation: the boundary fluxes of
 *                    step t added to the new state (k_flux's rows and
 *                    order) and then the gating of step t+1 from that
 *                    state (k_gate's gate_one), so a step is two launches.
 *
 * Fields of realisation i live at their own offset of one stacked buffer
 * (two of them ping-pong), each in the single-domain layout [z][species]
 * [y][x] of deviations from rest, so every per-voxel formula, the gating
 * key (seed, instance) and the flux row order are those of a cs_context
 * built from the same cs_problem: an ensemble member reproduces its own
 * single-domain run.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstdlib>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "csparks.h"
#include "cs_device.h"
#include "cs_mlp.h"

namespace {

constexpr int BRIDGE_BATCH_THREADS = 64;   /* one wavefront per realisation */
constexpr int BSTAT_THREADS = 256;
constexpr int NQUEUE = 8;                  /* one work queue per XCD        */

/* Geometry of one realisation inside the stacked buffers. */
struct BatchGeo {
  uint64_t u_off;    /* element offset of its fields (z-group 0, species 0) */
  uint64_t id_off;   /* byte offset of its domain ids                       */
  int nx, ny, nz;
  uint32_t ns;
  uint64_t plane, nv;
};

/* A tile of the stencil pass: `cols` consecutive plane columns of one
 * realisation from column col0, planes [z0, z1). */
struct BatchTile {
  uint32_t inst, col0;
  uint32_t z0, z1;
};

/* Per-realisation records of the bridge. */
struct BatchInst {
  uint32_t chan_off, nchan;   /* its channels in BatchChan                  */
  uint32_t row_off, nrows;    /* its rows in BRow                           */
  uint32_t seed, instance;    /* Philox key                                 */
  uint32_t uni_off;           /* first of its draws within a gating step    */
  uint32_t pad;
};

/* A channel of a discrete flux: the stored species 0 at its cleft voxel
 * (element of the stacked buffer) and its resting value, its flux and
 * index within the flux (the Philox counter), its state slot, and the
 * index of its draw among the realisation's draws of a gating step
 * (NO_ROW for a channel with a forced closing time, which draws none). */
struct BatchChan {
  uint64_t elem_to;
  double rest_to;
  uint32_t flux, ch, state, uni;
};

/* Where a channel sits, for the closure's features: its realisation, the
 * grid coordinates of its cleft (to) voxel, and the stored species 0 at
 * its SR (from) voxel with that voxel's resting value. */
struct BatchChanGeo {
  uint32_t inst;
  int x, y, z;
  uint64_t elem_from;
  double rest_from;
};

/* Neighbour entries of a channel: state slot of a channel of the same
 * realisation whose cleft voxel lies within Chebyshev distance 2 in the
 * cleft plane (itself included); bit 31 set when the distance is 2. */
constexpr uint32_t NB_FAR = 0x80000000u;

/* Step t's fluxes (fluxmask, from u into un) then gating of step gate_ti
 * from un (when gate is set), for every active realisation.  With ti_base
 * set (a captured step), the gating step is *ti_base + gate_ti. */
struct BridgeArgs {
  uint32_t fluxmask;
  uint32_t gate;
  uint64_t gate_ti;
  uint64_t uni_base;          /* first draw of this gating step            */
  const uint64_t* ti_base;
  uint32_t D;                 /* stochastic substep                          */
  uint32_t pad;
};

/* FM: the step's diffusing species fixed at compile time (with
 * species-major reactions), or BS_DYN for the mask and reaction mode taken
 * at run time. */
constexpr uint32_t BS_DYN = 0xFFFFFFFFu;

template <typename T, int NS, int PF, bool NT, uint32_t FM = BS_DYN>
__global__ void __launch_bounds__(BX * 8)
k_batch_stencil(const T* __restrict__ u, T* __restrict__ un, const uint8_t* __restrict__ ids,
                const DevParams<T>* __restrict__ P, const BatchGeo* __restrict__ geos,
                const BatchTile* __restrict__ tiles, uint32_t ntiles, uint32_t dmask_rt, uint32_t react_rt,
                uint32_t* __restrict__ queue /* [NQUEUE + 1], zero between launches */)
{
  const uint32_t dmask = FM != BS_DYN ? FM : dmask_rt, react = FM != BS_DYN ? 2u : react_rt;
  __shared__ BlockTables<T, NS> S;
  __shared__ uint32_t s_tile;
  const int nthr = BX * blockDim.y;
  const int tid = threadIdx.y * BX + threadIdx.x;
  S.load(P, tid, nthr);
  /* queue q owns tiles [q*per, (q+1)*per); blocks start on the queue of
   * their XCD (hardware dispatch is round-robin over the 8 XCDs) and move
   * on to the others when it runs dry */
  const uint32_t per = (ntiles + NQUEUE - 1) / NQUEUE;
  const uint32_t home = blockIdx.x % NQUEUE;
  for (uint32_t k = 0; k < NQUEUE; k++) {
    const uint32_t q = (home + k) % NQUEUE;
    const uint32_t lo = q * per, hi = min(ntiles, lo + per);
    for (;;) {
      __syncthreads();   /* s_tile of the previous tile has been read; S is loaded */
      if (tid == 0) s_tile = lo + atomicAdd(&queue[q], 1u);
      __syncthreads();
      const uint32_t t = s_tile;
      if (t >= hi) break;
      const BatchTile tl = tiles[t];
      const BatchGeo gg = geos[tl.inst];
      const Geo g{gg.nx, gg.ny, gg.nz, gg.ns, gg.plane, gg.nv};
      const uint32_t c = tl.col0 + (uint32_t)tid;
      if (c < gg.plane) {
        const int y = (int)(c / (uint32_t)g.nx), x = (int)(c - (uint32_t)y * (uint32_t)g.nx);
        stencil_columns<T, NS, PF, NT>(u + gg.u_off, un + gg.u_off, ids + gg.id_off, P, S, g, dmask, react, x, y,
                                       (int)tl.z0, (int)tl.z1);
      }
    }
  }
  /* the last block to finish empties the queues for the next launch (no
   * block of this launch touches them after counting itself done) */
  if (tid == 0) {
    __threadfence();
    if (atomicAdd(&queue[NQUEUE], 1u) == gridDim.x - 1) {
      for (int q = 0; q <= NQUEUE; q++) atomicExch(&queue[q], 0u);
    }
  }
}

/* The stencil pass with one block per tile: block b takes tile
 * xcd_tile(b) (groups of xcd consecutive tiles per XCD), so a tile costs
 * no queue claim and no barrier before its first loads, and the hardware
 * overlaps one block's start with the others' planes.  Each column's
 * update is k_batch_stencil's, bit for bit. */
template <typename T, int NS, int PF, bool NT, uint32_t FM = BS_DYN>
__global__ void __launch_bounds__(BX * 8)
k_batch_tiles(const T* __restrict__ u, T* __restrict__ un, const uint8_t* __restrict__ ids,
              const DevParams<T>* __restrict__ P, const BatchGeo* __restrict__ geos,
              const BatchTile* __restrict__ tiles, uint32_t ntiles, uint32_t dmask_rt, uint32_t react_rt, int xcd)
{
  const uint32_t dmask = FM != BS_DYN ? FM : dmask_rt, react = FM != BS_DYN ? 2u : react_rt;
  const int t = xcd_tile((int)blockIdx.x, (int)ntiles, xcd);
  if (t >= (int)ntiles) return;   /* whole padding blocks only: no barrier is skipped */
  __shared__ BlockTables<T, NS> S;
  const int nthr = BX * blockDim.y;
  const int tid = threadIdx.y * BX + threadIdx.x;
  const BatchTile tl = tiles[t];
  const BatchGeo gg = geos[tl.inst];
  S.load(P, tid, nthr);
  __syncthreads();
  const Geo g{gg.nx, gg.ny, gg.nz, gg.ns, gg.plane, gg.nv};
  const uint32_t c = tl.col0 + (uint32_t)tid;
  if (c < gg.plane) {
    const int y = (int)(c / (uint32_t)g.nx), x = (int)(c - (uint32_t)y * (uint32_t)g.nx);
    stencil_columns<T, NS, PF, NT>(u + gg.u_off, un + gg.u_off, ids + gg.id_off, P, S, g, dmask, react, x, y,
                                   (int)tl.z0, (int)tl.z1);
  }
}

/* One wavefront per active realisation: fluxes of a step, then gating of
 * the next.  The flux rows, their entry order and skip rules are k_flux's
 * (a row's sum in double, added once); the gating is gate_one on the
 * absolute Ca at each channel's cleft voxel, read after the fluxes, or on
 * the closure's prediction (cpred).  fpred, when given, scales each
 * channel's flux by the flux closure's ratio of the fine-grid release to
 * the coarse one. */
template <typename T>
__global__ void __launch_bounds__(BRIDGE_BATCH_THREADS)
k_batch_bridge(const T* __restrict__ u, T* __restrict__ un, const DevParams<T>* __restrict__ P,
               const BatchInst* __restrict__ insts, const uint8_t* __restrict__ active,
               const BatchChan* __restrict__ chans, const BRow* __restrict__ rows, const BEnt* __restrict__ ents,
               uint8_t* __restrict__ states, uint8_t* __restrict__ recruited, uint64_t* __restrict__ last_open,
               BridgeArgs a, const double* __restrict__ uniforms, const double* __restrict__ cpred,
               const double* __restrict__ fpred)
{
  const uint32_t i = blockIdx.x;
  if (active && !active[i]) return;
  const BatchInst in = insts[i];
  const double dt = P->dt, h = P->h;
  if (a.fluxmask) {
    for (uint32_t r = in.row_off + threadIdx.x; r < in.row_off + in.nrows; r += blockDim.x) {
      const BRow w = rows[r];
      if (w.skip) continue;
      double sum = 0.0;
      bool any = false;
      for (uint32_t e = w.start; e < w.end; e++) {
        const BEnt x = ents[e];
        const uint32_t f = x.flux_sign & 0xFFu;
        if (!((a.fluxmask >> f) & 1u)) continue;
        if (x.state != NO_ROW && !states[x.state]) continue;
        double J = flux_j(P->flux[f], (double)u[x.elem_from], (double)u[x.elem_to], x.rest_from, x.rest_to,
                          dt, h);
        if (fpred && x.state != NO_ROW) J *= fpred[x.state];
        sum += (x.flux_sign >> 8) ? J : -J;
        any = true;
      }
      if (any) un[w.elem] = (T)((double)un[w.elem] + sum);
    }
  }
  if (!a.gate) return;
  __syncthreads();   /* the fluxes are in un before gating reads it */
  const uint64_t gti = (a.ti_base ? *a.ti_base : 0) + a.gate_ti;
  const double t = (double)gti * dt, dts = dt * a.D;
  uint32_t open = 0;
  for (uint32_t c = threadIdx.x; c < in.nchan; c += blockDim.x) {
    const BatchChan b = chans[in.chan_off + c];
    const FluxParams& F = P->flux[b.flux];
    uint8_t sv = states[b.state];
    /* the cleft voxel's Ca, or the closure's prediction for the mouth */
    const double c0 = !F.stochastic ? 0.0 : cpred ? cpred[in.chan_off + c] : (double)un[b.elem_to] + b.rest_to;
    const double* draw = (uniforms && b.uni != NO_ROW) ? uniforms + a.uni_base + in.uni_off + b.uni : nullptr;
    sv = gate_one(F, sv, c0, gti, t, dts, in.seed, in.instance, draw, b.ch);
    states[b.state] = sv;
    if (sv) {
      recruited[b.state] = 1;
      open++;
    }
  }
  /* the latest gating step that left a channel of the realisation open */
  for (int off = 32; off > 0; off >>= 1) open += __shfl_xor(open, off, 64);
  if (threadIdx.x == 0 && open) last_open[i] = gti;
}

/* The closure's raw input k of channel c: the absolute Ca of the 3x3x3
 * coarse voxels around its cleft voxel (k = 9 (dz+1) + 3 (dy+1) + dx+1,
 * clamped at the grid faces), the absolute Ca of its SR voxel (27), and
 * the open channels within Chebyshev distance 1 (28) and 2 (29) in the
 * cleft plane, itself included; 0 for the padding inputs 30, 31. */
template <typename T>
__device__ inline float closure_raw(int k, uint32_t c, const BatchChanGeo& cg, const BatchGeo& gg,
                                    const T* __restrict__ u, const uint8_t* __restrict__ ids,
                                    const DevParams<T>* __restrict__ P, const uint8_t* __restrict__ states,
                                    const uint32_t* __restrict__ nb_start, const uint32_t* __restrict__ nb)
{
  if (k < 27) {
    const int x = min(max(cg.x + k % 3 - 1, 0), gg.nx - 1);
    const int y = min(max(cg.y + (k / 3) % 3 - 1, 0), gg.ny - 1);
    const int z = min(max(cg.z + k / 9 - 1, 0), gg.nz - 1);
    const uint64_t col = (uint64_t)y * gg.nx + x;
    const int d = ids[gg.id_off + (uint64_t)z * gg.plane + col] & 7;
    return (float)((double)u[gg.u_off + (uint64_t)z * gg.ns * gg.plane + col] + P->init64[d][0]);
  }
  if (k == 27) return (float)((double)u[cg.elem_from] + cg.rest_from);
  if (k == 28 || k == 29) {
    uint32_t n = 0;
    for (uint32_t e = nb_start[c]; e < nb_start[c + 1]; e++) {
      const uint32_t v = nb[e];
      if (k == 29 || !(v & NB_FAR)) n += states[v & ~NB_FAR];
    }
    return (float)n;
  }
  return 0.0f;
}

/* Feature k of a raw input: the logarithm of a concentration (k < 28),
 * a count as it is (28, 29). */
__device__ inline float closure_feature(int k, float raw)
{
  return k < 28 ? logf(fmaxf(raw, 1e-6f)) : raw;
}

/* Normalised input of feature f: the training set's standardisation (by
 * the reciprocal of the deviation, within an FP32 rounding of the
 * quotient); flags an input outside the monitored range. */
__device__ inline float closure_input(const MlpDev* __restrict__ N, int k, float f, bool& out_of_range)
{
  if (k >= MLP_IN) return 0.0f;
  const float x = (f - N->in_mean[k]) * __builtin_amdgcn_rcpf(N->in_std[k]);
  if (x < N->in_lo[k] - N->margin || x > N->in_hi[k] + N->margin) out_of_range = true;
  return x;
}

/* Raw closure inputs of every channel of the ensemble, [channel][30]. */
template <typename T>
__global__ void __launch_bounds__(256)
k_closure_features(const T* __restrict__ u, const uint8_t* __restrict__ ids, const DevParams<T>* __restrict__ P,
                   const BatchGeo* __restrict__ geos, const BatchChanGeo* __restrict__ cgeo,
                   const uint8_t* __restrict__ states, const uint32_t* __restrict__ nb_start,
                   const uint32_t* __restrict__ nb, uint32_t nchan, float* __restrict__ out)
{
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (uint64_t)nchan * MLP_IN) return;
  const uint32_t c = (uint32_t)(k / MLP_IN);
  const BatchChanGeo cg = cgeo[c];
  out[k] = closure_raw<T>((int)(k % MLP_IN), c, cg, geos[cg.inst], u, ids, P, states, nb_start, nb);
}

/* Absolute Ca at every channel's cleft voxel (the concentration gating
 * reads, species.c:2569). */
template <typename T>
__global__ void k_channel_ca(const T* __restrict__ u, const BatchChan* __restrict__ chans, uint32_t nchan,
                             double* __restrict__ out)
{
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < nchan) out[c] = (double)u[chans[c].elem_to] + chans[c].rest_to;
}

/* Absolute Ca at every channel's SR (from) voxel. */
template <typename T>
__global__ void k_channel_sr_ca(const T* __restrict__ u, const BatchChanGeo* __restrict__ cgeo, uint32_t nchan,
                                double* __restrict__ out)
{
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < nchan) out[c] = (double)u[cgeo[c].elem_from] + cgeo[c].rest_from;
}

/* The closure's predictions for every channel: 32 channels per wavefront
 * (lane l%32; both lane groups gather half of the inputs each, once), the
 * gating network N on the matrix cores gives the fine-grid Ca at each
 * channel mouth, which the bridge then gates on instead of the coarse
 * cleft voxel's Ca, and the flux network F (if given), on the same raw
 * inputs, the ratio of the fine release to the coarse one.  A channel
 * whose inputs, normalised for N, leave N's training range counts in its
 * realisation's monitor. */
template <typename T>
__global__ void __launch_bounds__(256)
k_closure_predict(const T* __restrict__ u, const uint8_t* __restrict__ ids, const DevParams<T>* __restrict__ P,
                  const MlpDev* __restrict__ N, const MlpDev* __restrict__ F, const BatchGeo* __restrict__ geos,
                  const uint8_t* __restrict__ active, const BatchChanGeo* __restrict__ cgeo,
                  const uint8_t* __restrict__ states, const uint32_t* __restrict__ nb_start,
                  const uint32_t* __restrict__ nb, uint32_t nchan, double* __restrict__ cpred,
                  double* __restrict__ fpred, uint32_t* __restrict__ flags)
{
  const int lane = threadIdx.x & 63, g = lane >> 5;
  const uint64_t tile = (uint64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  if (tile * 32 >= nchan) return;   /* whole wavefront */
  const uint32_t c = (uint32_t)(tile * 32 + (lane & 31));
  const bool valid = c < nchan;
  const uint32_t cc = valid ? c : 0;
  const BatchChanGeo cg = cgeo[cc];
  const BatchGeo gg = geos[cg.inst];
  /* the open-channel counts (inputs 28, 29, in the upper lane group):
   * one pass over the neighbour list for both, four entries at a time so
   * that four state loads are in flight */
  uint32_t n1 = 0, n2 = 0;
  if (valid && g == 1) {
    const uint32_t e1 = nb_start[cc + 1];
    uint32_t e = nb_start[cc];
    for (; e + 4 <= e1; e += 4) {
      uint32_t v[4], st[4];
#pragma unroll
      for (int q = 0; q < 4; q++) v[q] = nb[e + q];
#pragma unroll
      for (int q = 0; q < 4; q++) st[q] = states[v[q] & ~NB_FAR];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        n2 += st[q];
        n1 += (v[q] & NB_FAR) ? 0u : st[q];
      }
    }
    for (; e < e1; e++) {
      const uint32_t v = nb[e], st = states[v & ~NB_FAR];
      n2 += st;
      n1 += (v & NB_FAR) ? 0u : st;
    }
  }
  float feat[2][8];
#pragma unroll
  for (int ch = 0; ch < 2; ch++)
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int k = 16 * ch + 8 * g + j;
      feat[ch][j] = !valid ? 0.0f
                 : k == 28 ? (float)n1
                 : k == 29 ? (float)n2
                 : closure_raw<T>(k, cc, cg, gg, u, ids, P, states, nb_start, nb);
      feat[ch][j] = closure_feature(k, feat[ch][j]);   /* once, for both networks */
    }
  auto predict = [&](const MlpDev* __restrict__ M, bool& oor) {
    bf16x8 x[2];
#pragma unroll
    for (int ch = 0; ch < 2; ch++)
#pragma unroll
      for (int j = 0; j < 8; j++)
        x[ch][j] = to_bf16(valid ? closure_input(M, 16 * ch + 8 * g + j, feat[ch][j], oor) : 0.0f);
    f32x16 z1[2], z2[2];
    bf16x8 h1b[2][2];
    return mlp_forward(M, [&](int seg) { return mlp_opnd(M, seg, lane); }, x, z1, z2, h1b, [](int, int, float) {});
  };
  bool oor = false, oor_f = false;
  const float y = predict(N, oor);
  const float yf = F ? predict(F, oor_f) : 0.0f;
  const bool oor_any = oor || __shfl_xor((int)oor, 32, 64);
  if (g != 0 || !valid) return;
  if (active && !active[cg.inst]) return;
  if (oor_any) atomicAdd(&flags[cg.inst], 1u);
  cpred[c] = exp((double)y * N->out_std + N->out_mean);
  if (F) fpred[c] = exp((double)yf * F->out_std + F->out_mean);
}

/* Output reductions of every realisation: one block per realisation, a
 * fixed assignment of voxels to threads and a fixed-order tree, so the
 * FP64 sums are reproducible.  out[i][s][d][3] = min, max, sum(u dV) with
 * the resting value added as k_stats_final adds it (min from FLT_MAX, max
 * from FLT_MIN, species.c:2808-2814). */
template <typename T>
__global__ void __launch_bounds__(BSTAT_THREADS)
k_batch_stats(const T* __restrict__ u, const uint8_t* __restrict__ ids, const DevParams<T>* __restrict__ P,
              const BatchGeo* __restrict__ geos, double dV, double* __restrict__ out)
{
  __shared__ double sh[3][MAXD][BSTAT_THREADS];
  const BatchGeo gg = geos[blockIdx.x];
  const T* ui = u + gg.u_off;
  const uint8_t* ii = ids + gg.id_off;
  for (uint32_t s = 0; s < gg.ns; s++) {
    double mn[MAXD], mx[MAXD], sm[MAXD];
    uint64_t cnt[MAXD];
    for (int d = 0; d < MAXD; d++) { mn[d] = INFINITY; mx[d] = -INFINITY; sm[d] = 0.0; cnt[d] = 0; }
    for (uint64_t v = threadIdx.x; v < gg.nv; v += blockDim.x) {
      const uint64_t z = v / gg.plane, c = v - z * gg.plane;
      const double x = (double)ui[(z * gg.ns + s) * gg.plane + c];
      const int d = ii[v] & 7;
      mn[d] = fmin(mn[d], x);
      mx[d] = fmax(mx[d], x);
      sm[d] += x;
      cnt[d]++;
    }
    for (int d = 0; d < MAXD; d++) {
      sh[0][d][threadIdx.x] = mn[d];
      sh[1][d][threadIdx.x] = mx[d];
      sh[2][d][threadIdx.x] = sm[d] * dV;
    }
    __syncthreads();
    for (int w = BSTAT_THREADS / 2; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w)
        for (int d = 0; d < MAXD; d++) {
          sh[0][d][threadIdx.x] = fmin(sh[0][d][threadIdx.x], sh[0][d][threadIdx.x + w]);
          sh[1][d][threadIdx.x] = fmax(sh[1][d][threadIdx.x], sh[1][d][threadIdx.x + w]);
          sh[2][d][threadIdx.x] += sh[2][d][threadIdx.x + w];
        }
      __syncthreads();
    }
    /* the block's results, then the voxels per domain through the same
     * shared array */
    double rmn = 0.0, rmx = 0.0, rsm = 0.0;
    if (threadIdx.x < MAXD) {
      rmn = sh[0][threadIdx.x][0];
      rmx = sh[1][threadIdx.x][0];
      rsm = sh[2][threadIdx.x][0];
    }
    __syncthreads();
    for (int d = 0; d < MAXD; d++) sh[0][d][threadIdx.x] = (double)cnt[d];
    __syncthreads();
    if (threadIdx.x < MAXD) {
      const int d = threadIdx.x;
      double n_d = 0.0;
      for (int k = 0; k < BSTAT_THREADS; k++) n_d += sh[0][d][k];
      const double r = P->init64[d][s];
      double* o = out + (((uint64_t)blockIdx.x * MAXS + s) * MAXD + d) * 3;
      o[0] = n_d > 0 ? fmin((double)FLT_MAX, rmn + r) : (double)FLT_MAX;
      o[1] = n_d > 0 ? fmax((double)FLT_MIN, rmx + r) : (double)FLT_MIN;
      o[2] = rsm + n_d * r * dV;
    }
    __syncthreads();
  }
}

/* Linescan projections of every realisation (cs_linescan's sums): part[i]
 * holds, per plane z, the sums over y of u w_y (axis 0: one entry per x)
 * or over x of u w_x (axes 1, 2: one entry per y); one block per (plane,
 * realisation).  w[i] is the realisation's [n0 | n1 | n2] weights. */
template <typename T>
__global__ void __launch_bounds__(256)
k_batch_scan_planes(const T* __restrict__ u, const uint8_t* __restrict__ ids, const DevParams<T>* __restrict__ P,
                    const BatchGeo* __restrict__ geos, const uint64_t* __restrict__ woff,
                    const double* __restrict__ w, const uint64_t* __restrict__ poff, double* __restrict__ part,
                    uint32_t s, uint32_t axis, uint32_t dmask)
{
  const uint32_t i = blockIdx.y;
  const BatchGeo gg = geos[i];
  const int z = blockIdx.x;
  if (z >= gg.nz) return;
  const double* wi = w + woff[i];
  const T* us = u + gg.u_off + ((uint64_t)z * gg.ns + s) * gg.plane;
  const uint8_t* iz = ids + gg.id_off + (uint64_t)z * gg.plane;
  double* pz = part + poff[i] + (uint64_t)z * (axis == 0 ? gg.nx : gg.ny);
  if (axis == 0) {
    const double* wy = wi + gg.nx;
    for (int x = threadIdx.x; x < gg.nx; x += blockDim.x) {
      double acc = 0.0;
      for (int y = 0; y < gg.ny; y++) {
        const uint64_t c = (uint64_t)y * gg.nx + x;
        const int d = iz[c] & 7;
        if ((dmask >> d) & 1u) acc += ((double)us[c] + P->init64[d][s]) * wy[y];
      }
      pz[x] = acc;
    }
  } else {
    const double* wx = wi;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
    for (int y = wave; y < gg.ny; y += nwave) {
      double acc = 0.0;
      for (int x = lane; x < gg.nx; x += 64) {
        const uint64_t c = (uint64_t)y * gg.nx + x;
        const int d = iz[c] & 7;
        if ((dmask >> d) & 1u) acc += ((double)us[c] + P->init64[d][s]) * wx[x];
      }
      for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
      if (lane == 0) pz[y] = acc;
    }
  }
}

/* Projection of realisation i on the axis from its plane sums, in a fixed
 * order: axis 0 sum_z w_z part[z][x]; axis 1 sum_z w_z part[z][y];
 * axis 2 sum_y w_y part[z][y]. */
__global__ void __launch_bounds__(256)
k_batch_scan_final(const BatchGeo* __restrict__ geos, const uint64_t* __restrict__ woff,
                   const double* __restrict__ w, const uint64_t* __restrict__ poff,
                   const double* __restrict__ part, const uint64_t* __restrict__ qoff, double* __restrict__ proj,
                   uint32_t axis)
{
  const uint32_t i = blockIdx.x;
  const BatchGeo gg = geos[i];
  const double* wi = w + woff[i];
  const double* pi = part + poff[i];
  double* qi = proj + qoff[i];
  const int nk = axis == 0 ? gg.nx : axis == 1 ? gg.ny : gg.nz;
  const int row = axis == 0 ? gg.nx : gg.ny;
  for (int k = threadIdx.x; k < nk; k += blockDim.x) {
    double acc = 0.0;
    if (axis == 2) {
      const double* wy = wi + gg.nx;
      for (int y = 0; y < gg.ny; y++) acc += wy[y] * pi[(uint64_t)k * row + y];
    } else {
      const double* wz = wi + gg.nx + gg.ny;
      for (int z = 0; z < gg.nz; z++) acc += wz[z] * pi[(uint64_t)z * row + k];
    }
    qi[k] = acc;
  }
}

template <typename T>
__global__ void k_batch_zero(T* __restrict__ u, uint64_t n)
{
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    u[i] = T(0);
}

/* Set the device copy of the first time index of a captured range. */
__global__ void k_batch_set_ti(uint64_t* __restrict__ ti, uint64_t v)
{
  if (threadIdx.x == 0 && blockIdx.x == 0) *ti = v;
}

}  // namespace

/* ------------------------------------------------------------ ensemble */

struct cs_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  uint32_t count = 0, ns = 0, nd = 0, nf = 0;
  size_t esize = 8;
  cs_problem model{};                      /* instance 0's shared part       */
  std::vector<BatchGeo> geo;
  std::vector<std::array<uint32_t, 3>> n;  /* reference axes per realisation */
  std::vector<std::vector<uint8_t>> h_ids; /* device-order ids (host copies) */
  std::vector<uint32_t> nchan, stoch, state_off;
  std::vector<std::vector<uint32_t>> flux_state_off;  /* [inst][flux]        */
  std::vector<uint8_t> init_states;
  uint64_t total_elems = 0, total_ids = 0;
  void* d_u[2] = {nullptr, nullptr};
  int cur = 0;
  uint8_t* d_ids = nullptr;
  void* d_params = nullptr;
  BatchGeo* d_geo = nullptr;
  BatchTile* d_tiles = nullptr;
  uint32_t ntiles = 0;
  uint32_t* d_queue = nullptr;
  BatchInst* d_inst = nullptr;
  BatchChan* d_chan = nullptr;
  BRow* d_rows = nullptr;
  BEnt* d_ents = nullptr;
  uint8_t* d_states = nullptr;
  uint8_t* d_recruited = nullptr;
  uint8_t* d_active = nullptr;
  uint64_t* d_last_open = nullptr;
  /* closure gating and replay */
  uint32_t gating = CS_GATE_OWN;
  const MlpDev* d_net = nullptr;           /* the closure's network            */
  const MlpDev* d_fnet = nullptr;          /* the flux closure's network       */
  double* d_fpred = nullptr;               /* its ratios, per channel          */
  BatchChanGeo* d_cgeo = nullptr;
  uint32_t* d_nb_start = nullptr;
  uint32_t* d_nb = nullptr;
  double* d_cpred = nullptr;
  double* d_chca = nullptr;
  float* d_feat = nullptr;
  uint32_t* d_flags = nullptr;
  uint32_t nchan_total = 0;
  double* d_uniforms = nullptr;
  uint64_t uniform_cap = 0;
  uint32_t uni_stride = 0;                 /* stochastic draws per gating step */
  double* d_stats = nullptr;
  /* linescan scratch */
  double* d_lw = nullptr;
  double* d_lpart = nullptr;
  double* d_lproj = nullptr;
  uint64_t* d_loff = nullptr;              /* woff | poff | qoff               */
  uint64_t lw_n = 0, lpart_n = 0, lproj_n = 0;
  uint32_t substeps[MAXS] = {0};
  uint32_t flux_substep[MAXF] = {0};
  uint32_t ndiscrete = 0;
  uint32_t rfast = 0;
  uint32_t fsm = 0;
  /* launch shape */
  cs_batch_launch launch{4, 32, 1, 1, 0};
  int tile_xcd = 0;                        /* k_batch_tiles' XCD groups       */
  int grid = 0;                            /* persistent blocks               */
  /* captured step periods, keyed by (phase within the period, buffer) */
  bool use_graph = false;
  uint64_t* d_ti = nullptr;
  std::map<std::pair<uint64_t, int>, hipGraphExec_t> graphs;
  std::vector<hipGraph_t> graph_defs;
};

namespace {

void free_batch(cs_batch* b)
{
  if (!b) return;
  (void)hipSetDevice(b->device);
  for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
  for (auto g : b->graph_defs) (void)hipGraphDestroy(g);
  void* ptrs[] = {b->d_u[0], b->d_u[1], b->d_ids, b->d_params, b->d_geo, b->d_tiles, b->d_queue, b->d_inst,
                  b->d_chan, b->d_rows, b->d_ents, b->d_states, b->d_recruited, b->d_active, b->d_last_open,
                  b->d_cgeo, b->d_nb_start, b->d_nb, b->d_cpred, b->d_fpred, b->d_chca, b->d_feat, b->d_flags,
                  b->d_uniforms, b->d_stats, b->d_lw, b->d_lpart, b->d_lproj, b->d_loff, b->d_ti};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

/* The parts of a problem every realisation of an ensemble must share. */
int same_model(const cs_problem& a, const cs_problem& p, uint32_t i)
{
  auto bad = [i](const char* what) { return set_err(CS_EINVAL, "realisation %u: %s differs from realisation 0", i, what); };
  if (p.num_domains != a.num_domains || p.num_species != a.num_species) return bad("domain or species count");
  if (p.precision != a.precision || p.rng != a.rng) return bad("precision or generator");
  if (p.h != a.h || p.dt != a.dt || p.dV != a.dV) return bad("h, dt or dV");
  if (memcmp(p.init, a.init, sizeof a.init) || memcmp(p.substeps, a.substeps, sizeof a.substeps) ||
      memcmp(p.coef, a.coef, sizeof a.coef))
    return bad("resting values, substeps or coefficient tables");
  if (p.num_buffers != a.num_buffers || p.reaction_substep != a.reaction_substep || p.dt_react != a.dt_react ||
      memcmp(p.bsp0, a.bsp0, sizeof a.bsp0) || memcmp(p.bsp1, a.bsp1, sizeof a.bsp1) ||
      memcmp(p.k_on, a.k_on, sizeof a.k_on) || memcmp(p.k_off, a.k_off, sizeof a.k_off) ||
      memcmp(p.tot, a.tot, sizeof a.tot))
    return bad("reaction tables");
  if (p.num_fluxes != a.num_fluxes || p.stochastic_substep != a.stochastic_substep ||
      p.force_dirichlet != a.force_dirichlet || memcmp(p.fixed, a.fixed, sizeof a.fixed))
    return bad("fluxes, stochastic substep, Dirichlet flag or fixed species");
  for (uint32_t f = 0; f < a.num_fluxes; f++) {
    const cs_flux_desc &x = a.fluxes[f], &y = p.fluxes[f];
    if (x.kind != y.kind || x.species != y.species || x.substep != y.substep || x.inner != y.inner ||
        x.discrete != y.discrete || memcmp(x.params, y.params, sizeof x.params))
      return bad("a flux model");
  }
  return CS_OK;
}

/* Column tiles of every realisation: `cols` columns by at most zchunk
 * planes, a realisation's planes split into chunks whose sizes differ by
 * at most one (25 planes by 8: 7, 6, 6, 6 rather than 8, 8, 8, 1, since
 * a chunk also reads the planes either side of it). */
void make_tiles(const cs_batch* b, uint32_t cols, uint32_t zchunk, std::vector<BatchTile>& out)
{
  out.clear();
  for (uint32_t i = 0; i < b->count; i++) {
    const BatchGeo& g = b->geo[i];
    const uint32_t nz = (uint32_t)g.nz;
    const uint32_t k = (nz + std::max<uint32_t>(1, zchunk) - 1) / std::max<uint32_t>(1, zchunk);
    for (uint32_t q = 0; q < k; q++) {
      const uint32_t z0 = q * nz / k, z1 = (q + 1) * nz / k;
      for (uint64_t c = 0; c < g.plane; c += cols) out.push_back(BatchTile{i, (uint32_t)c, z0, z1});
    }
  }
}

template <typename T>
int upload_tiles(cs_batch* b)
{
  std::vector<BatchTile> tiles;
  make_tiles(b, BX * b->launch.rows, b->launch.zchunk, tiles);
  if (b->d_tiles) HIP_TRY(hipFree(b->d_tiles));
  b->d_tiles = nullptr;
  HIP_TRY(hipMalloc(&b->d_tiles, sizeof(BatchTile) * std::max<size_t>(tiles.size(), 1)));
  HIP_TRY(hipMemcpyAsync(b->d_tiles, tiles.data(), sizeof(BatchTile) * tiles.size(), hipMemcpyHostToDevice,
                         b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  b->ntiles = (uint32_t)tiles.size();
  /* persistent grid: as many blocks as stay resident, in whole rounds of
   * the eight XCDs, never more than there are tiles */
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, b->device));
  const int per_cu = std::max(1, 2048 / (BX * (int)b->launch.rows));
  int grid = prop.multiProcessorCount * std::min(per_cu, 8);
  grid = std::min<int>(grid, (int)((b->ntiles + NQUEUE - 1) / NQUEUE * NQUEUE));
  b->grid = std::max(grid, NQUEUE);
  /* one block per tile: groups of up to 16 consecutive tiles per XCD,
   * small enough that every XCD gets at least three groups */
  b->tile_xcd = 0;
  for (int G = 16; G >= 2; G /= 2)
    if ((int64_t)b->ntiles >= 8LL * 3 * G) { b->tile_xcd = G; break; }
  return CS_OK;
}

template <typename T, int NS, int PF, bool NT, uint32_t FM = BS_DYN>
void launch_bstencil_k(cs_batch* b, const void* uin, void* uout, uint32_t dmask, uint32_t react)
{
  if (b->launch.persistent)
    hipLaunchKernelGGL((k_batch_stencil<T, NS, PF, NT, FM>), dim3(b->grid), dim3(BX, b->launch.rows), 0, b->stream,
                       (const T*)uin, (T*)uout, b->d_ids, (const DevParams<T>*)b->d_params, b->d_geo, b->d_tiles,
                       b->ntiles, dmask, react, b->d_queue);
  else
    hipLaunchKernelGGL((k_batch_tiles<T, NS, PF, NT, FM>), dim3(xcd_launch((int)b->ntiles, b->tile_xcd)),
                       dim3(BX, b->launch.rows), 0, b->stream, (const T*)uin, (T*)uout, b->d_ids,
                       (const DevParams<T>*)b->d_params, b->d_geo, b->d_tiles, b->ntiles, dmask, react,
                       b->tile_xcd);
}

/* The step masks of the reference's five-slot species layout (Ca, dye,
 * CMDN-CSQN, ATP, TRPN: Ca and ATP every step, the dye and CMDN on their
 * own multiples) run kernels compiled for them in FP32 with one-plane
 * prefetch; other masks, layouts and FP64 read the mask at run time. */
template <typename T, int NS, int PF, bool NT>
void launch_bstencil_v(cs_batch* b, const void* uin, void* uout, uint32_t dmask, uint32_t react)
{
  if constexpr (sizeof(T) == 4 && NS == 5 && PF == 1) {
    if (react == 2) {
      switch (dmask) {
        case 0x9: return launch_bstencil_k<T, NS, PF, NT, 0x9>(b, uin, uout, dmask, react);
        case 0xB: return launch_bstencil_k<T, NS, PF, NT, 0xB>(b, uin, uout, dmask, react);
        case 0xD: return launch_bstencil_k<T, NS, PF, NT, 0xD>(b, uin, uout, dmask, react);
        case 0xF: return launch_bstencil_k<T, NS, PF, NT, 0xF>(b, uin, uout, dmask, react);
        default: break;
      }
    }
  }
  launch_bstencil_k<T, NS, PF, NT>(b, uin, uout, dmask, react);
}

template <typename T, int NS>
void launch_bstencil_ns(cs_batch* b, const void* uin, void* uout, uint32_t dmask, uint32_t react)
{
  const bool nt = b->launch.nontemporal != 0;
  if (b->launch.prefetch == 2) {
    if (nt) launch_bstencil_v<T, NS, 2, true>(b, uin, uout, dmask, react);
    else launch_bstencil_v<T, NS, 2, false>(b, uin, uout, dmask, react);
  } else {
    if (nt) launch_bstencil_v<T, NS, 1, true>(b, uin, uout, dmask, react);
    else launch_bstencil_v<T, NS, 1, false>(b, uin, uout, dmask, react);
  }
}

template <typename T>
void launch_bstencil(cs_batch* b, const void* uin, void* uout, uint32_t dmask, uint32_t react)
{
  switch (b->ns) {
    case 1: launch_bstencil_ns<T, 1>(b, uin, uout, dmask, react); break;
    case 2: launch_bstencil_ns<T, 2>(b, uin, uout, dmask, react); break;
    case 3: launch_bstencil_ns<T, 3>(b, uin, uout, dmask, react); break;
    case 4: launch_bstencil_ns<T, 4>(b, uin, uout, dmask, react); break;
    case 5: launch_bstencil_ns<T, 5>(b, uin, uout, dmask, react); break;
    case 6: launch_bstencil_ns<T, 6>(b, uin, uout, dmask, react); break;
    case 7: launch_bstencil_ns<T, 7>(b, uin, uout, dmask, react); break;
    default: launch_bstencil_ns<T, 8>(b, uin, uout, dmask, react); break;
  }
}

uint32_t b_diffusing(const cs_batch* b, uint64_t ti)
{
  uint32_t m = 0;
  for (uint32_t s = 0; s < b->ns; s++)
    if (b->substeps[s] && ti % b->substeps[s] == 0) m |= 1u << s;
  return m;
}

uint32_t b_react(const cs_batch* b, uint64_t ti, uint32_t dmask)
{
  if (b->model.num_buffers == 0 || ti % b->model.reaction_substep != 0) return 0;
  return (b->rfast && (dmask & 1u)) ? 2u : 1u;
}

uint32_t b_fluxmask(const cs_batch* b, uint64_t ti)
{
  uint32_t m = 0;
  for (uint32_t f = 0; f < b->nf; f++)
    if (ti % b->flux_substep[f] == 0) m |= 1u << f;
  return m;
}

bool b_gates(const cs_batch* b, uint64_t ti) { return b->ndiscrete && ti % b->model.stochastic_substep == 0; }

template <typename T>
void launch_bbridge(cs_batch* b, const void* u, void* un, const BridgeArgs& a, const double* d_uni,
                    const double* cpred = nullptr, const double* fpred = nullptr)
{
  hipLaunchKernelGGL(k_batch_bridge<T>, dim3(b->count), dim3(BRIDGE_BATCH_THREADS), 0, b->stream, (const T*)u,
                     (T*)un, (const DevParams<T>*)b->d_params, b->d_inst, b->d_active, b->d_chan, b->d_rows,
                     b->d_ents, b->d_states, b->d_recruited, b->d_last_open, a, d_uni, cpred, fpred);
}

/* The closure networks' predictions from the state u: the gating
 * network's mouth Ca (with the range monitor) and, when set, the flux
 * network's release ratios. */
template <typename T>
void launch_predict(cs_batch* b, const void* u)
{
  const unsigned blocks = (b->nchan_total + 127) / 128;
  if (!blocks) return;
  hipLaunchKernelGGL(k_closure_predict<T>, dim3(blocks), dim3(256), 0, b->stream, (const T*)u, b->d_ids,
                     (const DevParams<T>*)b->d_params, b->d_net, b->d_fnet, b->d_geo, b->d_active, b->d_cgeo,
                     b->d_states, b->d_nb_start, b->d_nb, b->nchan_total, b->d_cpred, b->d_fpred, b->d_flags);
}

/* Steps [first, first + n) on the stream: gating of the first step, then
 * per step the stencil pass and the bridge (its fluxes, and the gating of
 * the next step unless it is the last).  ev, when given, holds 2n events
 * that bracket every stencil launch.  With d_ti (a captured range), the
 * bridges read the first time index from the device and carry offsets. */
template <typename T>
int b_run(cs_batch* b, uint64_t first, uint64_t n, const double* d_uni, hipEvent_t* ev,
          const uint64_t* d_ti = nullptr)
{
  const uint32_t D = b->model.stochastic_substep;
  uint64_t gate_idx = 0;
  auto bridge = [&](uint32_t fluxmask, uint64_t gti, bool gate_next, const void* u, void* un) {
    BridgeArgs a{};
    a.fluxmask = fluxmask;
    a.gate = (gate_next && b->gating != CS_GATE_REPLAY && b_gates(b, gti)) ? 1u : 0u;
    a.gate_ti = d_ti ? gti - first : gti;
    a.ti_base = d_ti;
    a.uni_base = d_uni ? gate_idx * b->uni_stride : 0;
    a.D = D;
    if (a.gate) gate_idx++;
    const bool fclosure = b->gating == CS_GATE_CLOSURE && b->d_fnet;
    if (b->gating == CS_GATE_CLOSURE && (a.gate || fclosure)) {
      /* fluxes first (scaled by the flux network's ratios, predicted from
       * the state they read), then the predictions from the new state,
       * then gating on them; with the flux network every bridge predicts,
       * since the next step's fluxes read its ratios */
      if (a.fluxmask) {
        BridgeArgs f = a;
        f.gate = 0;
        launch_bbridge<T>(b, u, un, f, d_uni, nullptr, fclosure ? b->d_fpred : nullptr);
      }
      launch_predict<T>(b, un);
      if (a.gate) {
        a.fluxmask = 0;
        launch_bbridge<T>(b, un, un, a, d_uni, b->d_cpred);
      }
      return;
    }
    if (a.fluxmask || a.gate) launch_bbridge<T>(b, u, un, a, d_uni);
  };
  if (n == 0) return CS_OK;
  bridge(0, first, true, b->d_u[b->cur], b->d_u[b->cur]);
  for (uint64_t i = 0; i < n; i++) {
    const uint64_t ti = first + i;
    const uint32_t dmask = b_diffusing(b, ti);
    void* uin = b->d_u[b->cur];
    void* uout = b->d_u[b->cur ^ 1];
    if (ev) HIP_TRY(hipEventRecord(ev[2 * i], b->stream));
    launch_bstencil<T>(b, uin, uout, dmask, b_react(b, ti, dmask));
    if (ev) HIP_TRY(hipEventRecord(ev[2 * i + 1], b->stream));
    bridge(b->nf ? b_fluxmask(b, ti) : 0u, ti + 1, i + 1 < n, uin, uout);
    b->cur ^= 1;
  }
  HIP_TRY(hipGetLastError());
  return CS_OK;
}

uint64_t b_gates_in(const cs_batch* b, uint64_t first, uint64_t n)
{
  if (!b->ndiscrete) return 0;
  const uint64_t D = b->model.stochastic_substep;
  auto upto = [D](uint64_t k) { return k == 0 ? 0 : (k - 1) / D + 1; };
  return upto(first + n) - upto(first);
}

/* Period of the step pattern: the least common multiple of every substep
 * (diffusion, reaction, gating, fluxes). */
uint64_t b_period(const cs_batch* b)
{
  auto lcm = [](uint64_t a, uint64_t c) {
    uint64_t x = a, y = c;
    while (y) { const uint64_t t = x % y; x = y; y = t; }
    return a / x * c;
  };
  uint64_t p = 1;
  for (uint32_t s = 0; s < b->ns; s++)
    if (b->substeps[s]) p = lcm(p, b->substeps[s]);
  p = lcm(p, std::max<uint32_t>(1, b->model.reaction_substep));
  if (b->ndiscrete) p = lcm(p, std::max<uint32_t>(1, b->model.stochastic_substep));
  for (uint32_t f = 0; f < b->nf; f++) p = lcm(p, std::max<uint32_t>(1, b->flux_substep[f]));
  return p;
}

template <typename T>
int b_step(cs_batch* b, uint64_t first, uint64_t n, const double* host_uni, uint64_t count)
{
  HIP_TRY(hipSetDevice(b->device));
  const double* d_uni = nullptr;
  if (b->model.rng == CS_RNG_MT && b->uni_stride) {
    const uint64_t need = b_gates_in(b, first, n) * b->uni_stride;
    if (!host_uni) return set_err(CS_EINVAL, "this ensemble draws gating uniforms from the caller: use cs_batch_step_uniforms");
    if (count != need)
      return set_err(CS_EINVAL, "expected %llu uniforms, got %llu", (unsigned long long)need, (unsigned long long)count);
    if (need > b->uniform_cap) {
      if (b->d_uniforms) HIP_TRY(hipFree(b->d_uniforms));
      b->d_uniforms = nullptr;
      HIP_TRY(hipMalloc(&b->d_uniforms, sizeof(double) * need));
      b->uniform_cap = need;
    }
    if (need) {
      /* the caller's array may be freed once this call returns, and a copy
       * from pageable memory can still be reading it then: wait for it */
      HIP_TRY(hipMemcpyAsync(b->d_uniforms, host_uni, sizeof(double) * need, hipMemcpyHostToDevice, b->stream));
      HIP_TRY(hipStreamSynchronize(b->stream));
    }
    d_uni = b->d_uniforms;
    return b_run<T>(b, first, n, d_uni, nullptr);
  }
  if (host_uni && count) return set_err(CS_EINVAL, "this ensemble uses Philox gating; no uniforms expected");
  if (!b->use_graph) return b_run<T>(b, first, n, nullptr, nullptr);
  /* whole periods of the step pattern replay one captured graph each; the
   * rest runs directly.  The captured bridges read the period's first
   * time index from d_ti, set before every replay.  A period holds an even
   * number of steps, so it leaves the ping-pong buffers as it found them. */
  uint64_t P = b_period(b);
  if (P % 2) P *= 2;
  uint64_t done = 0;
  while (n - done >= P) {
    const uint64_t ti = first + done;
    const std::pair<uint64_t, int> key{ti % P, b->cur};
    auto it = b->graphs.find(key);
    if (it == b->graphs.end()) {
      hipGraph_t g = nullptr;
      HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      const int cur0 = b->cur;
      const int rc = b_run<T>(b, ti, P, nullptr, nullptr, b->d_ti);
      b->cur = cur0;
      const hipError_t e = hipStreamEndCapture(b->stream, &g);
      if (rc != CS_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (e != hipSuccess) return set_err(CS_EDEVICE, "graph capture failed: %s", hipGetErrorString(e));
      hipGraphExec_t ex = nullptr;
      const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
      if (ei != hipSuccess) {
        (void)hipGraphDestroy(g);
        return set_err(CS_EDEVICE, "graph instantiation failed: %s", hipGetErrorString(ei));
      }
      b->graph_defs.push_back(g);
      it = b->graphs.emplace(key, ex).first;
    }
    hipLaunchKernelGGL(k_batch_set_ti, dim3(1), dim3(64), 0, b->stream, b->d_ti, ti);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipGraphLaunch(it->second, b->stream));
    done += P;
  }
  if (done < n) return b_run<T>(b, first + done, n - done, nullptr, nullptr);
  return CS_OK;
}

void drop_graphs(cs_batch* b)
{
  for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
  for (auto g : b->graph_defs) (void)hipGraphDestroy(g);
  b->graphs.clear();
  b->graph_defs.clear();
}

}  // namespace
/* ------------------------------------------------------------- C ABI */

extern "C" {

int cs_batch_create(const cs_problem* problems, uint32_t count, cs_batch** out)
{
  if (!out) return set_err(CS_EINVAL, "out is NULL");
  *out = nullptr;
  if (!problems || count == 0) return set_err(CS_EINVAL, "an ensemble needs at least one problem");
  for (uint32_t i = 0; i < count; i++) {
    const int rc = same_model(problems[0], problems[i], i);
    if (rc != CS_OK) return rc;
  }
  const cs_problem& p0 = problems[0];
  if (p0.num_species < 1 || p0.num_species > MAXS) return set_err(CS_EINVAL, "num_species outside 1..%d", MAXS);
  if (p0.num_domains < 1 || p0.num_domains > MAXD) return set_err(CS_EINVAL, "num_domains outside 1..%d", MAXD);
  if (p0.num_fluxes > MAXF) return set_err(CS_EINVAL, "too many fluxes");
  if (p0.precision != CS_PREC_F64 && p0.precision != CS_PREC_F32) return set_err(CS_EINVAL, "unknown precision");
  if (p0.reaction_substep < 1 || p0.stochastic_substep < 1) return set_err(CS_EINVAL, "substeps must be >= 1");
  if (!(p0.dt > 0) || !(p0.h > 0)) return set_err(CS_EINVAL, "dt and h must be positive");
  for (uint32_t s = 0; s < p0.num_species; s++)
    if (p0.substeps[s] > 255) return set_err(CS_EINVAL, "substep too large");
  for (uint32_t bb = 0; bb < p0.num_buffers; bb++)
    for (uint32_t d = 0; d < p0.num_domains; d++)
      if (p0.bsp0[bb][d] >= p0.num_species || p0.bsp1[bb][d] >= p0.num_species)
        return set_err(CS_ERANGE, "buffer %u of domain %u names a species out of range", bb, d);
  bool seen_density = false;
  for (uint32_t f = 0; f < p0.num_fluxes; f++) {
    const cs_flux_desc& F = p0.fluxes[f];
    if (F.kind != CS_FLUX_RYR && F.kind != CS_FLUX_SERCA) return set_err(CS_EINVAL, "flux %u: unknown kind", f);
    if (F.species >= p0.num_species) return set_err(CS_EINVAL, "flux %u: bad species", f);
    if (F.substep < 1) return set_err(CS_EINVAL, "flux %u: substep must be >= 1", f);
    if (F.discrete) {
      if (seen_density) return set_err(CS_EINVAL, "discrete fluxes must come first");
      if (F.kind != CS_FLUX_RYR) return set_err(CS_EINVAL, "flux %u: discrete flux must be RyR", f);
    } else {
      seen_density = true;
    }
  }

  cs_batch* b = new cs_batch();
  auto bail = [&](int code) { free_batch(b); return code; };
  b->model = p0;
  b->model.domain_ids = nullptr;
  for (auto& F : b->model.fluxes) { F.pairs = nullptr; F.init_states = nullptr; F.num_pairs = 0; }
  b->count = count;
  b->ns = p0.num_species;
  b->nd = p0.num_domains;
  b->nf = p0.num_fluxes;
  b->esize = p0.precision == CS_PREC_F64 ? 8 : 4;
  for (uint32_t s = 0; s < b->ns; s++) b->substeps[s] = p0.substeps[s];
  for (uint32_t f = 0; f < b->nf; f++) {
    b->flux_substep[f] = p0.fluxes[f].substep;
    if (p0.fluxes[f].discrete) b->ndiscrete++;
    b->fsm |= 1u << p0.fluxes[f].species;
  }
  if (hipGetDevice(&b->device) != hipSuccess) return bail(set_err(CS_EDEVICE, "no HIP device"));
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(set_err(CS_EDEVICE, "hipStreamCreate failed"));

  /* geometry, ids and the flat flux records of every realisation */
  std::vector<uint8_t> ids_all;
  std::vector<BatchInst> insts(count);
  std::vector<BatchChan> chans;
  std::vector<BatchChanGeo> cgeo;
  std::vector<uint32_t> nbs{0}, nbl;
  std::vector<BRow> rows;
  std::vector<BEnt> ents;
  b->geo.resize(count);
  b->n.resize(count);
  b->h_ids.resize(count);
  b->flux_state_off.assign(count, std::vector<uint32_t>(MAXF, 0));
  b->nchan.assign(count, 0);
  b->stoch.assign(count, 0);
  b->state_off.assign(count, 0);
  uint64_t uoff = 0, idoff = 0;
  const uint64_t align = 256;
  for (uint32_t i = 0; i < count; i++) {
    const cs_problem& p = problems[i];
    if (p.n[0] < 1 || p.n[1] < 1 || p.n[2] < 1) return bail(set_err(CS_EINVAL, "realisation %u: empty grid", i));
    if (!p.domain_ids) return bail(set_err(CS_EINVAL, "realisation %u: domain_ids is NULL", i));
    const uint32_t n0 = p.n[0], n1 = p.n[1], n2 = p.n[2];
    const uint64_t nv = (uint64_t)n0 * n1 * n2;
    /* the column march addresses a realisation's z-groups with 32-bit
     * offsets: its whole field stack must stay below 2^32 bytes */
    if (nv * b->ns * b->esize >= (1ull << 32))
      return bail(set_err(CS_ERANGE, "realisation %u: %llu voxels exceed the 4 GiB field limit", i,
                          (unsigned long long)nv));
    for (uint64_t v = 0; v < nv; v++)
      if (p.domain_ids[v] >= p.num_domains)
        return bail(set_err(CS_ERANGE, "realisation %u: domain id out of range at voxel %llu", i,
                            (unsigned long long)v));
    BatchGeo& g = b->geo[i];
    g.nx = (int)n0;
    g.ny = (int)n1;
    g.nz = (int)n2;
    g.ns = b->ns;
    g.plane = (uint64_t)n0 * n1;
    g.nv = nv;
    g.u_off = uoff;
    g.id_off = idoff;
    b->n[i] = {n0, n1, n2};
    const uint64_t ebytes = nv * b->ns * b->esize;
    uoff += ((ebytes + align - 1) / align * align) / b->esize;
    idoff += (nv + align - 1) / align * align;
    /* ids in device order with the uniform-neighbourhood bit */
    std::vector<uint8_t>& hid = b->h_ids[i];
    hid.assign(nv, 0);
    for (uint32_t x = 0; x < n0; x++)
      for (uint32_t y = 0; y < n1; y++)
        for (uint32_t z = 0; z < n2; z++) {
          const uint64_t r = ((uint64_t)x * n1 + y) * n2 + z;
          const uint8_t d = p.domain_ids[r];
          bool uni = true;
          if (x > 0) uni &= p.domain_ids[r - (uint64_t)n1 * n2] == d;
          if (x + 1 < n0) uni &= p.domain_ids[r + (uint64_t)n1 * n2] == d;
          if (y > 0) uni &= p.domain_ids[r - n2] == d;
          if (y + 1 < n1) uni &= p.domain_ids[r + n2] == d;
          if (z > 0) uni &= p.domain_ids[r - 1] == d;
          if (z + 1 < n2) uni &= p.domain_ids[r + 1] == d;
          hid[((uint64_t)z * n1 + y) * n0 + x] = (uint8_t)(d | (uni ? 0x80 : 0));
        }
    ids_all.resize(g.id_off + nv);
    memcpy(ids_all.data() + g.id_off, hid.data(), nv);

    /* flux rows: every pair touching a (species, voxel), in k_flux's order */
    struct Ent { uint64_t key; uint32_t pair; uint8_t flux; int8_t sign; };
    std::vector<Ent> es;
    std::vector<uint32_t> pairs_dev;
    std::vector<uint32_t> pair_off(b->nf + 1, 0);
    const uint32_t nref[3] = {n0, n1, n2};
    for (uint32_t f = 0; f < b->nf; f++) {
      const cs_flux_desc& F = p.fluxes[f];
      if (F.num_pairs && !F.pairs) return bail(set_err(CS_EINVAL, "realisation %u flux %u: pairs is NULL", i, f));
      if (F.discrete && F.num_pairs && !F.init_states)
        return bail(set_err(CS_EINVAL, "realisation %u flux %u: init_states is NULL", i, f));
      pair_off[f + 1] = pair_off[f] + F.num_pairs;
      for (uint32_t k = 0; k < F.num_pairs; k++) {
        if (F.pairs[2 * k] >= nv || F.pairs[2 * k + 1] >= nv)
          return bail(set_err(CS_ERANGE, "realisation %u flux %u: pair voxel out of range", i, f));
        const uint32_t from = ref_to_dev(F.pairs[2 * k], nref), to = ref_to_dev(F.pairs[2 * k + 1], nref);
        pairs_dev.push_back(from);
        pairs_dev.push_back(to);
        const uint64_t sp = F.species;
        es.push_back({(sp << 32) | to, k, (uint8_t)f, (int8_t)1});
        if (F.inner) es.push_back({(sp << 32) | from, k, (uint8_t)f, (int8_t)-1});
      }
    }
    std::stable_sort(es.begin(), es.end(), [](const Ent& a, const Ent& c) { return a.key < c.key; });
    /* channel states and draws of this realisation */
    BatchInst& in = insts[i];
    in.seed = (uint32_t)p.seed;
    in.instance = (uint32_t)p.instance;
    in.chan_off = (uint32_t)chans.size();
    in.uni_off = b->uni_stride;
    b->state_off[i] = (uint32_t)b->init_states.size();
    uint32_t ndraw = 0;
    auto elem_of = [&](uint32_t s, uint32_t v) {
      const uint64_t z = v / g.plane;
      return g.u_off + (z * g.ns + s) * g.plane + (v - z * g.plane);
    };
    auto dom = [&](uint32_t v) { return (int)(hid[v] & 7); };
    for (uint32_t f = 0; f < b->nf; f++) {
      const cs_flux_desc& F = p.fluxes[f];
      if (!F.discrete) continue;
      b->flux_state_off[i][f] = (uint32_t)b->init_states.size();
      const bool stochastic = !(F.params[8] > 0);
      for (uint32_t k = 0; k < F.num_pairs; k++) {
        const uint32_t from = pairs_dev[2 * (pair_off[f] + k)], to = pairs_dev[2 * (pair_off[f] + k) + 1];
        const uint32_t st = (uint32_t)b->init_states.size();
        chans.push_back(BatchChan{elem_of(0, to), p.init[dom(to)][0], f, k, st, stochastic ? ndraw++ : NO_ROW});
        cgeo.push_back(BatchChanGeo{i, (int)(to % n0), (int)((to / n0) % n1), (int)(to / g.plane), elem_of(0, from),
                                    p.init[dom(from)][0]});
        b->init_states.push_back(F.init_states[k] ? 1 : 0);
      }
    }
    in.nchan = (uint32_t)chans.size() - in.chan_off;
    /* neighbours of each channel within Chebyshev distance 2 */
    for (uint32_t a = in.chan_off; a < in.chan_off + in.nchan; a++) {
      for (uint32_t c2 = in.chan_off; c2 < in.chan_off + in.nchan; c2++) {
        const int d = std::max({std::abs(cgeo[a].x - cgeo[c2].x), std::abs(cgeo[a].y - cgeo[c2].y),
                                std::abs(cgeo[a].z - cgeo[c2].z)});
        if (d <= 2) nbl.push_back(chans[c2].state | (d == 2 ? NB_FAR : 0u));
      }
      nbs.push_back((uint32_t)nbl.size());
    }
    b->nchan[i] = in.nchan;
    b->stoch[i] = ndraw;
    b->uni_stride += ndraw;
    in.row_off = (uint32_t)rows.size();
    for (size_t k = 0; k < es.size();) {
      size_t e2 = k;
      while (e2 < es.size() && es[e2].key == es[k].key) e2++;
      const uint32_t v = (uint32_t)(es[k].key & 0xFFFFFFFFu), sp = (uint32_t)(es[k].key >> 32);
      const uint32_t x = v % n0, y = (v / n0) % n1, z = (uint32_t)(v / g.plane);
      const bool face = x == 0 || y == 0 || z == 0 || x + 1 == n0 || y + 1 == n1 || z + 1 == n2;
      const bool skip = p.fixed[dom(v)][sp] || (sp == 0 && p.force_dirichlet && face);
      const uint32_t start = (uint32_t)ents.size();
      for (size_t e = k; e < e2; e++) {
        const uint32_t f = es[e].flux, gp = pair_off[f] + es[e].pair;
        const uint32_t from = pairs_dev[2 * gp], to = pairs_dev[2 * gp + 1];
        ents.push_back(BEnt{elem_of(sp, from), elem_of(sp, to), p.init[dom(from)][sp], p.init[dom(to)][sp], NO_ROW,
                            NO_ROW, p.fluxes[f].discrete ? b->flux_state_off[i][f] + es[e].pair : NO_ROW,
                            f | (es[e].sign > 0 ? 1u : 0u) << 8});
      }
      rows.push_back(BRow{elem_of(sp, v), start, (uint32_t)ents.size(), skip ? 1u : 0u, 0});
      k = e2;
    }
    in.nrows = (uint32_t)rows.size() - in.row_off;
  }
  b->total_elems = uoff;
  b->total_ids = idoff;
  b->nchan_total = (uint32_t)chans.size();

  auto upload = [&](void** dst, const void* src, size_t bytes) -> int {
    HIP_TRY(hipMalloc(dst, std::max<size_t>(bytes, 16)));
    if (bytes) HIP_TRY(hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, b->stream));
    return CS_OK;
  };
  int rc;
  if (hipMalloc(&b->d_ids, std::max<uint64_t>(idoff, 16)) != hipSuccess ||
      hipMalloc(&b->d_u[0], std::max<uint64_t>(uoff * b->esize, 16)) != hipSuccess ||
      hipMalloc(&b->d_u[1], std::max<uint64_t>(uoff * b->esize, 16)) != hipSuccess)
    return bail(set_err(CS_ENOMEM, "device allocation of %llu field bytes failed",
                        (unsigned long long)(2 * uoff * b->esize)));
  if (hipMemcpyAsync(b->d_ids, ids_all.data(), ids_all.size(), hipMemcpyHostToDevice, b->stream) != hipSuccess)
    return bail(set_err(CS_EDEVICE, "upload of ids failed"));
  if ((rc = upload((void**)&b->d_geo, b->geo.data(), sizeof(BatchGeo) * count)) != CS_OK ||
      (rc = upload((void**)&b->d_inst, insts.data(), sizeof(BatchInst) * count)) != CS_OK ||
      (rc = upload((void**)&b->d_chan, chans.data(), sizeof(BatchChan) * chans.size())) != CS_OK ||
      (rc = upload((void**)&b->d_cgeo, cgeo.data(), sizeof(BatchChanGeo) * cgeo.size())) != CS_OK ||
      (rc = upload((void**)&b->d_nb_start, nbs.data(), sizeof(uint32_t) * nbs.size())) != CS_OK ||
      (rc = upload((void**)&b->d_nb, nbl.data(), sizeof(uint32_t) * nbl.size())) != CS_OK ||
      (rc = upload((void**)&b->d_rows, rows.data(), sizeof(BRow) * rows.size())) != CS_OK ||
      (rc = upload((void**)&b->d_ents, ents.data(), sizeof(BEnt) * ents.size())) != CS_OK ||
      (rc = upload((void**)&b->d_states, b->init_states.data(), b->init_states.size())) != CS_OK ||
      (rc = upload((void**)&b->d_recruited, b->init_states.data(), b->init_states.size())) != CS_OK)
    return bail(rc);
  std::vector<uint8_t> act(count, 1);
  if ((rc = upload((void**)&b->d_active, act.data(), count)) != CS_OK) return bail(rc);
  if (hipMalloc(&b->d_last_open, sizeof(uint64_t) * count) != hipSuccess ||
      hipMalloc(&b->d_flags, sizeof(uint32_t) * count) != hipSuccess ||
      hipMalloc(&b->d_cpred, sizeof(double) * std::max<size_t>(chans.size(), 1)) != hipSuccess ||
      hipMalloc(&b->d_fpred, sizeof(double) * std::max<size_t>(chans.size(), 1)) != hipSuccess ||
      hipMalloc(&b->d_queue, sizeof(uint32_t) * (NQUEUE + 1)) != hipSuccess ||
      hipMalloc(&b->d_ti, sizeof(uint64_t)) != hipSuccess ||
      hipMalloc(&b->d_stats, sizeof(double) * count * MAXS * MAXD * 3) != hipSuccess)
    return bail(set_err(CS_ENOMEM, "device allocation of ensemble records failed"));
  if (hipMemsetAsync(b->d_queue, 0, sizeof(uint32_t) * (NQUEUE + 1), b->stream) != hipSuccess)
    return bail(set_err(CS_EDEVICE, "queue initialisation failed"));
  /* device tables: the shared model with instance 0's flux layout */
  if (b->esize == 8) {
    DevParams<double> dp;
    fill_params<double>(dp, &problems[0], b->ndiscrete);
    b->rfast = dp.rfast;
    if ((rc = upload(&b->d_params, &dp, sizeof dp)) != CS_OK) return bail(rc);
    if (hipStreamSynchronize(b->stream) != hipSuccess) return bail(set_err(CS_EDEVICE, "upload failed"));
  } else {
    DevParams<float> dp;
    fill_params<float>(dp, &problems[0], b->ndiscrete);
    b->rfast = dp.rfast;
    if ((rc = upload(&b->d_params, &dp, sizeof dp)) != CS_OK) return bail(rc);
    if (hipStreamSynchronize(b->stream) != hipSuccess) return bail(set_err(CS_EDEVICE, "upload failed"));
  }
  rc = b->esize == 8 ? upload_tiles<double>(b) : upload_tiles<float>(b);
  if (rc != CS_OK) return bail(rc);
  rc = cs_batch_reset(b);
  if (rc != CS_OK) return bail(rc);
  *out = b;
  return CS_OK;
}

int cs_batch_destroy(cs_batch* b)
{
  if (!b) return CS_OK;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  free_batch(b);
  return CS_OK;
}

int cs_batch_reset(cs_batch* b)
{
  if (!b) return set_err(CS_EINVAL, "ensemble is NULL");
  HIP_TRY(hipSetDevice(b->device));
  const uint64_t n = b->total_elems;
  const int blocks = (int)std::min<uint64_t>(8192, (n + 255) / 256 + 1);
  for (int k = 0; k < 2; k++) {
    if (b->esize == 8)
      hipLaunchKernelGGL(k_batch_zero<double>, dim3(blocks), dim3(256), 0, b->stream, (double*)b->d_u[k], n);
    else
      hipLaunchKernelGGL(k_batch_zero<float>, dim3(blocks), dim3(256), 0, b->stream, (float*)b->d_u[k], n);
  }
  HIP_TRY(hipGetLastError());
  b->cur = 0;
  if (!b->init_states.empty()) {
    HIP_TRY(hipMemcpyAsync(b->d_states, b->init_states.data(), b->init_states.size(), hipMemcpyHostToDevice,
                           b->stream));
    HIP_TRY(hipMemcpyAsync(b->d_recruited, b->init_states.data(), b->init_states.size(), hipMemcpyHostToDevice,
                           b->stream));
  }
  std::vector<uint64_t> lo(b->count, 0);
  std::vector<uint32_t> fl(b->count, 0);
  HIP_TRY(hipMemcpyAsync(b->d_last_open, lo.data(), sizeof(uint64_t) * b->count, hipMemcpyHostToDevice, b->stream));
  HIP_TRY(hipMemcpyAsync(b->d_flags, fl.data(), sizeof(uint32_t) * b->count, hipMemcpyHostToDevice, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  return CS_OK;
}

int cs_batch_info(const cs_batch* b, uint32_t* count, uint64_t* voxels, uint64_t* field_bytes)
{
  if (!b) return set_err(CS_EINVAL, "ensemble is NULL");
  uint64_t nv = 0;
  for (const auto& g : b->geo) nv += g.nv;
  if (count) *count = b->count;
  if (voxels) *voxels = nv;
  if (field_bytes) *field_bytes = 2 * b->total_elems * b->esize + b->total_ids;
  return CS_OK;
}

int cs_batch_upload_species(cs_batch* b, uint32_t inst, uint32_t s, const double* src)
{
  if (!b || !src) return set_err(CS_EINVAL, "NULL argument");
  if (inst >= b->count) return set_err(CS_ERANGE, "realisation %u out of range", inst);
  if (s >= b->ns) return set_err(CS_ERANGE, "species %u out of range", s);
  HIP_TRY(hipSetDevice(b->device));
  const BatchGeo& g = b->geo[inst];
  const uint32_t n0 = b->n[inst][0], n1 = b->n[inst][1], n2 = b->n[inst][2];
  std::vector<char> buf(g.nv * b->esize);
  for (uint32_t x = 0; x < n0; x++)
    for (uint32_t y = 0; y < n1; y++)
      for (uint32_t z = 0; z < n2; z++) {
        const uint64_t dv = ((uint64_t)z * n1 + y) * n0 + x;
        const double val = src[((uint64_t)x * n1 + y) * n2 + z] - b->model.init[b->h_ids[inst][dv] & 7][s];
        if (b->esize == 8) ((double*)buf.data())[dv] = val;
        else ((float*)buf.data())[dv] = (float)val;
      }
  const size_t pb = g.plane * b->esize;
  char* base = (char*)b->d_u[b->cur] + (g.u_off + (uint64_t)s * g.plane) * b->esize;
  HIP_TRY(hipMemcpy2DAsync(base, pb * b->ns, buf.data(), pb, pb, (size_t)g.nz, hipMemcpyHostToDevice, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  return CS_OK;
}

int cs_batch_download_species(cs_batch* b, uint32_t inst, uint32_t s, double* dst)
{
  if (!b || !dst) return set_err(CS_EINVAL, "NULL argument");
  if (inst >= b->count) return set_err(CS_ERANGE, "realisation %u out of range", inst);
  if (s >= b->ns) return set_err(CS_ERANGE, "species %u out of range", s);
  HIP_TRY(hipSetDevice(b->device));
  const BatchGeo& g = b->geo[inst];
  const uint32_t n0 = b->n[inst][0], n1 = b->n[inst][1], n2 = b->n[inst][2];
  std::vector<char> buf(g.nv * b->esize);
  const size_t pb = g.plane * b->esize;
  const char* base = (const char*)b->d_u[b->cur] + (g.u_off + (uint64_t)s * g.plane) * b->esize;
  HIP_TRY(hipMemcpy2DAsync(buf.data(), pb, base, pb * b->ns, pb, (size_t)g.nz, hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  for (uint32_t x = 0; x < n0; x++)
    for (uint32_t y = 0; y < n1; y++)
      for (uint32_t z = 0; z < n2; z++) {
        const uint64_t dv = ((uint64_t)z * n1 + y) * n0 + x;
        dst[((uint64_t)x * n1 + y) * n2 + z] =
            (b->esize == 8 ? ((const double*)buf.data())[dv] : (double)((const float*)buf.data())[dv]) +
            b->model.init[b->h_ids[inst][dv] & 7][s];
      }
  return CS_OK;
}

int cs_batch_uniforms_needed(const cs_batch* b, uint64_t first, uint64_t n, uint64_t* count)
{
  if (!b || !count) return set_err(CS_EINVAL, "NULL argument");
  *count = b->model.rng == CS_RNG_MT ? b_gates_in(b, first, n) * b->uni_stride : 0;
  return CS_OK;
}

int cs_batch_step(cs_batch* b, uint64_t first, uint64_t n)
{
  if (!b) return set_err(CS_EINVAL, "ensemble is NULL");
  if (first == 0) return set_err(CS_EINVAL, "time indices start at 1");
  return b->esize == 8 ? b_step<double>(b, first, n, nullptr, 0) : b_step<float>(b, first, n, nullptr, 0);
}

int cs_batch_step_uniforms(cs_batch* b, uint64_t first, uint64_t n, const double* uni, uint64_t count)
{
  if (!b) return set_err(CS_EINVAL, "ensemble is NULL");
  if (first == 0) return set_err(CS_EINVAL, "time indices start at 1");
  if (count && !uni) return set_err(CS_EINVAL, "uniforms is NULL");
  static const double dummy = 0.0;
  if (!uni) uni = &dummy;
  return b->esize == 8 ? b_step<double>(b, first, n, uni, count) : b_step<float>(b, first, n, uni, count);
}

int cs_batch_synchronize(cs_batch* b)
{
  if (!b) return set_err(CS_EINVAL, "ensemble is NULL");
  HIP_TRY(hipSetDevice(b->device));
  HIP_TRY(hipStreamSynchronize(b->stream));
  return CS_OK;
}

int cs_batch_reduce_stats(cs_batch* b, double* mn, double* mx, double* tot)
{
  if (!b || !mn || !mx || !tot) return set_err(CS_EINVAL, "NULL argument");
  HIP_TRY(hipSetDevice(b->device));
  if (b->esize == 8)
    hipLaunchKernelGGL(k_batch_stats<double>, dim3(b->count), dim3(BSTAT_THREADS), 0, b->stream,
                       (const double*)b->d_u[b->cur], b->d_ids, (const DevParams<double>*)b->d_params, b->d_geo,
                       b->model.dV, b->d_stats);
  else
    hipLaunchKernelGGL(k_batch_stats<float>, dim3(b->count), dim3(BSTAT_THREADS), 0, b->stream,
                       (const float*)b->d_u[b->cur], b->d_ids, (const DevParams<float>*)b->d_params, b->d_geo,
                       b->model.dV, b->d_stats);
  HIP_TRY(hipGetLastError());
  std::vector<double> host((size_t)b->count * MAXS * MAXD * 3);
  HIP_TRY(hipMemcpyAsync(host.data(), b->d_stats, sizeof(double) * host.size(), hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  for (uint32_t i = 0; i < b->count; i++)
    for (uint32_t d = 0; d < b->nd; d++)
      for (uint32_t s = 0; s < b->ns; s++) {
        const double* h = host.data() + (((size_t)i * MAXS + s) * MAXD + d) * 3;
        const size_t o = ((size_t)i * b->nd + d) * b->ns + s;
        mn[o] = h[0];
        mx[o] = h[1];
        tot[o] = h[2];
      }
  return CS_OK;
}

int cs_batch_linescan(cs_batch* b, uint32_t species, uint32_t axis, const double* weights, uint32_t domain_mask,
                      double* proj)
{
  if (!b || !weights || !proj) return set_err(CS_EINVAL, "NULL argument");
  if (species >= b->ns) return set_err(CS_ERANGE, "species %u out of range", species);
  if (axis > 2) return set_err(CS_EINVAL, "axis must be 0, 1 or 2");
  HIP_TRY(hipSetDevice(b->device));
  std::vector<uint64_t> off(3 * (size_t)b->count);
  uint64_t nw = 0, np = 0, nq = 0;
  for (uint32_t i = 0; i < b->count; i++) {
    const BatchGeo& g = b->geo[i];
    off[i] = nw;
    off[b->count + i] = np;
    off[2 * b->count + i] = nq;
    nw += (uint64_t)g.nx + g.ny + g.nz;
    np += (uint64_t)g.nz * (axis == 0 ? g.nx : g.ny);
    nq += axis == 0 ? g.nx : axis == 1 ? g.ny : g.nz;
  }
  auto grow = [&](double** p, uint64_t& cap, uint64_t need) -> int {
    if (need <= cap && *p) return CS_OK;
    if (*p) HIP_TRY(hipFree(*p));
    *p = nullptr;
    HIP_TRY(hipMalloc(p, sizeof(double) * std::max<uint64_t>(need, 1)));
    cap = need;
    return CS_OK;
  };
  int rc;
  if ((rc = grow(&b->d_lw, b->lw_n, nw)) != CS_OK || (rc = grow(&b->d_lpart, b->lpart_n, np)) != CS_OK ||
      (rc = grow(&b->d_lproj, b->lproj_n, nq)) != CS_OK)
    return rc;
  if (!b->d_loff) HIP_TRY(hipMalloc(&b->d_loff, sizeof(uint64_t) * 3 * b->count));
  HIP_TRY(hipMemcpyAsync(b->d_lw, weights, sizeof(double) * nw, hipMemcpyHostToDevice, b->stream));
  HIP_TRY(hipMemcpyAsync(b->d_loff, off.data(), sizeof(uint64_t) * off.size(), hipMemcpyHostToDevice, b->stream));
  int maxz = 1;
  for (const auto& g : b->geo) maxz = std::max(maxz, g.nz);
  const uint64_t* woff = b->d_loff;
  const uint64_t* poff = b->d_loff + b->count;
  const uint64_t* qoff = b->d_loff + 2 * b->count;
  if (b->esize == 8)
    hipLaunchKernelGGL(k_batch_scan_planes<double>, dim3(maxz, b->count), dim3(256), 0, b->stream,
                       (const double*)b->d_u[b->cur], b->d_ids, (const DevParams<double>*)b->d_params, b->d_geo,
                       woff, b->d_lw, poff, b->d_lpart, species, axis, domain_mask);
  else
    hipLaunchKernelGGL(k_batch_scan_planes<float>, dim3(maxz, b->count), dim3(256), 0, b->stream,
                       (const float*)b->d_u[b->cur], b->d_ids, (const DevParams<float>*)b->d_params, b->d_geo,
                       woff, b->d_lw, poff, b->d_lpart, species, axis, domain_mask);
  hipLaunchKernelGGL(k_batch_scan_final, dim3(b->count), dim3(256), 0, b->stream, b->d_geo, woff, b->d_lw, poff,
                     b->d_lpart, qoff, b->d_lproj, axis);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(proj, b->d_lproj, sizeof(double) * nq, hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  return CS_OK;
}

int cs_batch_get_channel_states(cs_batch* b, uint32_t inst, uint32_t f, uint8_t* states)
{
  if (!b || !states) return set_err(CS_EINVAL, "NULL argument");
  if (inst >= b->count) return set_err(CS_ERANGE, "realisation %u out of range", inst);
  if (f >= b->nf || !b->model.fluxes[f].discrete) return set_err(CS_ERANGE, "flux %u is not a discrete flux", f);
  HIP_TRY(hipSetDevice(b->device));
  /* the channels of flux f are contiguous from flux_state_off */
  const uint32_t first = b->flux_state_off[inst][f];
  uint32_t last = b->state_off[inst] + b->nchan[inst];
  for (uint32_t g = f + 1; g < b->nf; g++)
    if (b->model.fluxes[g].discrete) { last = b->flux_state_off[inst][g]; break; }
  const uint32_t np = last - first;
  if (np) HIP_TRY(hipMemcpyAsync(states, b->d_states + first, np, hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  return CS_OK;
}

int cs_batch_get_channel_recruited(cs_batch* b, uint32_t inst, uint32_t f, uint8_t* out)
{
  if (!b || !out) return set_err(CS_EINVAL, "NULL argument");
  if (inst >= b->count) return set_err(CS_ERANGE, "realisation %u out of range", inst);
  if (f >= b->nf || !b->model.fluxes[f].discrete) return set_err(CS_ERANGE, "flux %u is not a discrete flux", f);
  HIP_TRY(hipSetDevice(b->device));
  const uint32_t first = b->flux_state_off[inst][f];
  uint32_t last = b->state_off[inst] + b->nchan[inst];
  for (uint32_t g = f + 1; g < b->nf; g++)
    if (b->model.fluxes[g].discrete) { last = b->flux_state_off[inst][g]; break; }
  if (last > first) HIP_TRY(hipMemcpyAsync(out, b->d_recruited + first, last - first, hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  return CS_OK;
}

int cs_batch_channel_summary(cs_batch* b, uint32_t* nopen, uint64_t* last_open, uint32_t* nrecruited)
{
  if (!b) return set_err(CS_EINVAL, "ensemble is NULL");
  HIP_TRY(hipSetDevice(b->device));
  if (last_open)
    HIP_TRY(hipMemcpyAsync(last_open, b->d_last_open, sizeof(uint64_t) * b->count, hipMemcpyDeviceToHost, b->stream));
  std::vector<uint8_t> rec, st;
  if (!b->init_states.empty()) {
    rec.resize(b->init_states.size());
    st.resize(b->init_states.size());
    if (nrecruited) HIP_TRY(hipMemcpyAsync(rec.data(), b->d_recruited, rec.size(), hipMemcpyDeviceToHost, b->stream));
    if (nopen) HIP_TRY(hipMemcpyAsync(st.data(), b->d_states, st.size(), hipMemcpyDeviceToHost, b->stream));
  }
  HIP_TRY(hipStreamSynchronize(b->stream));
  for (uint32_t i = 0; i < b->count; i++) {
    uint32_t k = 0, o = 0;
    for (uint32_t c = 0; c < b->nchan[i]; c++) {
      if (nrecruited) k += rec[b->state_off[i] + c];
      if (nopen) o += st[b->state_off[i] + c];
    }
    if (nrecruited) nrecruited[i] = k;
    if (nopen) nopen[i] = o;
  }
  return CS_OK;
}

int cs_batch_set_active(cs_batch* b, const uint8_t* active)
{
  if (!b || !active) return set_err(CS_EINVAL, "NULL argument");
  HIP_TRY(hipSetDevice(b->device));
  std::vector<uint8_t> act(active, active + b->count);
  HIP_TRY(hipMemcpyAsync(b->d_active, act.data(), b->count, hipMemcpyHostToDevice, b->stream));
  /* the stencil pass keeps only the active realisations' tiles */
  std::vector<BatchTile> all, keep;
  make_tiles(b, BX * b->launch.rows, b->launch.zchunk, all);
  for (const BatchTile& t : all)
    if (act[t.inst]) keep.push_back(t);
  HIP_TRY(hipMemcpyAsync(b->d_tiles, keep.data(), sizeof(BatchTile) * keep.size(), hipMemcpyHostToDevice, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  b->ntiles = (uint32_t)keep.size();
  drop_graphs(b);
  return CS_OK;
}

int cs_batch_get_launch(const cs_batch* b, cs_batch_launch* out)
{
  if (!b || !out) return set_err(CS_EINVAL, "NULL argument");
  *out = b->launch;
  return CS_OK;
}

int cs_batch_set_launch(cs_batch* b, const cs_batch_launch* in)
{
  if (!b || !in) return set_err(CS_EINVAL, "NULL argument");
  if (in->rows < 1 || in->rows > 8) return set_err(CS_EINVAL, "rows must be 1..8");
  if (in->zchunk < 1) return set_err(CS_EINVAL, "zchunk must be >= 1");
  if (in->prefetch < 1 || in->prefetch > 2) return set_err(CS_EINVAL, "prefetch must be 1 or 2");
  if (in->persistent > 1) return set_err(CS_EINVAL, "persistent must be 0 or 1");
  HIP_TRY(hipSetDevice(b->device));
  HIP_TRY(hipStreamSynchronize(b->stream));
  b->launch = *in;
  b->launch.nontemporal = in->nontemporal ? 1 : 0;
  drop_graphs(b);
  int rc = b->esize == 8 ? upload_tiles<double>(b) : upload_tiles<float>(b);
  if (rc != CS_OK) return rc;
  /* keep retired realisations out */
  std::vector<uint8_t> act(b->count);
  HIP_TRY(hipMemcpy(act.data(), b->d_active, b->count, hipMemcpyDeviceToHost));
  return cs_batch_set_active(b, act.data());
}

int cs_batch_step_period(const cs_batch* b, uint64_t* period)
{
  if (!b || !period) return set_err(CS_EINVAL, "NULL argument");
  uint64_t P = b_period(b);
  *period = P % 2 ? 2 * P : P;
  return CS_OK;
}

int cs_batch_set_graph(cs_batch* b, uint32_t on)
{
  if (!b) return set_err(CS_EINVAL, "ensemble is NULL");
  b->use_graph = on != 0;
  if (!b->use_graph) drop_graphs(b);
  return CS_OK;
}

int cs_batch_time_steps(cs_batch* b, uint64_t first, uint64_t n, double* total_ms, double* stencil_ms)
{
  if (!b || !total_ms || !stencil_ms) return set_err(CS_EINVAL, "NULL argument");
  if (first == 0) return set_err(CS_EINVAL, "time indices start at 1");
  if (b->model.rng == CS_RNG_MT && b->uni_stride) return set_err(CS_EINVAL, "timing needs Philox gating");
  if (n > 100000) return set_err(CS_ERANGE, "timing ranges are limited to 100000 steps");
  HIP_TRY(hipSetDevice(b->device));
  std::vector<hipEvent_t> ev(2 * n + 2, nullptr);
  struct Guard {
    std::vector<hipEvent_t>& v;
    ~Guard() { for (auto e : v) if (e) (void)hipEventDestroy(e); }
  } guard{ev};
  for (auto& e : ev) HIP_TRY(hipEventCreate(&e));
  HIP_TRY(hipEventRecord(ev[2 * n], b->stream));
  int rc;
  if (b->use_graph) {
    /* the whole range as the caller runs it (graphs included); the
     * stencil's share is not separated (reported as 0): time a range with
     * graphs off for it */
    rc = b->esize == 8 ? b_step<double>(b, first, n, nullptr, 0) : b_step<float>(b, first, n, nullptr, 0);
  } else {
    rc = b->esize == 8 ? b_run<double>(b, first, n, nullptr, ev.data()) : b_run<float>(b, first, n, nullptr, ev.data());
  }
  if (rc != CS_OK) return rc;
  HIP_TRY(hipEventRecord(ev[2 * n + 1], b->stream));
  HIP_TRY(hipEventSynchronize(ev[2 * n + 1]));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, ev[2 * n], ev[2 * n + 1]));
  *total_ms = ms;
  double ksum = 0.0;
  if (!b->use_graph) {
    for (uint64_t i = 0; i < n; i++) {
      HIP_TRY(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]));
      ksum += ms;
    }
    *stencil_ms = n ? ksum / (double)n : 0.0;
  } else {
    *stencil_ms = 0.0;
  }
  return CS_OK;
}

int cs_batch_set_gating(cs_batch* b, uint32_t mode, const cs_mlp* net)
{
  if (!b) return set_err(CS_EINVAL, "ensemble is NULL");
  if (mode > CS_GATE_CLOSURE) return set_err(CS_EINVAL, "unknown gating mode %u", mode);
  if (mode == CS_GATE_CLOSURE) {
    if (!net) return set_err(CS_EINVAL, "closure gating needs a network");
    if (b->model.rng == CS_RNG_MT && b->uni_stride) return set_err(CS_EINVAL, "closure gating draws Philox uniforms");
    b->d_net = (const MlpDev*)cs_internal_mlp_device(net);
  }
  b->gating = mode;
  drop_graphs(b);
  return CS_OK;
}

int cs_batch_set_flux_closure(cs_batch* b, const cs_mlp* net)
{
  if (!b) return set_err(CS_EINVAL, "ensemble is NULL");
  b->d_fnet = net ? (const MlpDev*)cs_internal_mlp_device(net) : nullptr;
  drop_graphs(b);
  return CS_OK;
}

int cs_batch_copy_channel_states(cs_batch* dst, const cs_batch* src)
{
  if (!dst || !src) return set_err(CS_EINVAL, "NULL argument");
  if (dst->count != src->count || dst->nchan != src->nchan)
    return set_err(CS_EINVAL, "the ensembles' realisations have different channels");
  if (dst->device != src->device) return set_err(CS_EINVAL, "the ensembles live on different devices");
  HIP_TRY(hipSetDevice(dst->device));
  const size_t n = dst->init_states.size();
  if (!n) return CS_OK;
  /* the source's pending steps first, then the copies on dst's stream */
  HIP_TRY(hipStreamSynchronize(src->stream));
  HIP_TRY(hipMemcpyAsync(dst->d_states, src->d_states, n, hipMemcpyDeviceToDevice, dst->stream));
  HIP_TRY(hipMemcpyAsync(dst->d_recruited, src->d_recruited, n, hipMemcpyDeviceToDevice, dst->stream));
  HIP_TRY(hipMemcpyAsync(dst->d_last_open, src->d_last_open, sizeof(uint64_t) * dst->count, hipMemcpyDeviceToDevice,
                         dst->stream));
  return CS_OK;
}

int cs_batch_channel_count(const cs_batch* b, uint32_t* nchan)
{
  if (!b || !nchan) return set_err(CS_EINVAL, "NULL argument");
  *nchan = b->nchan_total;
  return CS_OK;
}

int cs_batch_closure_features(cs_batch* b, float* out)
{
  if (!b || !out) return set_err(CS_EINVAL, "NULL argument");
  HIP_TRY(hipSetDevice(b->device));
  const uint32_t n = b->nchan_total;
  if (!n) return CS_OK;
  if (!b->d_feat) HIP_TRY(hipMalloc(&b->d_feat, sizeof(float) * (size_t)n * MLP_IN));
  const unsigned blocks = (unsigned)(((uint64_t)n * MLP_IN + 255) / 256);
  if (b->esize == 8)
    hipLaunchKernelGGL(k_closure_features<double>, dim3(blocks), dim3(256), 0, b->stream, (const double*)b->d_u[b->cur],
                       b->d_ids, (const DevParams<double>*)b->d_params, b->d_geo, b->d_cgeo, b->d_states,
                       b->d_nb_start, b->d_nb, n, b->d_feat);
  else
    hipLaunchKernelGGL(k_closure_features<float>, dim3(blocks), dim3(256), 0, b->stream, (const float*)b->d_u[b->cur],
                       b->d_ids, (const DevParams<float>*)b->d_params, b->d_geo, b->d_cgeo, b->d_states,
                       b->d_nb_start, b->d_nb, n, b->d_feat);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out, b->d_feat, sizeof(float) * (size_t)n * MLP_IN, hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  return CS_OK;
}

int cs_batch_channel_ca(cs_batch* b, double* out)
{
  if (!b || !out) return set_err(CS_EINVAL, "NULL argument");
  HIP_TRY(hipSetDevice(b->device));
  const uint32_t n = b->nchan_total;
  if (!n) return CS_OK;
  if (!b->d_chca) HIP_TRY(hipMalloc(&b->d_chca, sizeof(double) * n));
  if (b->esize == 8)
    hipLaunchKernelGGL(k_channel_ca<double>, dim3((n + 255) / 256), dim3(256), 0, b->stream,
                       (const double*)b->d_u[b->cur], b->d_chan, n, b->d_chca);
  else
    hipLaunchKernelGGL(k_channel_ca<float>, dim3((n + 255) / 256), dim3(256), 0, b->stream,
                       (const float*)b->d_u[b->cur], b->d_chan, n, b->d_chca);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out, b->d_chca, sizeof(double) * n, hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  return CS_OK;
}

int cs_batch_channel_sr_ca(cs_batch* b, double* out)
{
  if (!b || !out) return set_err(CS_EINVAL, "NULL argument");
  HIP_TRY(hipSetDevice(b->device));
  const uint32_t n = b->nchan_total;
  if (!n) return CS_OK;
  if (!b->d_chca) HIP_TRY(hipMalloc(&b->d_chca, sizeof(double) * n));
  if (b->esize == 8)
    hipLaunchKernelGGL(k_channel_sr_ca<double>, dim3((n + 255) / 256), dim3(256), 0, b->stream,
                       (const double*)b->d_u[b->cur], b->d_cgeo, n, b->d_chca);
  else
    hipLaunchKernelGGL(k_channel_sr_ca<float>, dim3((n + 255) / 256), dim3(256), 0, b->stream,
                       (const float*)b->d_u[b->cur], b->d_cgeo, n, b->d_chca);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out, b->d_chca, sizeof(double) * n, hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  return CS_OK;
}

int cs_batch_channel_open(cs_batch* b, uint8_t* out)
{
  if (!b || !out) return set_err(CS_EINVAL, "NULL argument");
  HIP_TRY(hipSetDevice(b->device));
  if (b->nchan_total)
    HIP_TRY(hipMemcpyAsync(out, b->d_states, b->nchan_total, hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  return CS_OK;
}

int cs_batch_flux_predictions(cs_batch* b, double* out)
{
  if (!b || !out) return set_err(CS_EINVAL, "NULL argument");
  HIP_TRY(hipSetDevice(b->device));
  if (b->nchan_total)
    HIP_TRY(hipMemcpyAsync(out, b->d_fpred, sizeof(double) * b->nchan_total, hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  return CS_OK;
}

int cs_batch_closure_predictions(cs_batch* b, double* out)
{
  if (!b || !out) return set_err(CS_EINVAL, "NULL argument");
  HIP_TRY(hipSetDevice(b->device));
  if (b->nchan_total)
    HIP_TRY(hipMemcpyAsync(out, b->d_cpred, sizeof(double) * b->nchan_total, hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  return CS_OK;
}

int cs_batch_monitor(cs_batch* b, uint32_t* flags)
{
  if (!b || !flags) return set_err(CS_EINVAL, "NULL argument");
  HIP_TRY(hipSetDevice(b->device));
  HIP_TRY(hipMemcpyAsync(flags, b->d_flags, sizeof(uint32_t) * b->count, hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  return CS_OK;
}

int cs_batch_autotune(cs_batch* b, uint64_t first, uint32_t reps, cs_batch_launch* best, double* best_ms)
{
  if (!b) return set_err(CS_EINVAL, "ensemble is NULL");
  if (reps < 1) reps = 1;
  HIP_TRY(hipSetDevice(b->device));
  /* the diffusion/reaction pattern of one period, weighted by frequency */
  std::map<std::pair<uint32_t, uint32_t>, uint32_t> pattern;
  const uint64_t P = b_period(b);
  for (uint64_t k = 0; k < P; k++) {
    const uint32_t dm = b_diffusing(b, first + k);
    pattern[{dm, b_react(b, first + k, dm)}]++;
  }
  const cs_batch_launch keep = b->launch;
  int maxz = 1;
  for (const auto& g : b->geo) maxz = std::max(maxz, g.nz);
  std::vector<cs_batch_launch> cand;
  for (uint32_t rows : {1u, 2u, 4u, 8u})
    for (uint32_t zc : {8u, 16u, 32u, (uint32_t)maxz})
      for (uint32_t pf : {1u, 2u})
        for (uint32_t nt : {0u, 1u})
          for (uint32_t ps : {0u, 1u}) cand.push_back(cs_batch_launch{rows, zc, pf, nt, ps});
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  double win = 1e30;
  cs_batch_launch wcfg = keep;
  int rc = CS_OK;
  for (const cs_batch_launch& L : cand) {
    if ((rc = cs_batch_set_launch(b, &L)) != CS_OK) break;
    double tot = 0.0;
    uint32_t wsum = 0;
    for (const auto& kv : pattern) {
      /* reads the current state, writes the spare buffer: state unchanged */
      auto one = [&]() {
        if (b->esize == 8) launch_bstencil<double>(b, b->d_u[b->cur], b->d_u[b->cur ^ 1], kv.first.first, kv.first.second);
        else launch_bstencil<float>(b, b->d_u[b->cur], b->d_u[b->cur ^ 1], kv.first.first, kv.first.second);
      };
      one();
      float ms = 0;
      if (hipEventRecord(e0, b->stream) != hipSuccess) { rc = set_err(CS_EDEVICE, "event record failed"); break; }
      for (uint32_t r = 0; r < reps; r++) one();
      if (hipEventRecord(e1, b->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
        rc = set_err(CS_EDEVICE, "timing a candidate failed");
        break;
      }
      tot += kv.second * (double)ms / reps;
      wsum += kv.second;
    }
    if (rc != CS_OK) break;
    const double mean = tot / std::max<uint32_t>(wsum, 1);
    if (mean < win) { win = mean; wcfg = L; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != CS_OK) {
    cs_batch_set_launch(b, &keep);
    return rc;
  }
  rc = cs_batch_set_launch(b, &wcfg);
  if (rc != CS_OK) return rc;
  if (best) *best = wcfg;
  if (best_ms) *best_ms = win;
  return CS_OK;
}

}  // extern "C"
