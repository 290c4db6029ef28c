/*
 * csparks.hip -- MI355X (gfx950) time stepping for calcium_sparks.
 *
 * Device layout.  Fields are stored [z][species][y][x], x fastest, where
 * (x, y, z) are the reference's axes 0, 1, 2: the ns species planes of one
 * z lie next to each other.  Each value is the deviation of the field from
 * its voxel's resting value init[domain][species], which keeps FP32 fields
 * accurate next to large resting values.  Two such stacks ping-pong: a
 * pass reads stack `cur` and writes the complete next state into the
 * other, which replaces the reference's du buffer and its apply_du copy
 * (species.c:1080-1121).  Domain ids are one byte per voxel: bits 0-2 hold
 * the domain, bit 7 is set when all six in-grid neighbours share that
 * domain (the reference's stencil_test, kernel_generic.cpp:125-173,
 * inverted).
 *
 * A one-step pass is three launches on the context stream:
 *   k_gate    (every stochastic substep) one block updates the channel
 *             states from Ca at the channel mouths (evaluate_RyR_
 *             stochastically, boundaryfluxes.c:295-359);
 *   k_stencil the memory-bound pass: one thread per (x,y) column marches
 *             z with a register ring of planes z-1, z, z+1 and writes every
 *             species of the next state: 7-point stencil (species.c:1501-
 *             1650), buffer reactions (reaction_SIMD, kernel_generic.cpp:
 *             397-509), Dirichlet faces and fixed domain species
 *             (species.c:1037-1077);
 *   k_flux    one thread per membrane voxel gathers every boundary-pair
 *             flux touching it (species.c:1704-1807) -- a gather, so no
 *             floating-point atomics and a fixed summation order.
 * A fused pass computes two steps per trip through HBM:
 *   k_step2   both steps' stencils and reactions on tiles with a one-voxel
 *             halo, the intermediate state kept on chip;
 *   k_bridge  one block gates the channels and applies the boundary
 *             fluxes around the passes, with k_gate's and k_flux's code.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "csparks.h"
#include "cs_physics.h"

#include "cs_device.h"

thread_local std::string g_err;

int cs_internal_set_err(int code, const char* fmt, ...)
{
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

namespace {

/* ------------------------------------------------------------ kernels */

/* Launch shape of k_stencil, chosen by cs_autotune (or the default). */
struct LaunchCfg {
  int rows;      /* wavefronts per block along y (block = 64 x rows)       */
  int zchunk;    /* planes marched per block                                */
  int xcd;       /* 1: blocks b, b+8, b+16.. (one XCD) get adjacent tiles  */
  int pf;        /* planes loaded ahead of the one being computed (1 or 2)  */
  int nt;        /* 1: non-temporal stores of the next state               */
  int lds;       /* 1: x/y neighbours from an LDS plane tile                */
  int lin;       /* 1: waves take 64 consecutive (x,y) columns of the plane,
                    wrapping across rows, instead of 64 x rows tiles       */
  int yb;        /* >0: row-blocked kernel, yb rows per lane and `rows`
                    wavefronts side by side along x per block              */
  int fuse;      /* 1: pairs of eligible steps run as one k_step2 pass     */
  int fz;        /* planes per block of k_step2                             */
  int fx;        /* xcd_tile map of k_step2's blocks                        */
  int ff;        /* 1: k_step2 compiled for the pass's step masks, if any   */
};

template <typename T, int NS, int PF, bool NT>
__global__ void __launch_bounds__(BX * 8)
k_stencil(const T* __restrict__ u, T* __restrict__ un, const uint8_t* __restrict__ ids,
          const DevParams<T>* __restrict__ P, Geo g, uint32_t dmask, uint32_t react,
          int zchunk, int gx, int gy, int nblocks, int xcd, int lin)
{
  __shared__ BlockTables<T, NS> S;

  /* tile of this block; with xcd set, the hardware's round-robin of blocks
   * over the 8 XCDs hands each XCD a contiguous run of tiles, so halo rows
   * shared by neighbouring tiles hit in that XCD's L2 */
  int bid = blockIdx.x;
  bid = xcd_tile(bid, nblocks, xcd);
  if (bid >= nblocks) return;
  const int bx = bid % gx, by = (bid / gx) % gy, bz = bid / (gx * gy);

  const int nthr = BX * blockDim.y;
  const int tid = threadIdx.y * BX + threadIdx.x;
  S.load(P, tid, nthr);
  __syncthreads();

  int x, y;
  if (lin) {
    /* linear: gx counts blocks of 64*rows consecutive plane columns */
    const int64_t c = ((int64_t)(bid % gx) * (int)blockDim.y + threadIdx.y) * BX + threadIdx.x;
    if (c >= (int64_t)g.plane) return;
    y = (int)(c / g.nx);
    x = (int)(c - (int64_t)y * g.nx);
  } else {
    x = bx * BX + threadIdx.x;
    y = by * (int)blockDim.y + threadIdx.y;
    if (x >= g.nx || y >= g.ny) return;
  }
  const int z0 = (lin ? bid / gx : bz) * zchunk;
  const int z1 = min(g.nz, z0 + zchunk);
  if (z0 >= z1) return;

  stencil_columns<T, NS, PF, NT>(u, un, ids, P, S, g, dmask, react, x, y, z0, z1);
}

/* LDS-tiled variant: the block's 64 x rows columns of plane z, with a
 * one-voxel halo, are staged in a double-buffered LDS tile, so every value
 * is loaded from memory once per plane (plus the halo) and the x/y
 * neighbours come from LDS; z neighbours stay in the register ring.  Halo
 * coordinates are clamped to the grid, which makes a missing neighbour
 * equal to the voxel itself (no-flux faces).  Threads past the grid edge
 * keep participating in the barriers but store nothing. */
template <typename T, int NS, bool NT>
__global__ void __launch_bounds__(BX * 8)
k_stencil_lds(const T* __restrict__ u, T* __restrict__ un, const uint8_t* __restrict__ ids,
              const DevParams<T>* __restrict__ P, Geo g, uint32_t dmask, uint32_t react,
              int zchunk, int gx, int gy, int nblocks, int xcd)
{
  constexpr int TW = BX + 2;
  extern __shared__ unsigned char lds_raw[];
  __shared__ BlockTables<T, NS> S;

  int bid = blockIdx.x;
  bid = xcd_tile(bid, nblocks, xcd);
  if (bid >= nblocks) return;
  const int rows = (int)blockDim.y;
  const int bx = bid % gx, by = (bid / gx) % gy, bz = bid / (gx * gy);
  const int nthr = BX * rows;
  const int tid = threadIdx.y * BX + threadIdx.x;
  S.load(P, tid, nthr);

  /* tile[buf][s][row 0..rows+1][col 0..65] */
  T* tile = reinterpret_cast<T*>(lds_raw);
  const int tstride = (rows + 2) * TW;        /* one species plane */
  const int bstride = NS * tstride;           /* one buffer         */

  const int tx = threadIdx.x, ty = threadIdx.y;
  const int x = bx * BX + tx, y = by * rows + ty;
  const bool inside = x < g.nx && y < g.ny;
  const int xc = min(x, g.nx - 1), yc = min(y, g.ny - 1);
  const int z0 = bz * zchunk;
  const int z1 = min(g.nz, z0 + zchunk);
  const uint64_t col = (uint64_t)yc * g.nx + xc;
  /* halo cells this thread loads: an x halo (lanes 0 and 63) and a y halo
   * (first and last wave of the block), at clamped coordinates */
  const bool hx = tx == 0 || tx == BX - 1;
  const int hxc = tx == 0 ? max(x - 1, 0) : min(x + 1, g.nx - 1);
  const uint64_t hxcol = (uint64_t)yc * g.nx + hxc;
  const int hxslot = (ty + 1) * TW + (tx == 0 ? 0 : TW - 1);
  const bool hy = ty == 0 || ty == rows - 1;
  const int hyc = ty == 0 ? max(y - 1, 0) : min(y + 1, g.ny - 1);
  const uint64_t hycol = (uint64_t)hyc * g.nx + xc;
  const int hyslot = (ty == 0 ? 0 : rows + 1) * TW + tx + 1;
  const int myslot = (ty + 1) * TW + tx + 1;
  /* rows == 1: the single wave loads both y halos */
  const bool hy2 = rows == 1;
  const uint64_t hy2col = (uint64_t)min(y + 1, g.ny - 1) * g.nx + xc;
  const int hy2slot = 2 * TW + tx + 1;

  const int nb = react == 1 ? (int)P->nb : 0;
  const uint32_t rbound = P->rbound;
  const T dtr = P->dt_react;
  const bool dirichlet = P->dirichlet != 0;
  const bool xyface = x == 0 || y == 0 || x + 1 == g.nx || y + 1 == g.ny;
  const int ztop_plane = g.nz - 1;

  T um[NS], uc[NS], up[NS];
  T hxv[NS], hyv[NS], hy2v[NS];
  {
    /* z-group offsets: species s of plane z starts at (z*NS + s)*plane */
    const uint64_t pc = (uint64_t)z0 * NS * g.plane, pm = (uint64_t)max(z0 - 1, 0) * NS * g.plane;
#pragma unroll
    for (int s = 0; s < NS; s++) {
      uc[s] = u[pc + s * g.plane + col];
      um[s] = u[pm + s * g.plane + col];
      if (hx) hxv[s] = u[pc + s * g.plane + hxcol];
      if (hy) hyv[s] = u[pc + s * g.plane + hycol];
      if (hy2) hy2v[s] = u[pc + s * g.plane + hy2col];
    }
  }
  int buf = 0;
  for (int z = z0; z < z1; z++) {
    const uint64_t pz = (uint64_t)z * g.plane;
    const uint64_t v = pz + col;
    const bool ztop = z < ztop_plane;
    /* stage plane z */
    T* tb = tile + buf * bstride;
#pragma unroll
    for (int s = 0; s < NS; s++) {
      tb[s * tstride + myslot] = uc[s];
      if (hx) tb[s * tstride + hxslot] = hxv[s];
      if (hy) tb[s * tstride + hyslot] = hyv[s];
      if (hy2) tb[s * tstride + hy2slot] = hy2v[s];
    }
    /* prefetch plane z+1 (clamped) */
    {
      const uint64_t pp = (uint64_t)min(z + 1, ztop_plane) * NS * g.plane;
#pragma unroll
      for (int s = 0; s < NS; s++) {
        up[s] = u[pp + s * g.plane + col];
        if (z + 1 < z1) {
          if (hx) hxv[s] = u[pp + s * g.plane + hxcol];
          if (hy) hyv[s] = u[pp + s * g.plane + hycol];
          if (hy2) hy2v[s] = u[pp + s * g.plane + hy2col];
        }
      }
    }
    const uint8_t id = ids[v];
    __syncthreads();
    if (inside) {
      const int d = id & 7;
      const bool uni = (id & 0x80) != 0;
      int dxm = d, dxp = d, dym = d, dyp = d, dzm = d, dzp = d;
      if (!uni) {
        dxm = ids[v - (x > 0 ? 1 : 0)] & 7;
        dxp = ids[v + (x + 1 < g.nx ? 1 : 0)] & 7;
        dym = ids[v - (y > 0 ? (uint64_t)g.nx : 0)] & 7;
        dyp = ids[v + (y + 1 < g.ny ? (uint64_t)g.nx : 0)] & 7;
        dzm = z > 0 ? ids[v - g.plane] & 7 : d;
        dzp = ztop ? ids[v + g.plane] & 7 : d;
      }
      T nw[NS];
#pragma unroll
      for (int s = 0; s < NS; s++) {
        T acc = uc[s];
        if ((dmask >> s) & 1u) {
          const T* ts = tb + s * tstride + myslot;
          const T c = uc[s];
          const T exm = ts[-1] - c, exp_ = ts[1] - c;
          const T eym = ts[-TW] - c, eyp = ts[TW] - c;
          const T ezm = um[s] - c, ezp = up[s] - c;
          const T* C = S.coef[s] + d * MAXD;
          if (uni) {
            acc = c + C[d] * ((ezm + ezp) + (eym + eyp) + (exm + exp_));
          } else {
            /* deviations of different domains differ by their resting values */
            const T* D = S.dinit[s] + d * MAXD;
            acc = c + ((C[dzm] * (ezm + D[dzm]) + C[dzp] * (ezp + D[dzp])) +
                       (C[dym] * (eym + D[dym]) + C[dyp] * (eyp + D[dyp])) +
                       (C[dxm] * (exm + D[dxm]) + C[dxp] * (exp_ + D[dxp])));
          }
        }
        nw[s] = acc;
      }
      if (react == 2) react_species_major<T, NS>(nw, uc, S, d, rbound);
      else react_general<T, NS>(nw, uc, S, d, nb, dmask, dtr);
      if (dirichlet && (xyface || z == 0 || !ztop)) nw[0] = T(0);   /* the resting value */
      const uint32_t fm = S.fixed[d];
      if (fm) {
#pragma unroll
        for (int s = 0; s < NS; s++)
          if ((fm >> s) & 1u) nw[s] = T(0);
      }
#pragma unroll
      for (int s = 0; s < NS; s++) store_next(un + ((uint64_t)z * NS + s) * g.plane + col, nw[s], NT);
    }
#pragma unroll
    for (int s = 0; s < NS; s++) {
      um[s] = uc[s];
      uc[s] = up[s];
    }
    buf ^= 1;
  }
}

/* Row-blocked variant.  A wavefront owns 64 consecutive x of Y consecutive
 * rows and marches z; a block is `rows` wavefronts side by side along x, so
 * a block of ceil(nx/64) wavefronts covers whole rows.  Per plane and
 * species each lane loads its Y own values (fully used 256-byte rows), the
 * two rows bordering the Y rows (y halo) and, in lanes 0 and 63 only, the
 * x neighbour beyond the wavefront; every other x neighbour comes from the
 * adjacent lane through DPP and every y neighbour from the lane's own
 * registers.  All loads of a plane are issued before the first wait.
 * Coordinates are clamped (rows past ny, lanes past nx and planes past the
 * grid repeat the last one), so a missing neighbour equals the voxel itself:
 * the no-flux faces, with no per-face selects.  Byte offsets within one
 * species plane are 32-bit (the host checks plane * sizeof(T) < 2^31). */
template <typename T, int NS, int Y, int PF, bool NT>
__global__ void __launch_bounds__(BX * 8)
k_stencil_rows(const T* __restrict__ u, T* __restrict__ un, const uint8_t* __restrict__ ids,
               const DevParams<T>* __restrict__ P, Geo g, uint32_t dmask, uint32_t react,
               int zchunk, int gx, int gy, int nblocks, int xcd)
{
  __shared__ BlockTables<T, NS> S;

  int bid = blockIdx.x;
  bid = xcd_tile(bid, nblocks, xcd);
  if (bid >= nblocks) return;
  const int nthr = BX * blockDim.y;
  const int tid = threadIdx.y * BX + threadIdx.x;
  S.load(P, tid, nthr);
  __syncthreads();

  const int bx = bid % gx, by = (bid / gx) % gy, bz = bid / (gx * gy);
  const int lane = threadIdx.x;
  const int xw = (bx * (int)blockDim.y + (int)threadIdx.y) * BX;   /* first x of the wave */
  if (xw >= g.nx) return;                                           /* whole wavefront     */
  const int xr = xw + lane;
  const bool xin = xr < g.nx;
  const int x = xin ? xr : g.nx - 1;
  /* lane 0 fetches x-1, lane 63 fetches x+1 (both clamped) */
  const int xe = lane == 0 ? max(xr - 1, 0) : min(xr + 1, g.nx - 1);
  const bool edge_lane = lane == 0 || lane == BX - 1;
  const int y0 = by * Y;
  const int z0 = bz * zchunk;
  const int z1 = min(g.nz, z0 + zchunk);
  if (z0 >= z1) return;

  uint32_t roff[Y];   /* byte offset of (x, row r) within a species plane */
#pragma unroll
  for (int r = 0; r < Y; r++) roff[r] = ((uint32_t)min(y0 + r, g.ny - 1) * (uint32_t)g.nx + (uint32_t)x) * (uint32_t)sizeof(T);
  const uint32_t hmoff = ((uint32_t)max(y0 - 1, 0) * (uint32_t)g.nx + (uint32_t)x) * (uint32_t)sizeof(T);
  const uint32_t hpoff = ((uint32_t)min(y0 + Y, g.ny - 1) * (uint32_t)g.nx + (uint32_t)x) * (uint32_t)sizeof(T);
  const int32_t edelta = (xe - x) * (int32_t)sizeof(T);

  const int nb = react == 1 ? (int)P->nb : 0;
  const uint32_t rbound = P->rbound;
  const T dtr = P->dt_react;
  const bool dirichlet = P->dirichlet != 0;
  const bool xface = xr == 0 || xr + 1 == g.nx;
  const int ztop_plane = g.nz - 1;
  const size_t pstride = g.plane;   /* elements per plane */

  /* descriptors over the planes this block touches (see k_stencil) */
  const int zb = max(z0 - 1, 0), ze = min(g.nz, z1 + 2);
  const uint32_t ES = (uint32_t)sizeof(T);
  const uint32_t pbytes = (uint32_t)pstride * ES;
  const uint32_t gbytes = (uint32_t)NS * pbytes;   /* one z-group of all species */
  const Rsrc rin = make_rsrc(u + (size_t)zb * NS * pstride, (uint32_t)(ze - zb) * gbytes);
  const Rsrc rout = make_rsrc(un + (size_t)z0 * NS * pstride, (uint32_t)(z1 - z0) * gbytes);
  const Rsrc rid = make_rsrc(ids + (size_t)zb * pstride, (uint32_t)(ze - zb) * (uint32_t)pstride);
  uint32_t vox[Y];    /* voxel index of (x, row r) within a plane */
#pragma unroll
  for (int r = 0; r < Y; r++) vox[r] = roff[r] / ES;

  /* register ring over z (own rows): cm = z-1, cc = z, cp = z+1, cq = z+2 */
  T cm[NS][Y], cc[NS][Y], cp[NS][Y], cq[NS][Y];
#pragma unroll
  for (int s = 0; s < NS; s++)
#pragma unroll
    for (int r = 0; r < Y; r++) {
      cc[s][r] = bload<T>(rin, roff[r], (uint32_t)(z0 - zb) * gbytes + s * pbytes);
      cm[s][r] = bload<T>(rin, roff[r], (uint32_t)(max(z0 - 1, 0) - zb) * gbytes + s * pbytes);
      if (PF == 2) cp[s][r] = bload<T>(rin, roff[r], (uint32_t)(min(z0 + 1, ztop_plane) - zb) * gbytes + s * pbytes);
    }
  /* Drain the prologue loads here.  Otherwise the wait the compiler places
   * at the loop header for them also covers the previous plane's stores
   * (stores count in vmcnt on gfx9), so every plane would wait for its own
   * stores to complete. */
  __builtin_amdgcn_s_waitcnt(0);
  for (int z = z0; z < z1; z++) {
    const uint32_t sz = (uint32_t)(z - zb) * gbytes;
    const uint32_t szi = (uint32_t)(z - zb) * (uint32_t)pstride;
    /* this plane's y halos and x edges first, then the planes ahead, so a
     * wait for the former leaves the latter in flight */
    T hm[NS], hp[NS], ex[NS][Y];
#pragma unroll
    for (int s = 0; s < NS; s++) {
      hm[s] = cc[s][0];
      hp[s] = cc[s][Y - 1];
#pragma unroll
      for (int r = 0; r < Y; r++) ex[s][r] = cc[s][r];
      if ((dmask >> s) & 1u) {
        hm[s] = bload<T>(rin, hmoff, sz + s * pbytes);
        hp[s] = bload<T>(rin, hpoff, sz + s * pbytes);
        if (edge_lane) {
#pragma unroll
          for (int r = 0; r < Y; r++) ex[s][r] = bload<T>(rin, (uint32_t)((int32_t)roff[r] + edelta), sz + s * pbytes);
        }
      }
    }
    uint32_t idw[Y];
#pragma unroll
    for (int r = 0; r < Y; r++) idw[r] = bload_u8(rid, vox[r], szi);
    if (PF == 1) {
      const uint32_t sp = (uint32_t)(min(z + 1, ztop_plane) - zb) * gbytes;
#pragma unroll
      for (int s = 0; s < NS; s++)
#pragma unroll
        for (int r = 0; r < Y; r++) cp[s][r] = bload<T>(rin, roff[r], sp + s * pbytes);
    } else if (z + 1 < z1) {
      const uint32_t sq = (uint32_t)(min(z + 2, ztop_plane) - zb) * gbytes;
#pragma unroll
      for (int s = 0; s < NS; s++)
#pragma unroll
        for (int r = 0; r < Y; r++) cq[s][r] = bload<T>(rin, roff[r], sq + s * pbytes);
    }

    /* x neighbours of every row through DPP (all lanes active here) */
    T xm[NS][Y], xp[NS][Y];
#pragma unroll
    for (int s = 0; s < NS; s++)
#pragma unroll
      for (int r = 0; r < Y; r++) {
        xm[s][r] = dpp_from_left(cc[s][r], ex[s][r]);
        xp[s][r] = dpp_from_right(cc[s][r], ex[s][r]);
      }
    /* With one plane ahead, every value loaded this iteration is complete
     * before the first store, so only stores are outstanding at the loop
     * back edge and the ring copies need no wait there. */
    if (PF == 1) {
#pragma unroll
      for (int s = 0; s < NS; s++)
#pragma unroll
        for (int r = 0; r < Y; r++) asm volatile("" ::"v"(cp[s][r]));
    }

#pragma unroll
    for (int r = 0; r < Y; r++) {
      const int yr = y0 + r;
      const uint32_t vb = vox[r];
      const int d = idw[r] & 7;
      const bool uni = (idw[r] & 0x80) != 0;
      int dxm = d, dxp = d, dym = d, dyp = d, dzm = d, dzp = d;
      if (!uni) {
        const int yc = min(yr, g.ny - 1);
        dxm = bload_u8(rid, x > 0 ? vb - 1 : vb, szi) & 7;
        dxp = bload_u8(rid, x + 1 < g.nx ? vb + 1 : vb, szi) & 7;
        dym = bload_u8(rid, yc > 0 ? vb - (uint32_t)g.nx : vb, szi) & 7;
        dyp = bload_u8(rid, yc + 1 < g.ny ? vb + (uint32_t)g.nx : vb, szi) & 7;
        dzm = z > 0 ? bload_u8(rid, vb, szi - (uint32_t)pstride) & 7 : d;
        dzp = z < ztop_plane ? bload_u8(rid, vb, szi + (uint32_t)pstride) & 7 : d;
      }
      T uc[NS], nw[NS];
#pragma unroll
      for (int s = 0; s < NS; s++) uc[s] = cc[s][r];
#pragma unroll
      for (int s = 0; s < NS; s++) {
        T acc = uc[s];
        if ((dmask >> s) & 1u) {
          const T c = uc[s];
          const T ym = r > 0 ? cc[s][r - 1] : hm[s];
          const T yp = r < Y - 1 ? cc[s][r + 1] : hp[s];
          const T exm = xm[s][r] - c, exp_ = xp[s][r] - c;
          const T eym = ym - c, eyp = yp - c;
          const T ezm = cm[s][r] - c, ezp = cp[s][r] - c;
          const T* C = S.coef[s] + d * MAXD;
          if (uni) {
            acc = c + C[d] * ((ezm + ezp) + (eym + eyp) + (exm + exp_));
          } else {
            /* deviations of different domains differ by their resting values */
            const T* D = S.dinit[s] + d * MAXD;
            acc = c + ((C[dzm] * (ezm + D[dzm]) + C[dzp] * (ezp + D[dzp])) +
                       (C[dym] * (eym + D[dym]) + C[dyp] * (eyp + D[dyp])) +
                       (C[dxm] * (exm + D[dxm]) + C[dxp] * (exp_ + D[dxp])));
          }
        }
        nw[s] = acc;
      }
      if (react == 2) react_species_major<T, NS>(nw, uc, S, d, rbound);
      else react_general<T, NS>(nw, uc, S, d, nb, dmask, dtr);
      if (dirichlet && (xface || yr == 0 || yr + 1 == g.ny || z == 0 || z == ztop_plane)) nw[0] = T(0);   /* the resting value */
      const uint32_t fm = S.fixed[d];
      if (fm) {
#pragma unroll
        for (int s = 0; s < NS; s++)
          if ((fm >> s) & 1u) nw[s] = T(0);
      }
      if (xin && yr < g.ny) {
#pragma unroll
        for (int s = 0; s < NS; s++) bstore<T>(rout, roff[r], (uint32_t)(z - z0) * gbytes + s * pbytes, nw[s], NT);
      }
    }
#pragma unroll
    for (int s = 0; s < NS; s++)
#pragma unroll
      for (int r = 0; r < Y; r++) {
        cm[s][r] = cc[s][r];
        cc[s][r] = cp[s][r];
        if (PF == 2) cp[s][r] = cq[s][r];
      }
  }
}

/* Gate every channel of every discrete flux (fluxes in gating order,
 * channels strided over the block's threads) at step ti.  ca(to, gp)
 * gives the stored value of species 0 at the cleft voxel `to` of global
 * pair gp (species.c:2569). */
template <typename T, typename Ca>
__device__ inline void gate_channels(const DevParams<T>* __restrict__ P, const uint8_t* __restrict__ ids,
                                     const uint32_t* __restrict__ pairs, uint8_t* __restrict__ states,
                                     uint8_t* __restrict__ recruited, uint64_t ti, uint32_t D, uint32_t seed,
                                     uint32_t instance, const double* __restrict__ uniforms,
                                     uint64_t uniform_base, Ca ca)
{
  const double t = (double)ti * P->dt;
  const double dts = P->dt * D;
  const uint32_t nf = P->ndiscrete;
  for (uint32_t f = 0; f < nf; f++) {
    const FluxParams& F = P->flux[f];
    for (uint32_t ch = threadIdx.x; ch < F.npairs; ch += blockDim.x) {
      uint8_t st = states[F.state_off + ch];
      double c0 = 0.0;
      if (F.stochastic) {
        const uint32_t to = pairs[2 * (F.pair_off + ch) + 1];
        c0 = ca(to, F.pair_off + ch) + P->init64[ids[to] & 7][0];
      }
      st = gate_one(F, st, c0, ti, t, dts, seed, instance,
                    uniforms ? uniforms + uniform_base + F.uniform_off + ch : nullptr, ch);
      states[F.state_off + ch] = st;
      if (st) recruited[F.state_off + ch] = 1;
    }
  }
}

/* Channel gating: one block (gate_channels) with Ca from u. */
template <typename T>
__global__ void __launch_bounds__(GATE_THREADS)
k_gate(const T* __restrict__ u, const uint8_t* __restrict__ ids, Geo g, const uint32_t* __restrict__ pairs,
       uint8_t* __restrict__ states, uint8_t* __restrict__ recruited,
       const DevParams<T>* __restrict__ P, uint64_t ti_arg, const uint64_t* __restrict__ ti_ptr,
       uint32_t D, uint32_t seed, uint32_t instance, const double* __restrict__ uniforms,
       uint64_t uniform_base)
{
  const uint64_t ti = (ti_ptr ? *ti_ptr : 0) + ti_arg;
  gate_channels<T>(P, ids, pairs, states, recruited, ti, D, seed, instance, uniforms, uniform_base,
                   [&](uint32_t to, uint32_t) { return (double)u[elem(g, 0, to)]; });
}


/* The (species, voxel) rows of the boundary fluxes: row r lists, in a
 * fixed order, every pair touching its voxel with the sign of its
 * contribution (+ at the `to` voxel, - at an inner `from` voxel). */
struct FluxRows {
  const uint32_t* voxel;
  const uint8_t* species;
  const uint32_t* start;      /* [n + 1] */
  const uint32_t* ent_pair;   /* pair index within its flux */
  const uint8_t* ent_flux;
  const int8_t* ent_sign;
  uint32_t n;
};

/* Sum of the fluxes of row r over the fluxes in fluxmask, closed channels
 * skipped; val(s, voxel, gp, which) gives the stored value of species s
 * at the from (which 0) or to (which 1) voxel of global pair gp.  False
 * when the row takes nothing: no entry applies, its species is fixed in
 * the voxel's domain, or it is species 0 on a Dirichlet face. */
template <typename T, typename Val>
__device__ inline bool row_sum(const DevParams<T>* __restrict__ P, const uint8_t* __restrict__ ids, const Geo& g,
                               const FluxRows& rows, uint32_t r, const uint32_t* __restrict__ pairs,
                               const uint8_t* __restrict__ states, uint32_t fluxmask, double& sum, Val val)
{
  const uint32_t v = rows.voxel[r];
  const uint32_t s = rows.species[r];
  const int d = ids[v] & 7;
  if ((P->fixed_mask[d] >> s) & 1u) return false;
  if (s == 0 && P->dirichlet) {
    const uint32_t x = v % g.nx, y = (v / g.nx) % g.ny, z = (uint32_t)(v / g.plane);
    if (x == 0 || y == 0 || z == 0 || x + 1 == (uint32_t)g.nx || y + 1 == (uint32_t)g.ny ||
        z + 1 == (uint32_t)g.nz)
      return false;
  }
  const double dt = P->dt, h = P->h;
  bool any = false;
  sum = 0.0;
  for (uint32_t e = rows.start[r]; e < rows.start[r + 1]; e++) {
    const uint32_t f = rows.ent_flux[e];
    if (!((fluxmask >> f) & 1u)) continue;
    const FluxParams& F = P->flux[f];
    const uint32_t p = rows.ent_pair[e];
    if (F.discrete && !states[F.state_off + p]) continue;
    const uint32_t gp = F.pair_off + p;
    const uint32_t from = pairs[2 * gp], to = pairs[2 * gp + 1];
    const double J = flux_j(F, val(s, from, gp, 0u), val(s, to, gp, 1u), P->init64[ids[from] & 7][s],
                            P->init64[ids[to] & 7][s], dt, h);
    sum += rows.ent_sign[e] > 0 ? J : -J;
    any = true;
  }
  return any;
}

/* Boundary fluxes gathered per (species, voxel) row and added to the
 * next state un, from the state u. */
template <typename T>
__global__ void __launch_bounds__(FLUX_THREADS)
k_flux(const T* __restrict__ u, T* __restrict__ un, const uint8_t* __restrict__ ids, FluxRows rows,
       const uint32_t* __restrict__ pairs, const uint8_t* __restrict__ states,
       const DevParams<T>* __restrict__ P, Geo g, uint32_t fluxmask)
{
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows.n) return;
  double sum;
  if (row_sum<T>(P, ids, g, rows, r, pairs, states, fluxmask, sum,
                 [&](uint32_t s, uint32_t vox, uint32_t, uint32_t) { return (double)u[elem(g, s, vox)]; })) {
    T* o = un + elem(g, rows.species[r], rows.voxel[r]);
    *o = (T)((double)*o + sum);
  }
}

/* ------------------------------------------------ fused two-step pass */

/* Voxels of the boundary pairs for the fused pass, ordered by plane
 * column (y*nx + x) and then z, so that a lane marching one column meets
 * them in order.  Pair voxel k keeps its level-1 value of species 0 in
 * slot k; row[k] is its k_flux row (every flux acts on species 0), or
 * NO_ROW. */
struct F2Voxels {
  const uint32_t* col_start;   /* [plane + 1] first pair voxel of each column */
  const uint32_t* z;           /* plane of pair voxel k                        */
  const uint32_t* row;         /* k_flux row of pair voxel k, or NO_ROW         */
};

/* What one of the two steps of a fused pass does. */
struct F2Step {
  uint32_t dmask, react, fluxmask, gate;
  uint64_t ti, uni_base;
};

/* Two time steps in one pass over HBM.  A block owns a 64 x WY tile of
 * plane columns, of which the inner (64-2) x (WY-2) are its own: every
 * lane computes step 1 at plane z (level 1) for its column, the ring of
 * one column around the inner tile included, and the inner lanes then
 * compute step 2 at plane z-1 (level 2) from level-1 values, so the
 * intermediate state never leaves the chip.  Level 0 comes from HBM with
 * two planes in flight ahead of use: the x neighbours by DPP (the first
 * and last lane load their outer one) and the y neighbours by loads.
 * Level-1 values reach the x neighbours by DPP and the y neighbours
 * through a double-buffered LDS plane (one barrier per plane); z
 * neighbours sit in register rings.  Each
 * level is k_stencil's per-voxel update.  The boundary fluxes of step 1,
 * summed per k_flux row before the pass (k_bridge), are added at level 1
 * as k_flux adds them, and the level-1 value of species 0 at every pair
 * voxel is kept in its slot: step 2's gating and fluxes need only those
 * values, and its fluxes add to the result after the pass (k_bridge),
 * since no other voxel's step-2 update reads them. */
/* Step masks of a fused pass fixed at compile time: FM1/FM2 are the
 * diffusing species of its two steps and FR the reaction mode of both, or
 * F2_DYN for values taken from s1/s2 at run time.  With fixed masks the
 * species loops keep only the species that diffuse, where run-time bits
 * leave every species' code in the loop behind a test. */
constexpr uint32_t F2_DYN = 0xFFFFFFFFu;

template <typename T, int NS, int WY, bool NT, uint32_t FM1 = F2_DYN, uint32_t FM2 = F2_DYN, uint32_t FR = F2_DYN>
__global__ void __launch_bounds__(BX * WY)
k_step2(const T* __restrict__ u, T* __restrict__ un, const uint8_t* __restrict__ ids,
        const DevParams<T>* __restrict__ P, Geo g, F2Voxels fv, const double* __restrict__ fsum,
        const uint8_t* __restrict__ fany, T* __restrict__ l1, F2Step s1, F2Step s2, int zchunk, int tx, int ty,
        int nblocks, int xcd)
{
  __shared__ BlockTables<T, NS> S;
  __shared__ T ex[2][NS][WY][BX];

  const int bid = xcd_tile(blockIdx.x, nblocks, xcd);
  if (bid >= nblocks) return;   /* whole padding blocks only: no barrier is skipped */
  const int bxi = bid % tx, byi = (bid / tx) % ty, bzi = bid / (tx * ty);
  const int lane = threadIdx.x, wy = threadIdx.y;
  S.load(P, wy * BX + lane, BX * WY);

  const int nz = g.nz;
  const int x = bxi * (BX - 2) - 1 + lane, y = byi * (WY - 2) - 1 + wy;
  const bool ingrid = x >= 0 && x < g.nx && y >= 0 && y < g.ny;
  const bool inner = ingrid && lane >= 1 && lane < BX - 1 && wy >= 1 && wy < WY - 1;
  /* lanes outside the grid read the clamped column and store nothing: at
   * level 0 they mirror the face column, so the DPP neighbour of a face
   * lane is the voxel itself (no flux through the face); at level 2 the
   * face lanes select their own value instead */
  const int xc = min(max(x, 0), g.nx - 1), yc = min(max(y, 0), g.ny - 1);
  const int z0 = bzi * zchunk, z1 = min(nz, z0 + zchunk);

  const uint32_t ES = (uint32_t)sizeof(T);
  const uint32_t col = (uint32_t)yc * (uint32_t)g.nx + (uint32_t)xc;
  const uint32_t bc = col * ES;
  const uint32_t bxm = (xc > 0 ? col - 1 : col) * ES, bxp = (xc + 1 < g.nx ? col + 1 : col) * ES;
  const uint32_t bym = (yc > 0 ? col - g.nx : col) * ES, byp = (yc + 1 < g.ny ? col + g.nx : col) * ES;
  const bool has_xm = x > 0, has_xp = x + 1 < g.nx;
  const bool xyface = xc == 0 || yc == 0 || xc + 1 == g.nx || yc + 1 == g.ny;
  /* the first and last lane load their outer x neighbour at level 0 */
  const bool edge_lane = (lane == 0 && xc > 0) || (lane == BX - 1 && xc + 1 < g.nx);
  const uint32_t bedge = lane == 0 ? bxm : bxp;
  /* level 2: whether this wavefront holds a lane on an x face */
  const bool xclamp = __ballot(ingrid && (!has_xm || !has_xp)) != 0;
  /* LDS rows of the y neighbours (the voxel itself at the grid's y faces;
   * the tile's outer rows use their own row and are never inner) */
  const int wym = (y > 0 && wy > 0) ? wy - 1 : wy, wyp = (y + 1 < g.ny && wy + 1 < WY) ? wy + 1 : wy;

  /* level 0 spans planes [z0-2, z1+2) of u; the host keeps (zchunk + 4)
   * groups below 2^32 bytes */
  const int zb = max(z0 - 2, 0), ze = min(nz, z1 + 2);
  const size_t plane = g.plane;
  const uint32_t pbytes = (uint32_t)plane * ES, gbytes = (uint32_t)NS * pbytes;
  const Rsrc rin = make_rsrc(u + (size_t)zb * NS * plane, (uint32_t)(ze - zb) * gbytes);
  const Rsrc rout = make_rsrc(un + (size_t)z0 * NS * plane, (uint32_t)(z1 - z0) * gbytes);
  const Rsrc rid = make_rsrc(ids + (size_t)zb * plane, (uint32_t)(ze - zb) * (uint32_t)plane);
  auto gofs = [&](int zz) { return (uint32_t)(min(max(zz, zb), ze - 1) - zb) * gbytes; };
  auto iofs = [&](int zz) { return (uint32_t)(min(max(zz, zb), ze - 1) - zb) * (uint32_t)plane; };

  const uint32_t react_nb = (uint32_t)P->nb;
  uint32_t dm1 = FM1 != F2_DYN ? FM1 : s1.dmask, dm2 = FM2 != F2_DYN ? FM2 : s2.dmask;
  uint32_t re1 = FR != F2_DYN ? FR : s1.react, re2 = FR != F2_DYN ? FR : s2.react, rbound = P->rbound;
  __shared__ uint32_t msk[5];
  if (lane == 0 && wy == 0) {
    msk[0] = dm1;
    msk[1] = dm2;
    msk[2] = re1;
    msk[3] = re2;
    msk[4] = rbound;
  }
  const T dtr = P->dt_react;
  const bool dirichlet = P->dirichlet != 0;

  /* pair voxels of this column from the first level-1 plane on */
  const int zs = max(z0 - 1, 0);
  uint32_t pk = 0, pend = 0;
  if (ingrid) {
    pk = fv.col_start[col];
    pend = fv.col_start[col + 1];
  }
  while (pk < pend && (int)fv.z[pk] < zs) pk++;
  int pz = pk < pend ? (int)fv.z[pk] : 0x7FFFFFFF;

  /* One level at plane zz: vm, vc, vp are the level's input at planes
   * zz-1, zz, zz+1 of this column, ym/yp its y neighbours at zz; the
   * result (the next level at zz) goes to nw. */
  auto level = [&](uint32_t dmask, uint32_t react, const bool first, const int zz, const uint32_t id,
                   const T (&vm)[NS], const T (&vc)[NS], const T (&vp)[NS], const T (&ym)[NS], const T (&yp)[NS],
                   const T (&xe)[NS], T (&nw)[NS]) {
    const int d = id & 7;
    const bool uni = (id & 0x80) != 0;
    const bool ztop = zz < nz - 1;
    const uint32_t szi = iofs(zz);
#pragma unroll
    for (int s = 0; s < NS; s++) nw[s] = vc[s];
    T nxm[NS], nxp[NS];
#pragma unroll
    for (int s = 0; s < NS; s++)
      if ((dmask >> s) & 1u) {
        if (first) {
          nxm[s] = dpp_from_left(vc[s], xe[s]);
          nxp[s] = dpp_from_right(vc[s], xe[s]);
        } else {
          /* level 2: the first and last lane are never inner, so their
           * missing neighbour may read 0, which lets the move fold into
           * the difference that uses it */
          nxm[s] = dpp_from_left_zero(vc[s]);
          nxp[s] = dpp_from_right_zero(vc[s]);
          if (xclamp) {
            nxm[s] = has_xm ? nxm[s] : vc[s];
            nxp[s] = has_xp ? nxp[s] : vc[s];
          }
        }
      }
    if (__ballot(!uni) == 0) {
#pragma unroll
      for (int s = 0; s < NS; s++)
        if ((dmask >> s) & 1u) {
          const T c = vc[s];
          const T exm = nxm[s] - c, exp_ = nxp[s] - c;
          const T eym = ym[s] - c, eyp = yp[s] - c;
          const T ezm = vm[s] - c, ezp = vp[s] - c;
          nw[s] = c + S.coef[s][d * MAXD + d] * ((ezm + ezp) + (eym + eyp) + (exm + exp_));
        }
    } else {
      int dxm = d, dxp = d, dym = d, dyp = d, dzm = d, dzp = d;
      if (!uni) {
        dxm = bload_u8(rid, bxm / ES, szi) & 7;
        dxp = bload_u8(rid, bxp / ES, szi) & 7;
        dym = bload_u8(rid, bym / ES, szi) & 7;
        dyp = bload_u8(rid, byp / ES, szi) & 7;
        dzm = zz > 0 ? bload_u8(rid, col, iofs(zz - 1)) & 7 : d;
        dzp = ztop ? bload_u8(rid, col, iofs(zz + 1)) & 7 : d;
      }
#pragma unroll
      for (int s = 0; s < NS; s++)
        if ((dmask >> s) & 1u) {
          const T c = vc[s];
          const T exm = nxm[s] - c, exp_ = nxp[s] - c;
          const T eym = ym[s] - c, eyp = yp[s] - c;
          const T ezm = vm[s] - c, ezp = vp[s] - c;
          const T* C = S.coef[s] + d * MAXD;
          if (uni) {
            nw[s] = c + C[d] * ((ezm + ezp) + (eym + eyp) + (exm + exp_));
          } else {
            const T* Dn = S.dinit[s] + d * MAXD;
            nw[s] = c + ((C[dzm] * (ezm + Dn[dzm]) + C[dzp] * (ezp + Dn[dzp])) +
                         (C[dym] * (eym + Dn[dym]) + C[dyp] * (eyp + Dn[dyp])) +
                         (C[dxm] * (exm + Dn[dxm]) + C[dxp] * (exp_ + Dn[dxp])));
          }
        }
    }
    if (react == 2) {
      react_species_major<T, NS>(nw, vc, S, d, rbound);
    } else if (react == 1) {
      react_general<T, NS>(nw, vc, S, d, (int)react_nb, dmask, dtr);
    }
    if (dirichlet && (xyface || zz == 0 || !ztop)) nw[0] = T(0);   /* the resting value */
    const uint32_t fm = S.fixed[d];
    if (fm) {
#pragma unroll
      for (int s = 0; s < NS; s++)
        if ((fm >> s) & 1u) nw[s] = T(0);
    }
    /* level 1 at a pair voxel: step 1's flux sum, then the slot */
    if (first && pz == zz) {
      const uint32_t r = fv.row[pk];
      if (r != NO_ROW && fany[r]) nw[0] = (T)((double)nw[0] + fsum[r]);
      if (inner && zz >= z0 && zz < z1) l1[pk] = nw[0];
      pk++;
      pz = pk < pend ? (int)fv.z[pk] : 0x7FFFFFFF;
    }
  };

  /* register rings, rotated by renaming over their common period of six
   * planes: level 0 holds planes z-1 .. z+2 and receives z+3 (one slot is
   * free); level 1 holds planes z-2, z-1 and receives z.  The y (and
   * edge-lane x) neighbours of level 0 at plane z are reloaded for z+1 as
   * soon as level 1 has used them.  A lane without an outer x neighbour
   * in the grid loads its own column, the mirror of the face.  Values of
   * species a level does not diffuse are never read from the neighbour
   * registers. */
  const uint32_t bxe = edge_lane ? bedge : bc;
  T A0[NS], A1[NS], A2[NS], A3[NS], A4[NS], A5[NS];
  T Ym[NS], Yp[NS], Xe[NS];
  T B0[NS], B1[NS], B2[NS];
  uint32_t I0, I1, I2, I3, I4 = 0, I5 = 0;
#pragma unroll
  for (int s = 0; s < NS; s++) {
    A0[s] = bload<T>(rin, bc, gofs(zs - 1) + s * pbytes);
    A1[s] = bload<T>(rin, bc, gofs(zs) + s * pbytes);
    A2[s] = bload<T>(rin, bc, gofs(zs + 1) + s * pbytes);
    A3[s] = bload<T>(rin, bc, gofs(zs + 2) + s * pbytes);
    A4[s] = A5[s] = A1[s];
    Ym[s] = Yp[s] = Xe[s] = A1[s];
    if ((dm1 >> s) & 1u) {
      Ym[s] = bload<T>(rin, bym, gofs(zs) + s * pbytes);
      Yp[s] = bload<T>(rin, byp, gofs(zs) + s * pbytes);
      Xe[s] = bload<T>(rin, bxe, gofs(zs) + s * pbytes);
    }
  }
  I0 = bload_u8(rid, col, iofs(zs - 1));
  I1 = bload_u8(rid, col, iofs(zs));
  I2 = bload_u8(rid, col, iofs(zs + 1));
  I3 = bload_u8(rid, col, iofs(zs + 2));
  __builtin_amdgcn_s_waitcnt(0);
#pragma unroll
  for (int s = 0; s < NS; s++) B0[s] = B1[s] = B2[s] = A1[s];

  /* One plane: level 1 at z from level 0 (am, ac, ap = planes z-1, z,
   * z+1; Ym, Yp, Xe its neighbours at z), then level 2 at z-1 from level 1
   * (wm, wc, wn = planes z-2, z-1, z), both clamped at the grid's z faces
   * (plane nz is plane nz-1 again).  Every plane, the one past the top
   * face too, issues the same loads (clamped addresses): the level-0
   * neighbours of plane z+1, then plane z+3.  Loads complete in issue
   * order, so the next plane's wait for its neighbours does not take in
   * plane z+3, which has two planes to arrive, and no path through the
   * planes makes a wait cover loads that a later plane issued. */
  auto iter = [&](const int z, T (&am)[NS], T (&ac)[NS], T (&ap)[NS], T (&q0)[NS], T (&qn)[NS], T (&qf)[NS],
                  T (&wm)[NS], T (&wc)[NS], T (&wn)[NS], uint32_t& idm, uint32_t& ic, uint32_t& iqn) {
    (void)q0;
    (void)qf;
    __syncthreads();
    /* the step masks, read again from LDS at each plane (the barrier
     * orders the first read after the write): each species test is then a
     * bit test of one scalar register, not a flag hoisted out of the loop,
     * of which there are too many to keep in scalar registers */
    if constexpr (FM1 == F2_DYN) dm1 = __builtin_amdgcn_readfirstlane(msk[0]);
    if constexpr (FM2 == F2_DYN) dm2 = __builtin_amdgcn_readfirstlane(msk[1]);
    if constexpr (FR == F2_DYN) {
      re1 = __builtin_amdgcn_readfirstlane(msk[2]);
      re2 = __builtin_amdgcn_readfirstlane(msk[3]);
    }
    rbound = __builtin_amdgcn_readfirstlane(msk[4]);
    const bool below = z < nz;
    if (below) {
      level(dm1, re1, true, z, ic, am, ac, ap, Ym, Yp, Xe, wn);
    } else {
#pragma unroll
      for (int s = 0; s < NS; s++) wn[s] = wc[s];
    }
#pragma unroll
    for (int s = 0; s < NS; s++)
      if ((dm1 >> s) & 1u) {
        Ym[s] = bload<T>(rin, bym, gofs(z + 1) + s * pbytes);
        Yp[s] = bload<T>(rin, byp, gofs(z + 1) + s * pbytes);
        Xe[s] = bload<T>(rin, bxe, gofs(z + 1) + s * pbytes);
      }
#pragma unroll
    for (int s = 0; s < NS; s++) qn[s] = bload<T>(rin, bc, gofs(z + 3) + s * pbytes);
    iqn = bload_u8(rid, col, iofs(z + 3));
    if (below) {
#pragma unroll
      for (int s = 0; s < NS; s++)
        if ((dm2 >> s) & 1u) ex[z & 1][s][wy][lane] = wn[s];
    }
    const int zo = z - 1;
    if (zo >= z0) {
      /* y neighbours of level 1; only the species level 2 diffuses are
       * read, so the others are left as they are */
      T Lm[NS], Lp[NS], nw[NS];
#pragma unroll
      for (int s = 0; s < NS; s++) {
        Lm[s] = Lp[s] = T(0);
        if ((dm2 >> s) & 1u) {
          Lm[s] = ex[zo & 1][s][wym][lane];
          Lp[s] = ex[zo & 1][s][wyp][lane];
        }
      }
      level(dm2, re2, false, zo, idm, wm, wc, wn, Lm, Lp, wc, nw);
      if (inner) {
        const uint32_t so = (uint32_t)(zo - z0) * gbytes;
#pragma unroll
        for (int s = 0; s < NS; s++) bstore<T>(rout, bc, so + s * pbytes, nw[s], NT);
      }
    }
    /* at z = 0 the slot that holds plane z-2 next is plane 0 (clamped) */
    if (z == 0) {
#pragma unroll
      for (int s = 0; s < NS; s++) wc[s] = wn[s];
    }
  };

  /* Each plane after the first leaves the loop when it is past zt, so no
   * path runs a later plane after skipping one. */
  const int zt = z1 == nz ? nz : z1;
  for (int z = zs;; z += 6) {
    iter(z, A0, A1, A2, A3, A4, A5, B0, B1, B2, I0, I1, I4);
    if (z + 1 > zt) break;
    iter(z + 1, A1, A2, A3, A4, A5, A0, B1, B2, B0, I1, I2, I5);
    if (z + 2 > zt) break;
    iter(z + 2, A2, A3, A4, A5, A0, A1, B2, B0, B1, I2, I3, I0);
    if (z + 3 > zt) break;
    iter(z + 3, A3, A4, A5, A0, A1, A2, B0, B1, B2, I3, I4, I1);
    if (z + 4 > zt) break;
    iter(z + 4, A4, A5, A0, A1, A2, A3, B1, B2, B0, I4, I5, I2);
    if (z + 5 > zt) break;
    iter(z + 5, A5, A0, A1, A2, A3, A4, B2, B0, B1, I5, I0, I3);
    if (z + 6 > zt) break;
  }
}


/* Channel gating and boundary fluxes around fused passes, in one block
 * with barriers between the phases:
 *   tail, after a pass over steps t, t+1: k_gate and k_flux of step t+1,
 *     with the values of species 0 at the pair voxels taken from the
 *     pass's level-1 slots, the fluxes added to u (the state after t+1);
 *   head, before a pass over steps t', t'+1: k_gate of step t' on u, then
 *     the flux sum of every k_flux row from u, which the pass adds at
 *     level 1.
 * Either part may be off.  The gating (gate_one) and flux (flux_j) code,
 * the entry order and the skip rules are k_gate's and k_flux's, so the
 * results are the same.  phases selects the phases this launch runs
 * (BR_TAIL_GATE, BR_TAIL_FLUX, BR_HEAD_GATE, BR_HEAD_ROWS): one block runs
 * them all with barriers between them; a problem with more channels or
 * rows than a block takes well runs each phase as its own launch over
 * many blocks, the launches on one stream keeping the order. */
constexpr uint32_t BR_TAIL_GATE = 1, BR_TAIL_FLUX = 2, BR_HEAD_GATE = 4, BR_HEAD_ROWS = 8, BR_ALL = 15;

template <typename T>
__global__ void __launch_bounds__(BRIDGE_THREADS)
k_bridge(T* __restrict__ u, const DevParams<T>* __restrict__ P, const BChan* __restrict__ chans, uint32_t nchan,
         const BRow* __restrict__ rows, uint32_t nrows, const BEnt* __restrict__ ents, const T* __restrict__ l1,
         uint8_t* __restrict__ states, uint8_t* __restrict__ recruited, double* __restrict__ fsum,
         uint8_t* __restrict__ fany, F2Step tail, uint32_t do_tail, F2Step head, uint32_t do_head, uint32_t D,
         uint32_t seed, uint32_t instance, const double* __restrict__ uniforms, uint32_t phases)
{
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x, nth = gridDim.x * blockDim.x;
  __shared__ FluxParams F[MAXF];
  {
    const uint32_t words = (uint32_t)(sizeof(FluxParams) * MAXF / sizeof(uint64_t));
    const uint64_t* src = reinterpret_cast<const uint64_t*>(P->flux);
    uint64_t* dst = reinterpret_cast<uint64_t*>(F);
    for (uint32_t k = threadIdx.x; k < words; k += blockDim.x) dst[k] = src[k];
  }
  __syncthreads();
  const double dt = P->dt, h = P->h, dts = dt * D;

  auto gate = [&](const F2Step& st, auto ca) {
    const double t = (double)st.ti * dt;
    for (uint32_t c = tid; c < nchan; c += nth) {
      const BChan b = chans[c];
      const FluxParams& Fp = F[b.flux];
      const uint32_t k = Fp.state_off + b.ch;
      uint8_t sv = states[k];
      const double c0 = Fp.stochastic ? ca(b) + b.rest_to : 0.0;
      sv = gate_one(Fp, sv, c0, st.ti, t, dts, seed, instance,
                    uniforms ? uniforms + st.uni_base + Fp.uniform_off + b.ch : nullptr, b.ch);
      states[k] = sv;
      if (sv) recruited[k] = 1;
    }
  };
  auto row_flux = [&](uint32_t r, uint32_t fluxmask, auto val, double& sum) -> bool {
    const BRow w = rows[r];
    sum = 0.0;
    if (w.skip) return false;
    bool any = false;
    for (uint32_t e = w.start; e < w.end; e++) {
      const BEnt x = ents[e];
      const uint32_t f = x.flux_sign & 0xFFu;
      if (!((fluxmask >> f) & 1u)) continue;
      if (x.state != NO_ROW && !states[x.state]) continue;
      const double J = flux_j(F[f], val(x, false), val(x, true), x.rest_from, x.rest_to, dt, h);
      sum += (x.flux_sign >> 8) ? J : -J;
      any = true;
    }
    return any;
  };

  if (do_tail) {
    if (tail.gate && (phases & BR_TAIL_GATE)) gate(tail, [&](const BChan& b) { return (double)l1[b.slot_to]; });
    __syncthreads();
    if (tail.fluxmask && (phases & BR_TAIL_FLUX))
      for (uint32_t r = tid; r < nrows; r += nth) {
        double sum;
        if (row_flux(r, tail.fluxmask,
                     [&](const BEnt& x, bool to) { return (double)l1[to ? x.slot_to : x.slot_from]; }, sum)) {
          T* o = u + rows[r].elem;
          *o = (T)((double)*o + sum);
        }
      }
    __syncthreads();
  }
  if (do_head) {
    if (head.gate && (phases & BR_HEAD_GATE)) gate(head, [&](const BChan& b) { return (double)u[b.elem_to]; });
    __syncthreads();
    if (phases & BR_HEAD_ROWS)
      for (uint32_t r = tid; r < nrows; r += nth) {
        double sum;
        const bool any = row_flux(r, head.fluxmask,
                                  [&](const BEnt& x, bool to) { return (double)u[to ? x.elem_to : x.elem_from]; },
                                  sum);
        fsum[r] = sum;
        fany[r] = any ? 1 : 0;
      }
  }
}

/* Per-(species, domain) min / max / sum(u*dV): block partials in a fixed
 * order, then one pass over the partials (no atomics, reproducible). */
template <typename T>
__device__ inline void block_reduce3(double (&mn)[MAXD], double (&mx)[MAXD], double (&sm)[MAXD],
                                     double* sh /* [3][MAXD][STAT_THREADS/64] */)
{
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nw = STAT_THREADS / 64;
#pragma unroll
  for (int d = 0; d < MAXD; d++) {
    for (int off = 32; off > 0; off >>= 1) {
      mn[d] = fmin(mn[d], __shfl_down(mn[d], off));
      mx[d] = fmax(mx[d], __shfl_down(mx[d], off));
      sm[d] += __shfl_down(sm[d], off);
    }
    if (lane == 0) {
      sh[(0 * MAXD + d) * nw + wave] = mn[d];
      sh[(1 * MAXD + d) * nw + wave] = mx[d];
      sh[(2 * MAXD + d) * nw + wave] = sm[d];
    }
  }
}

/* Voxels per domain, passed by value to k_stats_final. */
struct DomCounts {
  uint64_t n[MAXD];
};

/* Block partials of the output reductions, per (species, domain): the
 * minimum and maximum of the stored deviations and their sum times dV.
 * The resting value is added once per domain by k_stats_final: rounding
 * is monotone, so min(u' + r) = min(u') + r exactly, and the sum is
 * sum(u') dV + n r dV.  A wavefront whose lanes share one domain updates
 * only that domain's accumulators. */
template <typename T>
__global__ void __launch_bounds__(STAT_THREADS)
k_stats(const T* __restrict__ u, const uint8_t* __restrict__ ids, Geo g, uint32_t ns, double dV,
        double* __restrict__ partial /* [grid][ns][MAXD][3] */)
{
  /* block (bx, by) takes planes by, by + gridDim.y, ... and within each the
   * columns bx*blockDim + t strided by the grid width: a fixed assignment,
   * so the FP64 sums do not depend on scheduling */
  __shared__ double sh[3 * MAXD * (STAT_THREADS / 64)];
  const int nw = STAT_THREADS / 64;
  const uint32_t blk = blockIdx.y * gridDim.x + blockIdx.x;
  const T inf = (T)INFINITY;
  for (uint32_t s = 0; s < ns; s++) {
    T mn[MAXD], mx[MAXD];
    double sm[MAXD];
#pragma unroll
    for (int d = 0; d < MAXD; d++) { mn[d] = inf; mx[d] = -inf; sm[d] = 0.0; }
    auto upd = [&](int d, T v) {
      mn[d] = v < mn[d] ? v : mn[d];
      mx[d] = v > mx[d] ? v : mx[d];
      sm[d] += (double)v;
    };
    /* eight strided elements per lane are loaded before any is used, so
     * each wavefront keeps several loads in flight */
    constexpr int U = 8;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (int z = blockIdx.y; z < g.nz; z += gridDim.y) {
      const T* us = u + ((uint64_t)z * ns + s) * g.plane;
      const uint8_t* iz = ids + (uint64_t)z * g.plane;
      for (uint64_t c0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c0 < g.plane; c0 += U * stride) {
        T v[U];
        int dv[U];
#pragma unroll
        for (int k = 0; k < U; k++) {
          const uint64_t c = c0 + k * stride;
          dv[k] = -1;   /* past the plane: no domain */
          v[k] = T(0);
          if (c < g.plane) {
            dv[k] = iz[c] & 7;
            v[k] = us[c];
          }
        }
#pragma unroll
        for (int k = 0; k < U; k++) {
          const int d0 = __builtin_amdgcn_readfirstlane(dv[k]);
          if (__ballot(dv[k] != d0) == 0) {
            switch (d0) {
              case 0: upd(0, v[k]); break;
              case 1: upd(1, v[k]); break;
              case 2: upd(2, v[k]); break;
              case 3: upd(3, v[k]); break;
              case 4: upd(4, v[k]); break;
              default: break;   /* every lane past the plane */
            }
          } else {
#pragma unroll
            for (int d = 0; d < MAXD; d++)
              if (d == dv[k]) upd(d, v[k]);
          }
        }
      }
    }
    double dmn[MAXD], dmx[MAXD];
#pragma unroll
    for (int d = 0; d < MAXD; d++) {
      dmn[d] = (double)mn[d];
      dmx[d] = (double)mx[d];
      sm[d] *= dV;
    }
    block_reduce3<T>(dmn, dmx, sm, sh);
    __syncthreads();
    if (threadIdx.x < MAXD) {
      const int d = threadIdx.x;
      double a = sh[(0 * MAXD + d) * nw], b = sh[(1 * MAXD + d) * nw], c = sh[(2 * MAXD + d) * nw];
      for (int w = 1; w < nw; w++) {
        a = fmin(a, sh[(0 * MAXD + d) * nw + w]);
        b = fmax(b, sh[(1 * MAXD + d) * nw + w]);
        c += sh[(2 * MAXD + d) * nw + w];
      }
      double* o = partial + (((uint64_t)blk * ns + s) * MAXD + d) * 3;
      o[0] = a; o[1] = b; o[2] = c;
    }
    __syncthreads();
  }
}

/* One block per (species, domain) entry: threads stride over the block
 * partials, a fixed-order tree in LDS combines them, and the resting
 * value r joins: min = min(FLT_MAX, min u' + r), max = max(FLT_MIN,
 * max u' + r) -- the reference's start values (species.c:2821-2953) --
 * and sum = sum u' dV + n r dV. */
template <typename T>
__global__ void __launch_bounds__(STAT_THREADS)
k_stats_final(const double* __restrict__ partial, uint32_t nblocks, uint32_t ns,
              const DevParams<T>* __restrict__ P, DomCounts cnt, double dV,
              double* __restrict__ out /* [ns][MAXD][3] */)
{
  __shared__ double sa[STAT_THREADS], sb[STAT_THREADS], sc[STAT_THREADS];
  const uint32_t i = blockIdx.x;   /* s * MAXD + d */
  double a = INFINITY, b = -INFINITY, c = 0.0;
  for (uint32_t k = threadIdx.x; k < nblocks; k += blockDim.x) {
    const double* p = partial + ((uint64_t)k * ns * MAXD + i) * 3;
    a = p[0] < a ? p[0] : a;
    b = p[1] > b ? p[1] : b;
    c += p[2];
  }
  sa[threadIdx.x] = a; sb[threadIdx.x] = b; sc[threadIdx.x] = c;
  __syncthreads();
  for (uint32_t w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      sa[threadIdx.x] = fmin(sa[threadIdx.x], sa[threadIdx.x + w]);
      sb[threadIdx.x] = fmax(sb[threadIdx.x], sb[threadIdx.x + w]);
      sc[threadIdx.x] += sc[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const uint32_t s = i / MAXD, d = i % MAXD;
    const double r = P->init64[d][s];
    out[3 * i + 0] = fmin((double)FLT_MAX, sa[0] + r);
    out[3 * i + 1] = fmax((double)FLT_MIN, sb[0] + r);
    out[3 * i + 2] = sc[0] + (double)cnt.n[d] * r * dV;
  }
}

/* Linescan reductions (species.c:1960-2035).  The reference's 3-D
 * Gaussian weight is a product of one weight per axis, so the linescan of
 * a species is a 1-D convolution of its weighted projection on the scan
 * axis; these kernels form that projection's partial sums in FP64, over
 * voxels whose domain bit is set in dmask, in a fixed order.
 *
 * k_linescan_rows: rows[z*ny + y] = sum_x u[z][y][x] wx[x]; one
 * wavefront per (y, z) row, butterfly reduction across the wavefront. */
template <typename T>
__global__ void __launch_bounds__(256)
k_linescan_rows(const T* __restrict__ u, uint32_t s, const uint8_t* __restrict__ ids,
                const DevParams<T>* __restrict__ P, Geo g, uint32_t dmask, const double* __restrict__ wx,
                double* __restrict__ rows)
{
  const uint64_t row = (uint64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  if (row >= (uint64_t)g.ny * (uint64_t)g.nz) return;   /* whole wavefront */
  const int lane = threadIdx.x & 63;
  const uint64_t z = row / g.ny, y = row - z * g.ny;
  const uint64_t ib = row * (uint64_t)g.nx;                                 /* ids   */
  const uint64_t ub = ((z * g.ns + s) * g.ny + y) * (uint64_t)g.nx;       /* field */
  double acc = 0.0;
  for (int x = lane; x < g.nx; x += 64) {
    const int d = ids[ib + x] & 7;
    if ((dmask >> d) & 1u) acc += ((double)u[ub + x] + P->init64[d][s]) * wx[x];
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if (lane == 0) rows[row] = acc;
}

/* k_linescan_cols: part[(z*gridDim.y + c)*nx + x] = sum over the rows y
 * of chunk c of u[z][y][x] wy[y]; one block per (plane, row chunk), a
 * lane per x. */
template <typename T>
__global__ void __launch_bounds__(256)
k_linescan_cols(const T* __restrict__ u, uint32_t s, const uint8_t* __restrict__ ids,
                const DevParams<T>* __restrict__ P, Geo g, uint32_t dmask, const double* __restrict__ wy, int ychunk,
                double* __restrict__ part)
{
  const int z = blockIdx.x, c = blockIdx.y;
  const int y0 = c * ychunk, y1 = min(g.ny, y0 + ychunk);
  const T* us = u + ((uint64_t)z * g.ns + s) * g.plane;
  for (int x = threadIdx.x; x < g.nx; x += blockDim.x) {
    double acc = 0.0;
    for (int y = y0; y < y1; y++) {
      const uint64_t v = ((uint64_t)z * g.ny + y) * g.nx + x;
      const int d = ids[v] & 7;
      if ((dmask >> d) & 1u) acc += ((double)us[(uint64_t)y * g.nx + x] + P->init64[d][s]) * wy[y];
    }
    part[((uint64_t)z * gridDim.y + c) * g.nx + x] = acc;
  }
}

/* Projection on the scan axis from the partial sums of k_linescan_cols
 * (axis 0, gyc row chunks per plane) or k_linescan_rows (axes 1, 2), one
 * thread per output, in cs_linescan's order: axis 0 proj[x] = sum_z
 * wz[z] (sum_c part[z][c][x]); axis 1 proj[y] = sum_z wz[z] rows[z][y];
 * axis 2 proj[z] = sum_y wy[y] rows[z][y].  w holds wx | wy | wz. */
__global__ void __launch_bounds__(256)
k_linescan_proj(const double* __restrict__ part, const double* __restrict__ w, Geo g, int gyc, uint32_t axis,
                double* __restrict__ proj)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const double* wy = w + g.nx;
  const double* wz = wy + g.ny;
  if (axis == 0) {
    if (k >= g.nx) return;
    double acc = 0.0;
    for (int z = 0; z < g.nz; z++) {
      double col = 0.0;
      for (int c = 0; c < gyc; c++) col += part[((uint64_t)z * gyc + c) * g.nx + k];
      acc += wz[z] * col;
    }
    proj[k] = acc;
  } else if (axis == 1) {
    if (k >= g.ny) return;
    double acc = 0.0;
    for (int z = 0; z < g.nz; z++) acc += wz[z] * part[(uint64_t)z * g.ny + k];
    proj[k] = acc;
  } else {
    if (k >= g.nz) return;
    double acc = 0.0;
    for (int y = 0; y < g.ny; y++) acc += wy[y] * part[(uint64_t)k * g.ny + y];
    proj[k] = acc;
  }
}

/* Plain copy of a known number of bytes (16 B per lane per iteration),
 * used to calibrate memory-traffic counters against algorithmic bytes. */
__global__ void k_calibration_copy(const uint4* __restrict__ a, uint4* __restrict__ b, uint64_t n16)
{
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16;
       i += (uint64_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

/* Fill every voxel with init[domain][species], i.e. a zero deviation from
 * the resting state (zero fill kernel rather than a memset, so the fill can
 * sit in a captured graph). */
template <typename T>
__global__ void k_fill_init(T* __restrict__ u, Geo g)
{
  const uint64_t n = g.nv * g.ns;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    u[i] = T(0);
}

}  // namespace

/* Scratch of cs_linescan_noalloc: the axis weights, the partial sums and
 * the projection of one call, so that call allocates nothing.  Each GPU
 * the library runs on holds its own copy until the process ends.  Host
 * code never passes this name to a kernel: it takes the device address
 * from hipGetSymbolAddress once per device (scratch_on_device). */
static __device__ __align__(256) double scratch[CS_LINESCAN_NOALLOC_MAX_DOUBLES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[MAX_DEVICES] = {nullptr};

/* Device address of `scratch` on the current device, looked up once per
 * device (nullptr with the error recorded on failure). */
double* scratch_on_device(int device)
{
  if (device < 0 || device >= MAX_DEVICES) {
    set_err(CS_EDEVICE, "device ordinal %d beyond %d", device, MAX_DEVICES);
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      set_err(CS_EDEVICE, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_ptr[device] = (double*)ptr;
  }
  return g_scratch_ptr[device];
}

}  // namespace

/* ------------------------------------------------------------ context */

struct cs_context {
  int device = 0;
  hipStream_t stream = nullptr;
  uint32_t precision = CS_PREC_F64;
  uint32_t ns = 0, nd = 0, nf = 0, ndiscrete = 0;
  uint32_t n[3] = {0, 0, 0};
  Geo geo{};
  size_t esize = 8;
  void* d_u[2] = {nullptr, nullptr};
  int cur = 0;
  uint8_t* d_ids = nullptr;
  void* d_params = nullptr;
  /* boundary fluxes */
  uint32_t* d_pairs = nullptr;
  uint8_t* d_states = nullptr;
  uint8_t* d_recruited = nullptr;   /* channel open at any gating step so far */
  uint32_t nrows = 0;
  uint32_t* d_row_voxel = nullptr;
  uint8_t* d_row_species = nullptr;
  uint32_t* d_row_start = nullptr;
  uint32_t* d_ent_pair = nullptr;
  uint8_t* d_ent_flux = nullptr;
  int8_t* d_ent_sign = nullptr;
  /* gating */
  double* d_uniforms = nullptr;
  uint64_t uniform_cap = 0;
  uint32_t stoch_per_gate = 0;
  /* reductions */
  double* d_partial = nullptr;
  double* d_stats = nullptr;
  double* d_lweights = nullptr;   /* linescan weights, nx + ny + nz        */
  double* d_lpart = nullptr;      /* linescan partial sums                  */
  double* scratch = nullptr;      /* this device's `scratch` (no-alloc scan)  */
  /* host copies */
  cs_problem prob{};
  std::vector<uint32_t> state_off, npairs;
  std::vector<uint8_t> init_states;
  uint32_t substeps[MAXS] = {0};
  uint32_t flux_substep[MAXF] = {0};
  LaunchCfg launch{4, 32, 1, 1, 0, 0, 0, 0, 0, 32, 1};
  uint32_t rfast = 0;   /* buffer reactions admit the species-major form */
  std::vector<uint8_t> h_ids;   /* domain ids in device order (host copy)  */
  uint64_t dcount[MAXD] = {0};  /* voxels per domain                        */
  /* fused two-step pass (k_step2) */
  uint32_t* d_f2_col = nullptr;      /* pair voxels by column, see F2Voxels   */
  uint32_t* d_f2_z = nullptr;
  uint32_t* d_f2_row = nullptr;
  uint32_t* d_f2_pslot = nullptr;    /* slots of each pair's from and to voxel */
  void* d_f2_l1 = nullptr;           /* level-1 species 0 at the pair voxels  */
  double* d_f2_fsum = nullptr;       /* step-1 flux sum per k_flux row        */
  uint8_t* d_f2_fany = nullptr;      /* the row takes its sum                 */
  void* d_bchan = nullptr;           /* k_bridge records: BChan, BRow, BEnt   */
  void* d_brow = nullptr;
  void* d_bent = nullptr;
  uint32_t nbchan = 0;
  bool f2_ok = false;                /* the problem admits fused passes       */
  std::string f2_why;                /* why not, when it does not             */
  uint32_t fsm = 0;                  /* species the fluxes act on (bit 0)     */
};

namespace {

/* Dynamic LDS of the tiled kernel: two buffers of ns planes of
 * (rows + 2) x 66 values; kept under 60 KiB so the default per-block limit
 * holds with the static parameter tables. */
size_t lds_tile_bytes(size_t esize, uint32_t ns, int rows)
{
  return esize * 2 * ns * (size_t)(rows + 2) * (BX + 2);
}
constexpr size_t LDS_TILE_LIMIT = 60 * 1024;

/* Largest z chunk whose planes (chunk + 3 z-groups) fit the 32-bit
 * offsets of the stencil kernels' buffer descriptors. */
int max_zchunk(const Geo& g, size_t esize)
{
  const uint64_t group = g.plane * esize * g.ns;
  return (int)std::min<uint64_t>((uint64_t)g.nz, ((1ull << 32) - 1) / group - 3);
}

/* Largest z chunk of k_step2, whose blocks span (chunk + 4) z-groups
 * under 32-bit descriptor offsets (0 when not even one plane fits). */
int max_fused_zchunk(const Geo& g, size_t esize)
{
  const uint64_t group = g.plane * esize * g.ns;
  const uint64_t fit = ((1ull << 32) - 1) / group;
  return fit > 4 ? (int)std::min<uint64_t>((uint64_t)g.nz, fit - 4) : 0;
}

/* Default launch shape, the fastest measured on MI355X for the 32M-voxel
 * single dyad in FP32.  One-step passes: linear columns, blocks of 4
 * wavefronts, 8 planes per block, groups of 16 tiles per XCD, one plane
 * ahead, non-temporal stores of the next state.  Fused passes: the z
 * chunk and XCD groups below. */
LaunchCfg default_launch(const Geo& g, size_t esize, int cus)
{
  LaunchCfg L{4, 8, 16, 1, 1, 0, 1, 0, 0, 32, 8, 1};
  (void)cus;
  L.zchunk = std::min(L.zchunk, max_zchunk(g, esize));
  /* k_step2: z chunks of 32 planes and XCD groups of up to 8 consecutive
   * tiles, small enough that every XCD gets at least three groups
   * (hardware order for small grids).  On the 400x400x200 dyad (FP32,
   * plain stores, two runs each on one box) a pass takes 459 us at chunk
   * 32 / groups 8, 463 at 40 / 32, 465 at 40 / 16, 468 at 32 / 16, 469 at
   * 40 / 8 and 476 at 32 / 32; cs_autotune times the others */
  const int wy = esize == 4 ? 16 : 8;
  const long tiles = (long)((g.nx + BX - 3) / (BX - 2)) * ((g.ny + wy - 3) / (wy - 2));
  const int fzmax = max_fused_zchunk(g, esize);
  L.fz = std::max(1, std::min({L.fz, g.nz, std::max(fzmax, 1)}));
  const long nb = tiles * ((g.nz + L.fz - 1) / L.fz);
  L.fx = 0;
  for (int G = 8; G >= 2; G /= 2)
    if (nb >= 8L * 3 * G) { L.fx = G; break; }
  return L;
}

void free_ctx(cs_context* c)
{
  if (!c) return;
  hipSetDevice(c->device);
  void* ptrs[] = {c->d_u[0], c->d_u[1], c->d_ids, c->d_params, c->d_pairs, c->d_states, c->d_recruited,
                  c->d_row_voxel, c->d_row_species, c->d_row_start, c->d_ent_pair,
                  c->d_ent_flux, c->d_ent_sign, c->d_uniforms, c->d_partial, c->d_stats,
                  c->d_lweights, c->d_lpart, c->d_f2_col, c->d_f2_z,
                  c->d_f2_row, c->d_f2_pslot, c->d_f2_l1, c->d_f2_fsum, c->d_f2_fany, c->d_bchan,
                  c->d_brow, c->d_bent};
  for (void* p : ptrs)
    if (p) hipFree(p);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
}



int validate(const cs_problem* p)
{
  if (!p) return set_err(CS_EINVAL, "problem is NULL");
  if (p->n[0] < 1 || p->n[1] < 1 || p->n[2] < 1)
    return set_err(CS_EINVAL, "grid must have at least one voxel per axis");
  const uint64_t nv = (uint64_t)p->n[0] * p->n[1] * p->n[2];
  if (nv > 0xFFFFFFFFull) return set_err(CS_ERANGE, "more than 2^32 voxels");
  if (p->num_species < 1 || p->num_species > MAXS)
    return set_err(CS_EINVAL, "num_species %u outside 1..%d", p->num_species, MAXS);
  if (p->num_domains < 1 || p->num_domains > MAXD)
    return set_err(CS_EINVAL, "num_domains %u outside 1..%d", p->num_domains, MAXD);
  if (p->precision != CS_PREC_F64 && p->precision != CS_PREC_F32)
    return set_err(CS_EINVAL, "unknown precision %u", p->precision);
  /* the stencil kernels address a block's z-groups (every species of a
   * plane) with 32-bit byte offsets: at least 4 groups must fit */
  const uint64_t group_bytes = (uint64_t)p->n[0] * p->n[1] * (p->precision == CS_PREC_F64 ? 8 : 4) *
                               (p->num_species ? p->num_species : 1);
  if (4 * group_bytes >= (1ull << 32))
    return set_err(CS_ERANGE, "an x-y plane of all species (%llu bytes) exceeds 2^30",
                   (unsigned long long)group_bytes);
  if (!p->domain_ids) return set_err(CS_EINVAL, "domain_ids is NULL");
  if (p->num_buffers > MAXS) return set_err(CS_EINVAL, "too many buffers");
  if (p->num_fluxes > MAXF) return set_err(CS_EINVAL, "too many fluxes");
  if (p->reaction_substep < 1 || p->stochastic_substep < 1)
    return set_err(CS_EINVAL, "substeps must be >= 1");
  if (!(p->dt > 0) || !(p->h > 0)) return set_err(CS_EINVAL, "dt and h must be positive");
  bool seen_density = false;
  for (uint32_t f = 0; f < p->num_fluxes; f++) {
    const cs_flux_desc& F = p->fluxes[f];
    if (F.kind != CS_FLUX_RYR && F.kind != CS_FLUX_SERCA)
      return set_err(CS_EINVAL, "flux %u: unknown kind", f);
    if (F.species >= p->num_species) return set_err(CS_EINVAL, "flux %u: bad species", f);
    if (F.substep < 1) return set_err(CS_EINVAL, "flux %u: substep must be >= 1", f);
    if (F.num_pairs && !F.pairs) return set_err(CS_EINVAL, "flux %u: pairs is NULL", f);
    if (F.discrete) {
      if (seen_density) return set_err(CS_EINVAL, "discrete fluxes must come first");
      if (F.kind != CS_FLUX_RYR) return set_err(CS_EINVAL, "flux %u: discrete flux must be RyR", f);
      if (F.num_pairs && !F.init_states) return set_err(CS_EINVAL, "flux %u: init_states is NULL", f);
    } else {
      seen_density = true;
    }
    for (uint32_t i = 0; i < 2 * F.num_pairs; i++)
      if (F.pairs[i] >= nv) return set_err(CS_ERANGE, "flux %u: pair voxel out of range", f);
  }
  for (uint64_t v = 0; v < nv; v++)
    if (p->domain_ids[v] >= p->num_domains)
      return set_err(CS_ERANGE, "domain id %u at voxel %llu out of range", p->domain_ids[v],
                     (unsigned long long)v);
  for (uint32_t s = 0; s < p->num_species; s++)
    if (p->substeps[s] > 255) return set_err(CS_EINVAL, "substep too large");
  for (uint32_t b = 0; b < p->num_buffers; b++)
    for (uint32_t d = 0; d < p->num_domains; d++)
      if (p->bsp0[b][d] >= p->num_species || p->bsp1[b][d] >= p->num_species)
        return set_err(CS_ERANGE, "buffer %u of domain %u names a species out of range", b, d);
  return CS_OK;
}

template <typename T>
int upload_params(cs_context* c)
{
  DevParams<T> dp;
  fill_params<T>(dp, &c->prob, c->ndiscrete);
  c->rfast = dp.rfast;
  if (!c->d_params) HIP_TRY(hipMalloc(&c->d_params, sizeof(DevParams<T>)));
  HIP_TRY(hipMemcpyAsync(c->d_params, &dp, sizeof dp, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));   /* dp lives on this stack frame */
  return CS_OK;
}

template <typename T>
int fill_init(cs_context* c)
{
  const int blocks = (int)std::min<uint64_t>(4096, (c->geo.nv + 255) / 256);
  for (int b = 0; b < 2; b++)
    hipLaunchKernelGGL(k_fill_init<T>, dim3(blocks), dim3(256), 0, c->stream, (T*)c->d_u[b], c->geo);
  HIP_TRY(hipGetLastError());
  if (c->ndiscrete && !c->init_states.empty()) {
    HIP_TRY(hipMemcpyAsync(c->d_states, c->init_states.data(), c->init_states.size(),
                           hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->d_recruited, c->init_states.data(), c->init_states.size(),
                           hipMemcpyHostToDevice, c->stream));
  }
  c->cur = 0;
  return CS_OK;
}

template <typename T, int NS, int PF, bool NT>
void launch_stencil_v(cs_context* c, const LaunchCfg& L, const void* uin, void* uout, uint32_t dmask,
                      uint32_t react)
{
  const Geo& g = c->geo;
  const bool lin = L.lin && !L.lds;
  const int gz = (g.nz + L.zchunk - 1) / L.zchunk;
  int gx, gy;
  if (L.yb) {
    gx = (g.nx + BX * L.rows - 1) / (BX * L.rows);
    gy = (g.ny + L.yb - 1) / L.yb;
    const int nblocks = gx * gy * gz;
    const int launch = xcd_launch(nblocks, L.xcd);
    if (L.yb == 4)
      hipLaunchKernelGGL((k_stencil_rows<T, NS, 4, PF, NT>), dim3(launch), dim3(BX, L.rows), 0, c->stream,
                         (const T*)uin, (T*)uout, c->d_ids, (const DevParams<T>*)c->d_params, g, dmask,
                         react, L.zchunk, gx, gy, nblocks, L.xcd);
    else
      hipLaunchKernelGGL((k_stencil_rows<T, NS, 2, PF, NT>), dim3(launch), dim3(BX, L.rows), 0, c->stream,
                         (const T*)uin, (T*)uout, c->d_ids, (const DevParams<T>*)c->d_params, g, dmask,
                         react, L.zchunk, gx, gy, nblocks, L.xcd);
    return;
  }
  if (lin) {
    gx = (int)((g.plane + (uint64_t)BX * L.rows - 1) / ((uint64_t)BX * L.rows));
    gy = 1;
  } else {
    gx = (g.nx + BX - 1) / BX;
    gy = (g.ny + L.rows - 1) / L.rows;
  }
  const int nblocks = gx * gy * gz;
  const int launch = xcd_launch(nblocks, L.xcd);
  if (L.lds) {
    const size_t lds = sizeof(T) * 2 * NS * (L.rows + 2) * (BX + 2);
    hipLaunchKernelGGL((k_stencil_lds<T, NS, NT>), dim3(launch), dim3(BX, L.rows), lds, c->stream,
                       (const T*)uin, (T*)uout, c->d_ids, (const DevParams<T>*)c->d_params, g, dmask,
                       react, L.zchunk, gx, gy, nblocks, L.xcd);
  } else {
    hipLaunchKernelGGL((k_stencil<T, NS, PF, NT>), dim3(launch), dim3(BX, L.rows), 0, c->stream,
                       (const T*)uin, (T*)uout, c->d_ids, (const DevParams<T>*)c->d_params, g, dmask,
                       react, L.zchunk, gx, gy, nblocks, L.xcd, lin ? 1 : 0);
  }
}

template <typename T, int NS>
void launch_stencil_ns(cs_context* c, const LaunchCfg& L, const void* uin, void* uout, uint32_t dmask,
                       uint32_t react)
{
  if (L.pf == 2 && (!L.lds || L.yb)) {
    if (L.nt) launch_stencil_v<T, NS, 2, true>(c, L, uin, uout, dmask, react);
    else launch_stencil_v<T, NS, 2, false>(c, L, uin, uout, dmask, react);
  } else {
    if (L.nt) launch_stencil_v<T, NS, 1, true>(c, L, uin, uout, dmask, react);
    else launch_stencil_v<T, NS, 1, false>(c, L, uin, uout, dmask, react);
  }
}

template <typename T>
void launch_stencil_cfg(cs_context* c, const LaunchCfg& L, const void* uin, void* uout, uint32_t dmask,
                        uint32_t react)
{
  switch (c->ns) {
    case 1: launch_stencil_ns<T, 1>(c, L, uin, uout, dmask, react); break;
    case 2: launch_stencil_ns<T, 2>(c, L, uin, uout, dmask, react); break;
    case 3: launch_stencil_ns<T, 3>(c, L, uin, uout, dmask, react); break;
    case 4: launch_stencil_ns<T, 4>(c, L, uin, uout, dmask, react); break;
    case 5: launch_stencil_ns<T, 5>(c, L, uin, uout, dmask, react); break;
    case 6: launch_stencil_ns<T, 6>(c, L, uin, uout, dmask, react); break;
    case 7: launch_stencil_ns<T, 7>(c, L, uin, uout, dmask, react); break;
    default: launch_stencil_ns<T, 8>(c, L, uin, uout, dmask, react); break;
  }
}

template <typename T>
void launch_stencil(cs_context* c, uint32_t dmask, uint32_t react)
{
  launch_stencil_cfg<T>(c, c->launch, c->d_u[c->cur], c->d_u[c->cur ^ 1], dmask, react);
}

FluxRows flux_rows(const cs_context* c)
{
  return FluxRows{c->d_row_voxel, c->d_row_species, c->d_row_start, c->d_ent_pair,
                  c->d_ent_flux, c->d_ent_sign, c->nrows};
}

/* Buffer reactions of step ti: 0 none, 1 in the reference's slot order,
 * 2 in species-major form (allowed when species 0 diffuses this step). */
uint32_t react_mode(const cs_context* c, uint64_t ti, uint32_t dmask)
{
  if (c->prob.num_buffers == 0 || ti % c->prob.reaction_substep != 0) return 0;
  return (c->rfast && (dmask & 1u)) ? 2u : 1u;
}

/* One step: gate (if due), stencil+reaction, fluxes; then swap. */
template <typename T>
int step_one(cs_context* c, uint64_t ti, const double* d_uni, uint64_t uni_base, hipEvent_t e0,
             hipEvent_t e1)
{
  const cs_problem& p = c->prob;
  uint32_t dmask = 0;
  for (uint32_t s = 0; s < c->ns; s++)
    if (c->substeps[s] && ti % c->substeps[s] == 0) dmask |= 1u << s;
  const uint32_t react = react_mode(c, ti, dmask);
  uint32_t fluxmask = 0;
  for (uint32_t f = 0; f < c->nf; f++)
    if (ti % c->flux_substep[f] == 0) fluxmask |= 1u << f;
  const T* ucur = (const T*)c->d_u[c->cur];
  if (c->ndiscrete && ti % p.stochastic_substep == 0) {
    hipLaunchKernelGGL(k_gate<T>, dim3(1), dim3(GATE_THREADS), 0, c->stream, ucur, c->d_ids, c->geo, c->d_pairs,
                       c->d_states, c->d_recruited, (const DevParams<T>*)c->d_params, ti, (const uint64_t*)nullptr,
                       p.stochastic_substep, (uint32_t)p.seed, (uint32_t)p.instance, d_uni, uni_base);
  }
  if (e0) HIP_TRY(hipEventRecord(e0, c->stream));
  launch_stencil<T>(c, dmask, react);
  if (e1) HIP_TRY(hipEventRecord(e1, c->stream));
  if (c->nrows && fluxmask) {
    const uint32_t blocks = (c->nrows + FLUX_THREADS - 1) / FLUX_THREADS;
    hipLaunchKernelGGL(k_flux<T>, dim3(blocks), dim3(FLUX_THREADS), 0, c->stream, ucur,
                       (T*)c->d_u[c->cur ^ 1], c->d_ids, flux_rows(c), c->d_pairs, c->d_states,
                       (const DevParams<T>*)c->d_params, c->geo, fluxmask);
  }
  HIP_TRY(hipGetLastError());
  c->cur ^= 1;
  return CS_OK;
}

/* Species whose diffusion substep fires at step ti. */
uint32_t diffusing(const cs_context* c, uint64_t ti)
{
  uint32_t m = 0;
  for (uint32_t s = 0; s < c->ns; s++)
    if (c->substeps[s] && ti % c->substeps[s] == 0) m |= 1u << s;
  return m;
}

F2Step step_info(const cs_context* c, uint64_t ti, uint64_t uni_base)
{
  F2Step st{};
  st.dmask = diffusing(c, ti);
  st.react = react_mode(c, ti, st.dmask);
  for (uint32_t f = 0; f < c->nf; f++)
    if (ti % c->flux_substep[f] == 0) st.fluxmask |= 1u << f;
  st.gate = (c->ndiscrete && ti % c->prob.stochastic_substep == 0) ? 1u : 0u;
  st.ti = ti;
  st.uni_base = uni_base;
  return st;
}

/* Steps ti and ti+1 may run as one k_step2 pass: the problem admits it
 * and every flux species (species 0 too when channels gate) diffuses on
 * both steps, so its neighbours are at hand at both levels. */
bool can_fuse(const cs_context* c, const LaunchCfg& L, uint64_t ti)
{
  if (!L.fuse || !c->f2_ok) return false;
  return (c->fsm & ~diffusing(c, ti)) == 0 && (c->fsm & ~diffusing(c, ti + 1)) == 0;
}

template <typename T, int NS, uint32_t FM1 = F2_DYN, uint32_t FM2 = F2_DYN, uint32_t FR = F2_DYN>
void launch_step2_k(cs_context* c, const LaunchCfg& L, const void* uin, void* uout, const F2Step& a,
                    const F2Step& b)
{
  constexpr int WY = sizeof(T) == 4 ? 16 : 8;
  const Geo& g = c->geo;
  const int tx = (g.nx + BX - 3) / (BX - 2), ty = (g.ny + WY - 3) / (WY - 2);
  const int zc = std::max(1, std::min(L.fz, g.nz));
  const int tz = (g.nz + zc - 1) / zc;
  const int nblocks = tx * ty * tz;
  const F2Voxels fv{c->d_f2_col, c->d_f2_z, c->d_f2_row};
  /* plain stores: on the 400x400x200 dyad non-temporal ones cost 2% (472 against
   * 463 us per pass at z chunk 40 and groups of 32) */
  hipLaunchKernelGGL((k_step2<T, NS, WY, false, FM1, FM2, FR>), dim3(xcd_launch(nblocks, L.fx)), dim3(BX, WY), 0,
                     c->stream, (const T*)uin, (T*)uout, c->d_ids, (const DevParams<T>*)c->d_params, g, fv,
                     (const double*)c->d_f2_fsum, (const uint8_t*)c->d_f2_fany, (T*)c->d_f2_l1, a, b, zc, tx, ty,
                     nblocks, L.fx);
}

/* Pairs of step masks compiled with fixed masks, for the species layout of
 * the reference's species files (parameters/*.h5: Ca, Fluo, CMDN-CSQN,
 * ATP, TRPN): Ca and ATP diffuse on every step (0x9), the dye and CMDN on
 * their own multiples of dt (0x2, 0x4), TRPN never.  These are every pair
 * of consecutive steps when the two slow species' substeps are at least 2;
 * the reaction runs species-major on both steps.  Other pairs, species
 * counts and FP64 run the kernel with run-time masks. */
template <typename T, int NS>
void launch_step2_ns(cs_context* c, const LaunchCfg& L, const void* uin, void* uout, const F2Step& a,
                     const F2Step& b)
{
  if constexpr (sizeof(T) == 4 && NS == 5) {
    if (L.ff && a.react == 2 && b.react == 2) {
      switch ((a.dmask << 8) | b.dmask) {
        case 0x0909: return launch_step2_k<T, NS, 0x9, 0x9, 2>(c, L, uin, uout, a, b);
        case 0x0B09: return launch_step2_k<T, NS, 0xB, 0x9, 2>(c, L, uin, uout, a, b);
        case 0x090B: return launch_step2_k<T, NS, 0x9, 0xB, 2>(c, L, uin, uout, a, b);
        case 0x0D09: return launch_step2_k<T, NS, 0xD, 0x9, 2>(c, L, uin, uout, a, b);
        case 0x090D: return launch_step2_k<T, NS, 0x9, 0xD, 2>(c, L, uin, uout, a, b);
        case 0x0F09: return launch_step2_k<T, NS, 0xF, 0x9, 2>(c, L, uin, uout, a, b);
        case 0x090F: return launch_step2_k<T, NS, 0x9, 0xF, 2>(c, L, uin, uout, a, b);
        case 0x0B0D: return launch_step2_k<T, NS, 0xB, 0xD, 2>(c, L, uin, uout, a, b);
        case 0x0D0B: return launch_step2_k<T, NS, 0xD, 0xB, 2>(c, L, uin, uout, a, b);
        default: break;
      }
    }
  }
  launch_step2_k<T, NS>(c, L, uin, uout, a, b);
}

template <typename T>
void launch_step2(cs_context* c, const LaunchCfg& L, const void* uin, void* uout, const F2Step& a,
                  const F2Step& b)
{
  switch (c->ns) {
    case 1: launch_step2_ns<T, 1>(c, L, uin, uout, a, b); break;
    case 2: launch_step2_ns<T, 2>(c, L, uin, uout, a, b); break;
    case 3: launch_step2_ns<T, 3>(c, L, uin, uout, a, b); break;
    case 4: launch_step2_ns<T, 4>(c, L, uin, uout, a, b); break;
    case 5: launch_step2_ns<T, 5>(c, L, uin, uout, a, b); break;
    case 6: launch_step2_ns<T, 6>(c, L, uin, uout, a, b); break;
    case 7: launch_step2_ns<T, 7>(c, L, uin, uout, a, b); break;
    default: break;   /* more species never fuse (f2_ok is false) */
  }
}

/* k_bridge on the state u: the tail of a pass (step `tail`) and/or the
 * head of the next (step `head`). */
template <typename T>
void launch_bridge(cs_context* c, void* u, const F2Step& tail, bool do_tail, const F2Step& head, bool do_head,
                   const double* d_uni)
{
  auto go = [&](uint32_t blocks, uint32_t phases) {
    hipLaunchKernelGGL(k_bridge<T>, dim3(blocks), dim3(BRIDGE_THREADS), 0, c->stream, (T*)u,
                       (const DevParams<T>*)c->d_params, (const BChan*)c->d_bchan, c->nbchan,
                       (const BRow*)c->d_brow, c->nrows, (const BEnt*)c->d_bent, (const T*)c->d_f2_l1, c->d_states,
                       c->d_recruited, c->d_f2_fsum, c->d_f2_fany, tail, do_tail ? 1u : 0u, head,
                       do_head ? 1u : 0u, (uint32_t)c->prob.stochastic_substep, (uint32_t)c->prob.seed,
                       (uint32_t)c->prob.instance, d_uni, phases);
  };
  const size_t most = std::max<size_t>(c->nrows, c->nbchan);
  if (most <= F2_ONE_BLOCK) {
    go(1, BR_ALL);
    return;
  }
  /* each phase over enough blocks for its channels or rows */
  auto blocks = [](size_t n) { return (uint32_t)std::max<size_t>(1, (n + BRIDGE_THREADS - 1) / BRIDGE_THREADS); };
  if (do_tail) {
    if (tail.gate) go(blocks(c->nbchan), BR_TAIL_GATE);
    if (tail.fluxmask) go(blocks(c->nrows), BR_TAIL_FLUX);
  }
  if (do_head) {
    if (head.gate) go(blocks(c->nbchan), BR_HEAD_GATE);
    go(blocks(c->nrows), BR_HEAD_ROWS);
  }
}

uint64_t gates_in(const cs_context* c, uint64_t first, uint64_t n)
{
  if (!c->ndiscrete) return 0;
  const uint64_t D = c->prob.stochastic_substep;
  /* multiples of D in [first, first+n) */
  const uint64_t last = first + n;
  auto upto = [D](uint64_t k) { return k == 0 ? 0 : (k - 1) / D + 1; }; /* multiples in [0,k) incl 0 */
  return upto(last) - upto(first);
}

template <typename T>
int run_steps(cs_context* c, uint64_t first, uint64_t n, const double* host_uni, uint64_t count,
              bool timing, double* total_ms, double* kms, uint64_t* npasses = nullptr,
              uint64_t* nfused = nullptr)
{
  HIP_TRY(hipSetDevice(c->device));
  const double* d_uni = nullptr;
  if (c->prob.rng == CS_RNG_MT && c->ndiscrete && c->stoch_per_gate) {
    const uint64_t need = gates_in(c, first, n) * c->stoch_per_gate;
    if (!host_uni) return set_err(CS_EINVAL, "this context draws gating uniforms from the caller: use cs_step_uniforms");
    if (count != need)
      return set_err(CS_EINVAL, "expected %llu uniforms, got %llu", (unsigned long long)need,
                     (unsigned long long)count);
    if (need > c->uniform_cap) {
      if (c->d_uniforms) HIP_TRY(hipFree(c->d_uniforms));
      c->d_uniforms = nullptr;
      HIP_TRY(hipMalloc(&c->d_uniforms, sizeof(double) * need));
      c->uniform_cap = need;
    }
    if (need) {
      /* the caller's array may be freed once this call returns, and a copy
       * from pageable memory can still be reading it then: wait for it */
      HIP_TRY(hipMemcpyAsync(c->d_uniforms, host_uni, sizeof(double) * need, hipMemcpyHostToDevice,
                             c->stream));
      HIP_TRY(hipStreamSynchronize(c->stream));
    }
    d_uni = c->d_uniforms;
  } else if (host_uni && count) {
    return set_err(CS_EINVAL, "this context uses Philox gating; no uniforms expected");
  }
  /* timing: one event pair around every stencil launch plus one pair
   * around the range, all on the context stream, read after the range */
  std::vector<hipEvent_t> evs;
  struct EvGuard {
    std::vector<hipEvent_t>& v;
    ~EvGuard() { for (auto e : v) if (e) hipEventDestroy(e); }
  } guard{evs};
  if (timing) {
    if (n > 100000) return set_err(CS_ERANGE, "timing ranges are limited to 100000 steps");
    evs.assign(2 * n + 2, nullptr);
    for (auto& e : evs) HIP_TRY(hipEventCreate(&e));
    HIP_TRY(hipEventRecord(evs[2 * n], c->stream));
  }
  uint64_t gate_idx = 0, passes = 0, fused = 0;
  const uint64_t spg = c->stoch_per_gate;
  auto gates = [c](uint64_t ti) { return c->ndiscrete && ti % c->prob.stochastic_substep == 0; };
  /* primed: gating and flux sums of the next pass's first step are done
   * (by the previous pass's k_bridge) */
  bool primed = false;
  for (uint64_t i = 0; i < n;) {
    const uint64_t ti = first + i;
    hipEvent_t ea = timing ? evs[2 * passes] : nullptr, eb = timing ? evs[2 * passes + 1] : nullptr;
    if (i + 1 < n && can_fuse(c, c->launch, ti)) {
      if (!primed) {
        const F2Step h = step_info(c, ti, gate_idx * spg);
        if (gates(ti)) gate_idx++;
        launch_bridge<T>(c, c->d_u[c->cur], h, false, h, true, d_uni);
      }
      const F2Step s1 = step_info(c, ti, 0);
      const F2Step s2 = step_info(c, ti + 1, gate_idx * spg);
      if (gates(ti + 1)) gate_idx++;
      if (ea) HIP_TRY(hipEventRecord(ea, c->stream));
      launch_step2<T>(c, c->launch, c->d_u[c->cur], c->d_u[c->cur ^ 1], s1, s2);
      if (eb) HIP_TRY(hipEventRecord(eb, c->stream));
      const bool next = i + 3 < n && can_fuse(c, c->launch, ti + 2);
      F2Step h2 = s2;
      if (next) {
        h2 = step_info(c, ti + 2, gate_idx * spg);
        if (gates(ti + 2)) gate_idx++;
      }
      launch_bridge<T>(c, c->d_u[c->cur ^ 1], s2, true, h2, next, d_uni);
      HIP_TRY(hipGetLastError());
      c->cur ^= 1;
      primed = next;
      i += 2;
      fused++;
    } else {
      const int rc = step_one<T>(c, ti, d_uni, gate_idx * spg, ea, eb);
      if (rc != CS_OK) return rc;
      if (gates(ti)) gate_idx++;
      primed = false;
      i += 1;
    }
    passes++;
  }
  if (timing) {
    HIP_TRY(hipEventRecord(evs[2 * n + 1], c->stream));
    HIP_TRY(hipEventSynchronize(evs[2 * n + 1]));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, evs[2 * n], evs[2 * n + 1]));
    *total_ms = ms;
    double ksum = 0.0;
    for (uint64_t i = 0; i < passes; i++) {
      HIP_TRY(hipEventElapsedTime(&ms, evs[2 * i], evs[2 * i + 1]));
      ksum += ms;
    }
    *kms = passes ? ksum / (double)passes : 0.0;
    if (npasses) *npasses = passes;
    if (nfused) *nfused = fused;
  }
  return CS_OK;
}

int build_flux_rows(cs_context* c, const cs_problem* p)
{
  struct Ent { uint64_t key; uint32_t pair; uint8_t flux; int8_t sign; };
  std::vector<Ent> ents;
  std::vector<uint32_t> pairs_dev;
  for (uint32_t f = 0; f < p->num_fluxes; f++) {
    const cs_flux_desc& F = p->fluxes[f];
    for (uint32_t i = 0; i < F.num_pairs; i++) {
      const uint32_t from = ref_to_dev(F.pairs[2 * i], c->n);
      const uint32_t to = ref_to_dev(F.pairs[2 * i + 1], c->n);
      pairs_dev.push_back(from);
      pairs_dev.push_back(to);
      const uint64_t sp = F.species;
      ents.push_back({(sp << 32) | to, i, (uint8_t)f, (int8_t)1});
      if (F.inner) ents.push_back({(sp << 32) | from, i, (uint8_t)f, (int8_t)-1});
    }
  }
  std::stable_sort(ents.begin(), ents.end(), [](const Ent& a, const Ent& b) { return a.key < b.key; });
  std::vector<uint32_t> row_voxel, row_start, ent_pair;
  std::vector<uint8_t> row_species, ent_flux;
  std::vector<int8_t> ent_sign;
  for (size_t i = 0; i < ents.size(); i++) {
    if (i == 0 || ents[i].key != ents[i - 1].key) {
      row_voxel.push_back((uint32_t)(ents[i].key & 0xFFFFFFFFu));
      row_species.push_back((uint8_t)(ents[i].key >> 32));
      row_start.push_back((uint32_t)i);
    }
    ent_pair.push_back(ents[i].pair);
    ent_flux.push_back(ents[i].flux);
    ent_sign.push_back(ents[i].sign);
  }
  row_start.push_back((uint32_t)ents.size());
  c->nrows = (uint32_t)row_voxel.size();
  auto up = [c](auto** dptr, const auto& vec) -> int {
    using E = typename std::decay<decltype(vec)>::type::value_type;
    const size_t bytes = sizeof(E) * std::max<size_t>(vec.size(), 1);
    HIP_TRY(hipMalloc((void**)dptr, bytes));
    if (!vec.empty())
      HIP_TRY(hipMemcpyAsync(*dptr, vec.data(), sizeof(E) * vec.size(), hipMemcpyHostToDevice, c->stream));
    return CS_OK;
  };
  int rc;
  if ((rc = up(&c->d_pairs, pairs_dev)) != CS_OK) return rc;
  if ((rc = up(&c->d_row_voxel, row_voxel)) != CS_OK) return rc;
  if ((rc = up(&c->d_row_species, row_species)) != CS_OK) return rc;
  if ((rc = up(&c->d_row_start, row_start)) != CS_OK) return rc;
  if ((rc = up(&c->d_ent_pair, ent_pair)) != CS_OK) return rc;
  if ((rc = up(&c->d_ent_flux, ent_flux)) != CS_OK) return rc;
  if ((rc = up(&c->d_ent_sign, ent_sign)) != CS_OK) return rc;
  /* channel states */
  c->state_off.assign(p->num_fluxes, 0);
  c->npairs.assign(p->num_fluxes, 0);
  c->init_states.clear();
  c->ndiscrete = 0;
  c->stoch_per_gate = 0;
  for (uint32_t f = 0; f < p->num_fluxes; f++) {
    const cs_flux_desc& F = p->fluxes[f];
    c->npairs[f] = F.num_pairs;
    c->flux_substep[f] = F.substep;
    if (!F.discrete) continue;
    c->ndiscrete++;
    c->state_off[f] = (uint32_t)c->init_states.size();
    for (uint32_t i = 0; i < F.num_pairs; i++) c->init_states.push_back(F.init_states[i] ? 1 : 0);
    if (!(F.params[8] > 0)) c->stoch_per_gate += F.num_pairs;
  }
  if ((rc = up(&c->d_states, c->init_states)) != CS_OK) return rc;
  if ((rc = up(&c->d_recruited, c->init_states)) != CS_OK) return rc;

  /* pair voxels of the fused pass (k_step2, k_bridge) */
  const Geo& g = c->geo;
  c->f2_ok = true;
  c->f2_why.clear();
  c->fsm = 0;
  auto why = [c](const char* m) {
    if (c->f2_ok) { c->f2_ok = false; c->f2_why = m; }
  };
  if (c->ns > 7) why("more than 7 species");
  if (max_fused_zchunk(g, c->esize) < 1) why("planes too large for 32-bit offsets");
  for (uint32_t f = 0; f < p->num_fluxes; f++)
    if (p->fluxes[f].num_pairs) {
      c->fsm |= 1u << p->fluxes[f].species;
      if (p->fluxes[f].species != 0) why("a flux acts on a species other than 0");
    }
  std::vector<uint64_t> keys;
  keys.reserve(pairs_dev.size());
  for (uint32_t v : pairs_dev) keys.push_back((uint64_t)(v % g.plane) << 32 | (uint64_t)(v / g.plane));
  std::sort(keys.begin(), keys.end());
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  std::unordered_map<uint32_t, uint32_t> slot;
  std::vector<uint32_t> col_start(g.plane + 1, 0), vz(keys.size()), vrow(keys.size(), NO_ROW);
  for (size_t k = 0; k < keys.size(); k++) {
    const uint64_t colv = keys[k] >> 32, z = keys[k] & 0xFFFFFFFFu;
    slot[(uint32_t)(z * g.plane + colv)] = (uint32_t)k;
    vz[k] = (uint32_t)z;
    col_start[colv + 1]++;
  }
  for (size_t k = 1; k < col_start.size(); k++) col_start[k] += col_start[k - 1];
  for (uint32_t r = 0; r < (uint32_t)row_voxel.size(); r++)
    if (row_species[r] == 0) vrow[slot[row_voxel[r]]] = r;
  std::vector<uint32_t> pslot(pairs_dev.size());
  for (size_t k = 0; k < pairs_dev.size(); k++) pslot[k] = slot[pairs_dev[k]];
  if ((rc = up(&c->d_f2_col, col_start)) != CS_OK) return rc;
  if ((rc = up(&c->d_f2_z, vz)) != CS_OK) return rc;
  if ((rc = up(&c->d_f2_row, vrow)) != CS_OK) return rc;
  if ((rc = up(&c->d_f2_pslot, pslot)) != CS_OK) return rc;

  /* k_bridge's flat records */
  std::vector<uint32_t> pair_off(p->num_fluxes + 1, 0);
  for (uint32_t f = 0; f < p->num_fluxes; f++) pair_off[f + 1] = pair_off[f] + p->fluxes[f].num_pairs;
  auto dom = [c](uint32_t v) { return (int)(c->h_ids[v] & 7); };
  std::vector<BChan> bchan;
  for (uint32_t f = 0; f < p->num_fluxes; f++) {
    if (!p->fluxes[f].discrete) continue;
    for (uint32_t ch = 0; ch < p->fluxes[f].num_pairs; ch++) {
      const uint32_t gp = pair_off[f] + ch, to = pairs_dev[2 * gp + 1];
      bchan.push_back(BChan{elem(g, 0, to), p->init[dom(to)][0], pslot[2 * gp + 1], f, ch, 0});
    }
  }
  std::vector<BRow> brow(row_voxel.size());
  std::vector<BEnt> bent(ent_pair.size());
  for (size_t r = 0; r < row_voxel.size(); r++) {
    const uint32_t v = row_voxel[r], sp = row_species[r];
    const uint32_t x = v % (uint32_t)g.nx, y = (v / (uint32_t)g.nx) % (uint32_t)g.ny, z = (uint32_t)(v / g.plane);
    const bool face = x == 0 || y == 0 || z == 0 || x + 1 == (uint32_t)g.nx || y + 1 == (uint32_t)g.ny ||
                      z + 1 == (uint32_t)g.nz;
    const bool skip = p->fixed[dom(v)][sp] || (sp == 0 && p->force_dirichlet && face);
    brow[r] = BRow{elem(g, sp, v), row_start[r], row_start[r + 1], skip ? 1u : 0u, 0};
    for (uint32_t e = row_start[r]; e < row_start[r + 1]; e++) {
      const uint32_t f = ent_flux[e], gp = pair_off[f] + ent_pair[e];
      const uint32_t from = pairs_dev[2 * gp], to = pairs_dev[2 * gp + 1];
      bent[e] = BEnt{elem(g, sp, from), elem(g, sp, to), p->init[dom(from)][sp], p->init[dom(to)][sp],
                     pslot[2 * gp], pslot[2 * gp + 1],
                     p->fluxes[f].discrete ? c->state_off[f] + ent_pair[e] : NO_ROW,
                     f | (ent_sign[e] > 0 ? 1u : 0u) << 8};
    }
  }
  c->nbchan = (uint32_t)bchan.size();
  auto upb = [c](void** dptr, const void* src, size_t bytes) -> int {
    HIP_TRY(hipMalloc(dptr, std::max<size_t>(bytes, 16)));
    if (bytes) HIP_TRY(hipMemcpyAsync(*dptr, src, bytes, hipMemcpyHostToDevice, c->stream));
    return CS_OK;
  };
  if ((rc = upb(&c->d_bchan, bchan.data(), sizeof(BChan) * bchan.size())) != CS_OK) return rc;
  if ((rc = upb(&c->d_brow, brow.data(), sizeof(BRow) * brow.size())) != CS_OK) return rc;
  if ((rc = upb(&c->d_bent, bent.data(), sizeof(BEnt) * bent.size())) != CS_OK) return rc;
  HIP_TRY(hipMalloc(&c->d_f2_l1, c->esize * std::max<size_t>(keys.size(), 1)));
  HIP_TRY(hipMalloc((void**)&c->d_f2_fsum, sizeof(double) * std::max<size_t>(row_voxel.size(), 1)));
  HIP_TRY(hipMalloc((void**)&c->d_f2_fany, std::max<size_t>(row_voxel.size(), 1)));
  /* the host vectors above end here */
  HIP_TRY(hipStreamSynchronize(c->stream));
  return CS_OK;
}

}  // namespace

/* ------------------------------------------------------------- C ABI */

extern "C" {

const char* cs_last_error(void) { return g_err.c_str(); }

int cs_device_info(int* device, char* name, size_t name_len)
{
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, dev));
  if (device) *device = dev;
  if (name && name_len) {
    snprintf(name, name_len, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
  }
  return CS_OK;
}

int cs_create(const cs_problem* p, cs_context** out)
{
  if (!out) return set_err(CS_EINVAL, "out is NULL");
  *out = nullptr;
  int rc = validate(p);
  if (rc != CS_OK) return rc;
  cs_context* c = new cs_context();
  c->prob = *p;
  c->prob.domain_ids = nullptr;
  for (auto& F : c->prob.fluxes) { F.pairs = nullptr; F.init_states = nullptr; }
  c->precision = p->precision;
  c->esize = p->precision == CS_PREC_F64 ? 8 : 4;
  c->ns = p->num_species;
  c->nd = p->num_domains;
  c->nf = p->num_fluxes;
  for (int a = 0; a < 3; a++) c->n[a] = p->n[a];
  for (uint32_t s = 0; s < c->ns; s++) c->substeps[s] = p->substeps[s];
  Geo& g = c->geo;
  g.nx = (int)p->n[0];
  g.ny = (int)p->n[1];
  g.nz = (int)p->n[2];
  g.plane = (uint64_t)g.nx * g.ny;
  g.nv = g.plane * g.nz;
  g.ns = p->num_species;
  auto bail = [&](int code) { free_ctx(c); return code; };
  if (hipGetDevice(&c->device) != hipSuccess) return bail(set_err(CS_EDEVICE, "no HIP device"));
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(set_err(CS_EDEVICE, "hipStreamCreate failed"));
  c->scratch = scratch_on_device(c->device);
  if (!c->scratch) return bail(CS_EDEVICE);

  /* domain ids in device order with the uniform-neighbourhood bit */
  std::vector<uint8_t> ids(g.nv);
  const uint32_t n0 = p->n[0], n1 = p->n[1], n2 = p->n[2];
  for (uint32_t x = 0; x < n0; x++)
    for (uint32_t y = 0; y < n1; y++)
      for (uint32_t z = 0; z < n2; z++) {
        const uint64_t r = ((uint64_t)x * n1 + y) * n2 + z;
        const uint8_t d = p->domain_ids[r];
        bool uni = true;
        if (x > 0) uni &= p->domain_ids[r - (uint64_t)n1 * n2] == d;
        if (x + 1 < n0) uni &= p->domain_ids[r + (uint64_t)n1 * n2] == d;
        if (y > 0) uni &= p->domain_ids[r - n2] == d;
        if (y + 1 < n1) uni &= p->domain_ids[r + n2] == d;
        if (z > 0) uni &= p->domain_ids[r - 1] == d;
        if (z + 1 < n2) uni &= p->domain_ids[r + 1] == d;
        ids[((uint64_t)z * n1 + y) * n0 + x] = (uint8_t)(d | (uni ? 0x80 : 0));
      }
  if (hipMalloc(&c->d_ids, g.nv) != hipSuccess) return bail(set_err(CS_ENOMEM, "device allocation of ids failed"));
  for (int b = 0; b < 2; b++)
    if (hipMalloc(&c->d_u[b], c->esize * c->ns * g.nv) != hipSuccess)
      return bail(set_err(CS_ENOMEM, "device allocation of %zu bytes failed", c->esize * c->ns * g.nv));
  if (hipMemcpyAsync(c->d_ids, ids.data(), g.nv, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return bail(set_err(CS_EDEVICE, "upload of ids failed"));
  c->h_ids.swap(ids);
  for (uint8_t id : c->h_ids)
    if ((id & 7) < MAXD) c->dcount[id & 7]++;
  rc = build_flux_rows(c, p);
  if (rc != CS_OK) return bail(rc);
  rc = c->esize == 8 ? upload_params<double>(c) : upload_params<float>(c);
  if (rc != CS_OK) return bail(rc);
  if (hipMalloc(&c->d_partial, sizeof(double) * STAT_BLOCKS * MAXS * MAXD * 3) != hipSuccess ||
      hipMalloc(&c->d_stats, sizeof(double) * MAXS * MAXD * 3) != hipSuccess)
    return bail(set_err(CS_ENOMEM, "device allocation of reduction buffers failed"));
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, c->device) != hipSuccess) return bail(set_err(CS_EDEVICE, "device query failed"));
  const int cus = std::max(prop.multiProcessorCount, 1);
  c->launch = default_launch(g, c->esize, cus);
  /* fused passes by default in FP32, where they measured faster; FP64
   * blocks hold half the rows (register file), so there one-step passes
   * win.  A fused pass holds one block per CU, so a grid whose fused launch
   * has fewer blocks than half the CUs leaves most of them idle: on the
   * 84^3 G1 geometry at 12 nm (36 blocks) a step takes 79.6 us fused
   * against 36.5 us in one-step passes (FP32, one box), while the
   * 200x200x100 dyad (240 blocks) takes 52.4 against 62.3 us. */
  const long fused_tiles = (long)((g.nx + BX - 3) / (BX - 2)) * ((g.ny + 16 - 3) / (16 - 2));
  const long fused_blocks = fused_tiles * ((g.nz + c->launch.fz - 1) / std::max(c->launch.fz, 1));
  c->launch.fuse = (c->f2_ok && c->esize == 4 && 2 * fused_blocks >= cus) ? 1 : 0;
  rc = c->esize == 8 ? fill_init<double>(c) : fill_init<float>(c);
  if (rc != CS_OK) return bail(rc);
  if (hipStreamSynchronize(c->stream) != hipSuccess) return bail(set_err(CS_EDEVICE, "setup failed"));
  *out = c;
  return CS_OK;
}

int cs_destroy(cs_context* c)
{
  if (!c) return CS_OK;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  free_ctx(c);
  return CS_OK;
}

int cs_reset(cs_context* c)
{
  if (!c) return set_err(CS_EINVAL, "context is NULL");
  HIP_TRY(hipSetDevice(c->device));
  int rc = c->esize == 8 ? fill_init<double>(c) : fill_init<float>(c);
  if (rc != CS_OK) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return CS_OK;
}

int cs_upload_species(cs_context* c, uint32_t s, const double* src)
{
  if (!c || !src) return set_err(CS_EINVAL, "NULL argument");
  if (s >= c->ns) return set_err(CS_ERANGE, "species %u out of range", s);
  HIP_TRY(hipSetDevice(c->device));
  const uint64_t nv = c->geo.nv;
  std::vector<char> buf(nv * c->esize);
  const uint32_t n0 = c->n[0], n1 = c->n[1], n2 = c->n[2];
  for (uint32_t x = 0; x < n0; x++)
    for (uint32_t y = 0; y < n1; y++)
      for (uint32_t z = 0; z < n2; z++) {
        const uint64_t dv = ((uint64_t)z * n1 + y) * n0 + x;
        /* the device holds the deviation from the voxel's resting value */
        const double val = src[((uint64_t)x * n1 + y) * n2 + z] - c->prob.init[c->h_ids[dv] & 7][s];
        if (c->esize == 8) ((double*)buf.data())[dv] = val;
        else ((float*)buf.data())[dv] = (float)val;
      }
  /* one plane of species s per z-group of the device stack */
  const size_t pb = c->geo.plane * c->esize;
  char* base = (char*)c->d_u[c->cur] + (uint64_t)s * pb;
  HIP_TRY(hipMemcpy2DAsync(base, pb * c->ns, buf.data(), pb, pb, (size_t)c->geo.nz, hipMemcpyHostToDevice,
                           c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return CS_OK;
}

int cs_download_species(cs_context* c, uint32_t s, double* dst)
{
  if (!c || !dst) return set_err(CS_EINVAL, "NULL argument");
  if (s >= c->ns) return set_err(CS_ERANGE, "species %u out of range", s);
  HIP_TRY(hipSetDevice(c->device));
  const uint64_t nv = c->geo.nv;
  std::vector<char> buf(nv * c->esize);
  const size_t pb = c->geo.plane * c->esize;
  const char* base = (const char*)c->d_u[c->cur] + (uint64_t)s * pb;
  HIP_TRY(hipMemcpy2DAsync(buf.data(), pb, base, pb * c->ns, pb, (size_t)c->geo.nz, hipMemcpyDeviceToHost,
                           c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  const uint32_t n0 = c->n[0], n1 = c->n[1], n2 = c->n[2];
  for (uint32_t x = 0; x < n0; x++)
    for (uint32_t y = 0; y < n1; y++)
      for (uint32_t z = 0; z < n2; z++) {
        const uint64_t dv = ((uint64_t)z * n1 + y) * n0 + x;
        dst[((uint64_t)x * n1 + y) * n2 + z] =
            (c->esize == 8 ? ((const double*)buf.data())[dv] : (double)((const float*)buf.data())[dv]) +
            c->prob.init[c->h_ids[dv] & 7][s];
      }
  return CS_OK;
}

int cs_uniforms_needed(const cs_context* c, uint64_t first, uint64_t n, uint64_t* count)
{
  if (!c || !count) return set_err(CS_EINVAL, "NULL argument");
  *count = c->prob.rng == CS_RNG_MT ? gates_in(c, first, n) * c->stoch_per_gate : 0;
  return CS_OK;
}

int cs_step(cs_context* c, uint64_t first, uint64_t n)
{
  if (!c) return set_err(CS_EINVAL, "context is NULL");
  if (first == 0) return set_err(CS_EINVAL, "time indices start at 1");
  return c->esize == 8 ? run_steps<double>(c, first, n, nullptr, 0, false, nullptr, nullptr)
                       : run_steps<float>(c, first, n, nullptr, 0, false, nullptr, nullptr);
}

int cs_step_uniforms(cs_context* c, uint64_t first, uint64_t n, const double* uni, uint64_t count)
{
  if (!c) return set_err(CS_EINVAL, "context is NULL");
  if (first == 0) return set_err(CS_EINVAL, "time indices start at 1");
  if (count && !uni) return set_err(CS_EINVAL, "uniforms is NULL");
  static const double dummy = 0.0;
  if (!uni) uni = &dummy;
  return c->esize == 8 ? run_steps<double>(c, first, n, uni, count, false, nullptr, nullptr)
                       : run_steps<float>(c, first, n, uni, count, false, nullptr, nullptr);
}

int cs_synchronize(cs_context* c)
{
  if (!c) return set_err(CS_EINVAL, "context is NULL");
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return CS_OK;
}

int cs_fused_available(const cs_context* c, char* why, size_t why_len)
{
  if (!c) return set_err(CS_EINVAL, "context is NULL");
  if (why && why_len) snprintf(why, why_len, "%s", c->f2_ok ? "" : c->f2_why.c_str());
  return c->f2_ok ? 1 : 0;
}

int cs_time_passes(cs_context* c, uint64_t first, uint64_t n, double* total_ms, double* pass_ms,
                   uint64_t* passes, uint64_t* fused)
{
  if (!c || !total_ms || !pass_ms) return set_err(CS_EINVAL, "NULL argument");
  if (first == 0) return set_err(CS_EINVAL, "time indices start at 1");
  if (c->prob.rng == CS_RNG_MT && c->stoch_per_gate)
    return set_err(CS_EINVAL, "cs_time_passes needs Philox gating");
  return c->esize == 8 ? run_steps<double>(c, first, n, nullptr, 0, true, total_ms, pass_ms, passes, fused)
                       : run_steps<float>(c, first, n, nullptr, 0, true, total_ms, pass_ms, passes, fused);
}

int cs_time_steps(cs_context* c, uint64_t first, uint64_t n, double* total_ms, double* kernel_ms)
{
  if (!c || !total_ms || !kernel_ms) return set_err(CS_EINVAL, "NULL argument");
  if (first == 0) return set_err(CS_EINVAL, "time indices start at 1");
  if (c->prob.rng == CS_RNG_MT && c->stoch_per_gate)
    return set_err(CS_EINVAL, "cs_time_steps needs Philox gating");
  return c->esize == 8 ? run_steps<double>(c, first, n, nullptr, 0, true, total_ms, kernel_ms)
                       : run_steps<float>(c, first, n, nullptr, 0, true, total_ms, kernel_ms);
}

int cs_reduce_stats(cs_context* c, double* mn, double* mx, double* tot)
{
  if (!c || !mn || !mx || !tot) return set_err(CS_EINVAL, "NULL argument");
  HIP_TRY(hipSetDevice(c->device));
  /* a 2-D grid of at most STAT_BLOCKS blocks: columns of a plane by planes */
  const uint32_t gx = (uint32_t)std::min<uint64_t>(16, (c->geo.plane + STAT_THREADS - 1) / STAT_THREADS);
  const uint32_t gy = (uint32_t)std::min<int>(STAT_BLOCKS / 16, c->geo.nz);
  const uint32_t blocks = gx * gy;
  const double dV = c->prob.dV;
  DomCounts cnt{};
  for (uint32_t d = 0; d < MAXD; d++) cnt.n[d] = c->dcount[d];
  if (c->esize == 8) {
    hipLaunchKernelGGL(k_stats<double>, dim3(gx, gy), dim3(STAT_THREADS), 0, c->stream,
                       (const double*)c->d_u[c->cur], c->d_ids, c->geo, c->ns, dV, c->d_partial);
    hipLaunchKernelGGL(k_stats_final<double>, dim3(c->ns * MAXD), dim3(STAT_THREADS), 0, c->stream, c->d_partial,
                       blocks, c->ns, (const DevParams<double>*)c->d_params, cnt, dV, c->d_stats);
  } else {
    hipLaunchKernelGGL(k_stats<float>, dim3(gx, gy), dim3(STAT_THREADS), 0, c->stream,
                       (const float*)c->d_u[c->cur], c->d_ids, c->geo, c->ns, dV, c->d_partial);
    hipLaunchKernelGGL(k_stats_final<float>, dim3(c->ns * MAXD), dim3(STAT_THREADS), 0, c->stream, c->d_partial,
                       blocks, c->ns, (const DevParams<float>*)c->d_params, cnt, dV, c->d_stats);
  }
  HIP_TRY(hipGetLastError());
  double host[MAXS * MAXD * 3];
  HIP_TRY(hipMemcpyAsync(host, c->d_stats, sizeof(double) * c->ns * MAXD * 3, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (uint32_t d = 0; d < c->nd; d++)
    for (uint32_t s = 0; s < c->ns; s++) {
      const double* h = host + (s * MAXD + d) * 3;
      mn[d * c->ns + s] = h[0];
      mx[d * c->ns + s] = h[1];
      tot[d * c->ns + s] = h[2];
    }
  return CS_OK;
}

int cs_linescan(cs_context* c, uint32_t species, uint32_t axis, const double* wx, const double* wy,
                const double* wz, uint32_t domain_mask, double* proj)
{
  if (!c || !proj) return set_err(CS_EINVAL, "NULL argument");
  if (species >= c->ns) return set_err(CS_ERANGE, "species %u out of range", species);
  if (axis > 2) return set_err(CS_EINVAL, "axis must be 0, 1 or 2");
  if ((axis != 0 && !wx) || (axis != 1 && !wy) || (axis != 2 && !wz))
    return set_err(CS_EINVAL, "the weights of both transverse axes are required");
  HIP_TRY(hipSetDevice(c->device));
  const Geo& g = c->geo;
  const int ychunk = 64;
  const int gyc = (g.ny + ychunk - 1) / ychunk;
  const size_t nrows = (size_t)g.ny * g.nz, ncols = (size_t)g.nz * gyc * g.nx;
  if (!c->d_lweights) HIP_TRY(hipMalloc(&c->d_lweights, sizeof(double) * ((size_t)g.nx + g.ny + g.nz)));
  if (!c->d_lpart) HIP_TRY(hipMalloc(&c->d_lpart, sizeof(double) * std::max(nrows, ncols)));
  double* dwx = c->d_lweights;
  double* dwy = dwx + g.nx;
  const void* us = c->d_u[c->cur];
  std::vector<double> part;
  if (axis == 0) {
    HIP_TRY(hipMemcpyAsync(dwy, wy, sizeof(double) * g.ny, hipMemcpyHostToDevice, c->stream));
    if (c->esize == 8)
      hipLaunchKernelGGL(k_linescan_cols<double>, dim3(g.nz, gyc), dim3(256), 0, c->stream, (const double*)us,
                         species, c->d_ids, (const DevParams<double>*)c->d_params, g, domain_mask, dwy, ychunk,
                         c->d_lpart);
    else
      hipLaunchKernelGGL(k_linescan_cols<float>, dim3(g.nz, gyc), dim3(256), 0, c->stream, (const float*)us,
                         species, c->d_ids, (const DevParams<float>*)c->d_params, g, domain_mask, dwy, ychunk,
                         c->d_lpart);
    HIP_TRY(hipGetLastError());
    part.resize(ncols);
    HIP_TRY(hipMemcpyAsync(part.data(), c->d_lpart, sizeof(double) * ncols, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (int x = 0; x < g.nx; x++) {
      double acc = 0.0;
      for (int z = 0; z < g.nz; z++) {
        double col = 0.0;
        for (int k = 0; k < gyc; k++) col += part[((size_t)z * gyc + k) * g.nx + x];
        acc += wz[z] * col;
      }
      proj[x] = acc;
    }
    return CS_OK;
  }
  HIP_TRY(hipMemcpyAsync(dwx, wx, sizeof(double) * g.nx, hipMemcpyHostToDevice, c->stream));
  const unsigned blocks = (unsigned)((nrows + 3) / 4);
  if (c->esize == 8)
    hipLaunchKernelGGL(k_linescan_rows<double>, dim3(blocks), dim3(256), 0, c->stream, (const double*)us, species, c->d_ids,
                       (const DevParams<double>*)c->d_params, g, domain_mask, dwx, c->d_lpart);
  else
    hipLaunchKernelGGL(k_linescan_rows<float>, dim3(blocks), dim3(256), 0, c->stream, (const float*)us, species, c->d_ids,
                       (const DevParams<float>*)c->d_params, g, domain_mask, dwx, c->d_lpart);
  HIP_TRY(hipGetLastError());
  part.resize(nrows);
  HIP_TRY(hipMemcpyAsync(part.data(), c->d_lpart, sizeof(double) * nrows, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (axis == 1) {
    for (int y = 0; y < g.ny; y++) {
      double acc = 0.0;
      for (int z = 0; z < g.nz; z++) acc += wz[z] * part[(size_t)z * g.ny + y];
      proj[y] = acc;
    }
  } else {
    for (int z = 0; z < g.nz; z++) {
      double acc = 0.0;
      for (int y = 0; y < g.ny; y++) acc += wy[y] * part[(size_t)z * g.ny + y];
      proj[z] = acc;
    }
  }
  return CS_OK;
}

int cs_linescan_noalloc(cs_context* c, uint32_t species, uint32_t axis, const double* wx, const double* wy,
                        const double* wz, uint32_t domain_mask, double* proj)
{
  if (!c || !proj) return set_err(CS_EINVAL, "NULL argument");
  if (species >= c->ns) return set_err(CS_ERANGE, "species %u out of range", species);
  if (axis > 2) return set_err(CS_EINVAL, "axis must be 0, 1 or 2");
  if ((axis != 0 && !wx) || (axis != 1 && !wy) || (axis != 2 && !wz))
    return set_err(CS_EINVAL, "the weights of both transverse axes are required");
  if (!c->scratch) return set_err(CS_EDEVICE, "no scratch on this device");
  const Geo& g = c->geo;
  const int ychunk = 64;
  const int gyc = (g.ny + ychunk - 1) / ychunk;
  /* scratch segments, each starting on 256 bytes: weights | partial sums |
   * projection */
  auto up32 = [](uint64_t n) { return (n + 31) / 32 * 32; };
  const uint64_t nw = (uint64_t)g.nx + g.ny + g.nz;
  const uint64_t np = axis == 0 ? (uint64_t)g.nz * gyc * g.nx : (uint64_t)g.ny * g.nz;
  const uint64_t nk = axis == 0 ? g.nx : axis == 1 ? g.ny : g.nz;
  const uint64_t need = up32(nw) + up32(np) + up32(nk);
  if (need > CS_LINESCAN_NOALLOC_MAX_DOUBLES)
    return set_err(CS_ERANGE, "linescan needs %llu scratch doubles, beyond the %u this entry point holds",
                   (unsigned long long)need, (unsigned)CS_LINESCAN_NOALLOC_MAX_DOUBLES);
  HIP_TRY(hipSetDevice(c->device));
  double* dw = c->scratch;
  double* dpart = dw + up32(nw);
  double* dproj = dpart + up32(np);
  /* weights the kernels read; the scan axis's own weights are not read */
  if (wx && axis != 0) HIP_TRY(hipMemcpyAsync(dw, wx, sizeof(double) * g.nx, hipMemcpyHostToDevice, c->stream));
  if (wy && axis != 1)
    HIP_TRY(hipMemcpyAsync(dw + g.nx, wy, sizeof(double) * g.ny, hipMemcpyHostToDevice, c->stream));
  if (wz && axis != 2)
    HIP_TRY(hipMemcpyAsync(dw + g.nx + g.ny, wz, sizeof(double) * g.nz, hipMemcpyHostToDevice, c->stream));
  const void* us = c->d_u[c->cur];
  if (axis == 0) {
    if (c->esize == 8)
      hipLaunchKernelGGL(k_linescan_cols<double>, dim3(g.nz, gyc), dim3(256), 0, c->stream, (const double*)us,
                         species, c->d_ids, (const DevParams<double>*)c->d_params, g, domain_mask, dw + g.nx,
                         ychunk, dpart);
    else
      hipLaunchKernelGGL(k_linescan_cols<float>, dim3(g.nz, gyc), dim3(256), 0, c->stream, (const float*)us,
                         species, c->d_ids, (const DevParams<float>*)c->d_params, g, domain_mask, dw + g.nx,
                         ychunk, dpart);
  } else {
    const unsigned blocks = (unsigned)(((uint64_t)g.ny * g.nz + 3) / 4);
    if (c->esize == 8)
      hipLaunchKernelGGL(k_linescan_rows<double>, dim3(blocks), dim3(256), 0, c->stream, (const double*)us, species,
                         c->d_ids, (const DevParams<double>*)c->d_params, g, domain_mask, dw, dpart);
    else
      hipLaunchKernelGGL(k_linescan_rows<float>, dim3(blocks), dim3(256), 0, c->stream, (const float*)us, species,
                         c->d_ids, (const DevParams<float>*)c->d_params, g, domain_mask, dw, dpart);
  }
  hipLaunchKernelGGL(k_linescan_proj, dim3((unsigned)((nk + 255) / 256)), dim3(256), 0, c->stream, dpart, dw, g,
                     gyc, axis, dproj);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(proj, dproj, sizeof(double) * nk, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return CS_OK;
}

int cs_get_channel_states(cs_context* c, uint32_t f, uint8_t* states)
{
  if (!c || !states) return set_err(CS_EINVAL, "NULL argument");
  if (f >= c->nf || !c->prob.fluxes[f].discrete) return set_err(CS_ERANGE, "flux %u is not a discrete flux", f);
  HIP_TRY(hipSetDevice(c->device));
  if (c->npairs[f])
    HIP_TRY(hipMemcpyAsync(states, c->d_states + c->state_off[f], c->npairs[f], hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return CS_OK;
}

int cs_set_channel_states(cs_context* c, uint32_t f, const uint8_t* states)
{
  if (!c || !states) return set_err(CS_EINVAL, "NULL argument");
  if (f >= c->nf || !c->prob.fluxes[f].discrete) return set_err(CS_ERANGE, "flux %u is not a discrete flux", f);
  HIP_TRY(hipSetDevice(c->device));
  std::vector<uint8_t> tmp(states, states + c->npairs[f]), rec(c->npairs[f]);
  for (auto& v : tmp) v = v ? 1 : 0;
  if (c->npairs[f]) {
    HIP_TRY(hipMemcpyAsync(c->d_states + c->state_off[f], tmp.data(), c->npairs[f], hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(rec.data(), c->d_recruited + c->state_off[f], c->npairs[f], hipMemcpyDeviceToHost,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (uint32_t i = 0; i < c->npairs[f]; i++) rec[i] |= tmp[i];
    HIP_TRY(hipMemcpyAsync(c->d_recruited + c->state_off[f], rec.data(), c->npairs[f], hipMemcpyHostToDevice,
                           c->stream));
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  return CS_OK;
}

int cs_get_channel_recruited(cs_context* c, uint32_t f, uint8_t* out)
{
  if (!c || !out) return set_err(CS_EINVAL, "NULL argument");
  if (f >= c->nf || !c->prob.fluxes[f].discrete) return set_err(CS_ERANGE, "flux %u is not a discrete flux", f);
  HIP_TRY(hipSetDevice(c->device));
  if (c->npairs[f])
    HIP_TRY(hipMemcpyAsync(out, c->d_recruited + c->state_off[f], c->npairs[f], hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return CS_OK;
}

int cs_get_launch_config(const cs_context* c, cs_launch_config* out)
{
  if (!c || !out) return set_err(CS_EINVAL, "NULL argument");
  out->rows = (uint32_t)c->launch.rows;
  out->zchunk = (uint32_t)c->launch.zchunk;
  out->xcd_remap = (uint32_t)c->launch.xcd;
  out->prefetch = (uint32_t)c->launch.pf;
  out->nontemporal = (uint32_t)c->launch.nt;
  out->lds_tile = (uint32_t)c->launch.lds;
  out->linear = (uint32_t)c->launch.lin;
  out->yblock = (uint32_t)c->launch.yb;
  out->fuse = (uint32_t)c->launch.fuse;
  out->fuse_zchunk = (uint32_t)c->launch.fz;
  out->fuse_xcd = (uint32_t)c->launch.fx;
  out->fuse_fixed = (uint32_t)c->launch.ff;
  return CS_OK;
}

int cs_set_launch_config(cs_context* c, const cs_launch_config* in)
{
  if (!c || !in) return set_err(CS_EINVAL, "NULL argument");
  if (in->yblock) {
    if (in->yblock != 2 && in->yblock != 4) return set_err(CS_EINVAL, "yblock must be 0, 2 or 4");
    if (in->rows < 1 || in->rows > 8) return set_err(CS_EINVAL, "the row-blocked kernel takes 1 to 8 wavefronts per block");
    if (in->lds_tile || in->linear) return set_err(CS_EINVAL, "yblock excludes lds_tile and linear");
    if (c->geo.plane * c->esize >= (1ull << 31))
      return set_err(CS_EINVAL, "plane of %llu voxels too large for the row-blocked kernel",
                     (unsigned long long)c->geo.plane);
  } else if (!(in->rows == 1 || in->rows == 2 || in->rows == 4 || in->rows == 8)) {
    return set_err(CS_EINVAL, "rows must be 1, 2, 4 or 8");
  }
  if (in->zchunk < 1) return set_err(CS_EINVAL, "zchunk must be >= 1");
  if (in->prefetch != 1 && in->prefetch != 2) return set_err(CS_EINVAL, "prefetch must be 1 or 2");
  if (in->nontemporal > 1 || in->lds_tile > 1 || in->linear > 1)
    return set_err(CS_EINVAL, "flags must be 0 or 1");
  if (in->xcd_remap > 64) return set_err(CS_EINVAL, "xcd_remap is 0, 1 or a group size up to 64");
  if (in->lds_tile && in->linear) return set_err(CS_EINVAL, "the LDS tile kernel uses 2-D tiles");
  if (in->lds_tile && in->prefetch != 1) return set_err(CS_EINVAL, "the LDS tile kernel prefetches one plane");
  if (in->lds_tile && lds_tile_bytes(c->esize, c->ns, (int)in->rows) > LDS_TILE_LIMIT)
    return set_err(CS_EINVAL, "LDS tile of %u rows too large for %u species", in->rows, c->ns);
  LaunchCfg L;
  L.rows = (int)in->rows;
  if (!in->lds_tile && (int)std::min<uint32_t>(in->zchunk, (uint32_t)c->geo.nz) > max_zchunk(c->geo, c->esize))
    return set_err(CS_EINVAL, "zchunk %u exceeds %d for this grid", in->zchunk, max_zchunk(c->geo, c->esize));
  L.zchunk = (int)std::min<uint32_t>(in->zchunk, (uint32_t)c->geo.nz);
  L.xcd = (int)in->xcd_remap;
  L.pf = (int)in->prefetch;
  L.nt = (int)in->nontemporal;
  L.lds = (int)in->lds_tile;
  L.lin = (int)in->linear;
  L.yb = (int)in->yblock;
  if (in->fuse > 1) return set_err(CS_EINVAL, "fuse must be 0 or 1");
  if (in->fuse && !c->f2_ok) return set_err(CS_EINVAL, "fused passes unavailable: %s", c->f2_why.c_str());
  if (in->fuse_xcd > 64) return set_err(CS_EINVAL, "fuse_xcd is 0, 1 or a group size up to 64");
  if (in->fuse_fixed > 1) return set_err(CS_EINVAL, "fuse_fixed must be 0 or 1");
  const int fzmax = max_fused_zchunk(c->geo, c->esize);
  if (in->fuse_zchunk < 1) return set_err(CS_EINVAL, "fuse_zchunk must be >= 1");
  if (fzmax >= 1 && (int)std::min<uint32_t>(in->fuse_zchunk, (uint32_t)c->geo.nz) > fzmax)
    return set_err(CS_EINVAL, "fuse_zchunk %u exceeds %d for this grid", in->fuse_zchunk, fzmax);
  L.fuse = (int)in->fuse;
  L.fz = (int)std::min<uint32_t>(in->fuse_zchunk, (uint32_t)c->geo.nz);
  L.fx = (int)in->fuse_xcd;
  L.ff = (int)in->fuse_fixed;
  c->launch = L;
  return CS_OK;
}

int cs_autotune(cs_context* c, uint64_t first_time_ind, uint32_t reps, cs_launch_config* best,
                double* best_ms)
{
  if (!c) return set_err(CS_EINVAL, "context is NULL");
  if (first_time_ind == 0) return set_err(CS_EINVAL, "time indices start at 1");
  if (reps < 1) reps = 1;
  HIP_TRY(hipSetDevice(c->device));
  /* the substep pattern of the coming steps: which species diffuse and
   * whether the reactions fire differs from step to step, and so does the
   * best shape.  Each distinct (diffusing species, reactions) pair of one
   * period is timed once per candidate and weighted by how often it occurs */
  uint64_t period = 1;
  auto lcm = [](uint64_t a, uint64_t b) { uint64_t x = a, y = b; while (y) { uint64_t t = x % y; x = y; y = t; } return a / x * b; };
  for (uint32_t s = 0; s < c->ns; s++)
    if (c->substeps[s]) period = std::min<uint64_t>(lcm(period, c->substeps[s]), 100000);
  if (c->prob.num_buffers) period = std::min<uint64_t>(lcm(period, c->prob.reaction_substep), 100000);
  struct Pattern { uint32_t mask, react; uint64_t count; };
  std::vector<Pattern> pats;
  for (uint64_t k = 0; k < period; k++) {
    const uint64_t ti = first_time_ind + k;
    uint32_t dm = 0;
    for (uint32_t s = 0; s < c->ns; s++)
      if (c->substeps[s] && ti % c->substeps[s] == 0) dm |= 1u << s;
    const uint32_t re = react_mode(c, ti, dm);
    bool found = false;
    for (auto& q : pats)
      if (q.mask == dm && q.react == re) { q.count++; found = true; break; }
    if (!found) pats.push_back({dm, re, 1});
  }
  const int nz = c->geo.nz;
  std::vector<int> zcs;
  const int zmax = max_zchunk(c->geo, c->esize);
  for (int zc : {8, 12, 16, 24, 32, 50, 64})
    if (zc < nz && zc <= zmax) zcs.push_back(zc);
  if (nz <= zmax) zcs.push_back(nz);
  if (zcs.empty()) zcs.push_back(zmax);
  std::vector<LaunchCfg> cands;
  for (int zc : zcs)
    for (int xcd : {0, 1, 4, 16})
      for (int nt : {0, 1}) {
        for (int pf : {1, 2})
          for (int rows : {2, 4, 8}) {
            cands.push_back({rows, zc, xcd, pf, nt, 0, 0, 0});   /* 2-D tiles      */
            cands.push_back({rows, zc, xcd, pf, nt, 0, 1, 0});   /* linear columns */
          }
        for (int rows : {2, 4, 8})
          if (lds_tile_bytes(c->esize, c->ns, rows) <= LDS_TILE_LIMIT)
            cands.push_back({rows, zc, xcd, 1, nt, 1, 0, 0});    /* LDS tiles      */
        if (c->geo.plane * c->esize < (1ull << 31)) {
          const int full = std::min(8, (c->geo.nx + BX - 1) / BX);
          std::vector<int> wv{full};
          if (full > 2) wv.push_back((full + 1) / 2);
          for (int yb : {2, 4})
            for (int pf : {1, 2})
              for (int w : wv) cands.push_back({w, zc, xcd, pf, nt, 0, 0, yb});   /* row-blocked */
        }
      }
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  LaunchCfg winner = c->launch;
  double wms = 1e30;
  int rc = CS_OK;
  for (const LaunchCfg& L : cands) {
    /* reads the current state, writes the other buffer (scratch until the
     * next step overwrites it); the state is unchanged */
    auto go = [&](uint32_t dm, uint32_t re) {
      if (c->esize == 8) launch_stencil_cfg<double>(c, L, c->d_u[c->cur], c->d_u[c->cur ^ 1], dm, re);
      else launch_stencil_cfg<float>(c, L, c->d_u[c->cur], c->d_u[c->cur ^ 1], dm, re);
    };
    double mean = 0;
    for (const Pattern& q : pats) {
      go(q.mask, q.react);
      if (hipEventRecord(e0, c->stream) != hipSuccess) { rc = set_err(CS_EDEVICE, "event record failed"); goto done; }
      for (uint32_t r = 0; r < reps; r++) go(q.mask, q.react);
      if (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
        rc = set_err(CS_EDEVICE, "autotune launch failed: %s", hipGetErrorString(hipGetLastError()));
        goto done;
      }
      float ms = 0;
      hipEventElapsedTime(&ms, e0, e1);
      mean += (double)ms / reps * (double)q.count / (double)period;
    }
    if (mean < wms) { wms = mean; winner = L; }
  }
  winner.fuse = 0;
  winner.fz = c->launch.fz;
  winner.fx = c->launch.fx;
  winner.ff = c->launch.ff;
  /* fused two-step passes: the pairs (ti, ti+1) of two periods that may
   * fuse, grouped by their patterns, without gating (so the channel
   * states and recruitment flags stay as they are); a pass counts as two
   * steps against the best single-step shape */
  if (c->f2_ok) {
    struct Pair { F2Step a, b; uint64_t count; };
    std::vector<Pair> prs;
    uint64_t fusable = 0;
    for (uint64_t k = 0; k < 2 * period; k += 2) {
      const uint64_t ti = first_time_ind + k;
      LaunchCfg on = winner;
      on.fuse = 1;
      if (!can_fuse(c, on, ti)) continue;
      F2Step a = step_info(c, ti, 0), b = step_info(c, ti + 1, 0);
      a.gate = b.gate = 0;
      fusable++;
      bool found = false;
      for (auto& q : prs)
        if (q.a.dmask == a.dmask && q.a.react == a.react && q.a.fluxmask == a.fluxmask && q.b.dmask == b.dmask &&
            q.b.react == b.react && q.b.fluxmask == b.fluxmask) { q.count++; found = true; break; }
      if (!found) prs.push_back({a, b, 1});
    }
    const int fzmax = max_fused_zchunk(c->geo, c->esize);
    std::vector<int> fzs;
    for (int zc : {16, 25, 32, 50, 64, 100})
      if (zc < nz && zc <= fzmax) fzs.push_back(zc);
    if (nz <= fzmax) fzs.push_back(nz);
    if (std::find(fzs.begin(), fzs.end(), c->launch.fz) == fzs.end() && c->launch.fz <= fzmax)
      fzs.push_back(c->launch.fz);   /* the default's balanced chunk */
    std::vector<int> fxs{0, 1, 8};
    if (std::find(fxs.begin(), fxs.end(), c->launch.fx) == fxs.end())
      fxs.push_back(c->launch.fx);   /* the default's tile groups */
    double fms = 1e30;
    LaunchCfg fbest = winner;
    for (int zc : fzs)
      for (int fx : fxs) {
        LaunchCfg L = winner;
        L.fuse = 1;
        L.fz = zc;
        L.fx = fx;
        double mean = 0;
        for (const Pair& q : prs) {
          /* the pass and its bridge (tail and the next head, no gating) */
          auto go = [&]() {
            if (c->esize == 8) {
              launch_step2<double>(c, L, c->d_u[c->cur], c->d_u[c->cur ^ 1], q.a, q.b);
              launch_bridge<double>(c, c->d_u[c->cur ^ 1], q.b, true, q.a, true, nullptr);
            } else {
              launch_step2<float>(c, L, c->d_u[c->cur], c->d_u[c->cur ^ 1], q.a, q.b);
              launch_bridge<float>(c, c->d_u[c->cur ^ 1], q.b, true, q.a, true, nullptr);
            }
          };
          go();
          if (hipEventRecord(e0, c->stream) != hipSuccess) { rc = set_err(CS_EDEVICE, "event record failed"); goto done; }
          for (uint32_t r = 0; r < reps; r++) go();
          if (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
            rc = set_err(CS_EDEVICE, "autotune launch failed: %s", hipGetErrorString(hipGetLastError()));
            goto done;
          }
          float ms = 0;
          hipEventElapsedTime(&ms, e0, e1);
          mean += (double)ms / reps * (double)q.count / (double)fusable;
        }
        if (!prs.empty() && mean < fms) { fms = mean; fbest = L; }
      }
    if (!prs.empty() && fms < 2.0 * wms) {
      winner = fbest;
      wms = fms / 2.0;
    }
  }
done:
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (rc != CS_OK) return rc;
  HIP_TRY(hipGetLastError());
  c->launch = winner;
  if (best) cs_get_launch_config(c, best);
  if (best_ms) *best_ms = wms;
  return CS_OK;
}

int cs_calibration_copy(uint64_t bytes, uint32_t reps, double* ms_per_copy)
{
  if (bytes < 16 || reps < 1) return set_err(CS_EINVAL, "bytes >= 16 and reps >= 1 required");
  const uint64_t n16 = bytes / 16;
  void *a = nullptr, *b = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = CS_OK;
  float ms = 0;
  if (hipMalloc(&a, n16 * 16) != hipSuccess || hipMalloc(&b, n16 * 16) != hipSuccess) {
    rc = set_err(CS_ENOMEM, "calibration buffers");
    goto out;
  }
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess) {
    rc = set_err(CS_EDEVICE, "calibration stream");
    goto out;
  }
  hipLaunchKernelGGL(k_calibration_copy, dim3(4096), dim3(256), 0, st, (const uint4*)a, (uint4*)b, n16);
  hipEventRecord(e0, st);
  for (uint32_t r = 0; r < reps; r++)
    hipLaunchKernelGGL(k_calibration_copy, dim3(4096), dim3(256), 0, st, (const uint4*)a, (uint4*)b, n16);
  hipEventRecord(e1, st);
  if (hipEventSynchronize(e1) != hipSuccess) { rc = set_err(CS_EDEVICE, "calibration copy failed"); goto out; }
  hipEventElapsedTime(&ms, e0, e1);
  if (ms_per_copy) *ms_per_copy = ms / reps;
out:
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  if (st) hipStreamDestroy(st);
  if (a) hipFree(a);
  if (b) hipFree(b);
  return rc;
}

double cs_ryr_flux(double dt, double h, double u0, double u1, const double* params)
{
  return cs_ryr_flux_d(dt, h, u0, u1, params[6]);
}

double cs_serca_flux(double dt, double h, double u0, double u1, const double* params)
{
  return cs_serca_flux_d(dt, h, u0, u1, params[1]);
}

void cs_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  cs_philox(c, key[0], key[1]);
  for (int i = 0; i < 4; i++) out[i] = c[i];
}

}  // extern "C"
