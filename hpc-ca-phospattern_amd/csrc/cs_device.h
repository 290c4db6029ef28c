/*
 * cs_device.h -- device building blocks shared by the single-domain
 * kernels (csparks.hip) and the ensemble kernels (cs_batch.hip): the
 * parameter tables, the per-voxel update (7-point stencil, buffer
 * reactions, Dirichlet faces, fixed species), the column march of a
 * one-step pass, channel gating and the boundary-flux formulas, plus the
 * host helpers that derive the device tables from a cs_problem.
 *
 * Everything here has internal linkage; each translation unit gets its
 * own copy.  Errors are recorded through cs_internal_set_err, defined once
 * in csparks.hip, so cs_last_error() reports them from either unit.
 */
#ifndef CS_DEVICE_H
#define CS_DEVICE_H

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>

#include "csparks.h"
#include "cs_physics.h"

/* records the message for cs_last_error() and returns code */
int cs_internal_set_err(int code, const char* fmt, ...);

#define HIP_TRY(expr)                                                        \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess)                                                    \
      return cs_internal_set_err(CS_EDEVICE, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

namespace {

constexpr int MAXS = CS_MAX_SPECIES;
constexpr int MAXD = CS_MAX_DOMAINS;
constexpr int MAXF = CS_MAX_FLUXES;
constexpr int BX = 64;          /* one wavefront along x                  */
constexpr int GATE_THREADS = 256;
constexpr int FLUX_THREADS = 256;
constexpr int BRIDGE_THREADS = 1024;
constexpr size_t F2_ONE_BLOCK = 16384;  /* rows and channels k_bridge runs in one block; more
                                          run each phase as a launch of many blocks */
constexpr int STAT_THREADS = 256;
constexpr int STAT_BLOCKS = 4096;

#define set_err cs_internal_set_err

struct FluxParams {
  uint32_t kind, species, discrete, inner;
  uint32_t npairs, pair_off, state_off, stochastic;
  uint32_t group, uniform_off;   /* gating group, offset among draws     */
  double   half_1mK0;            /* (1-K_0)/2 for RyR                    */
  double   density;              /* SERCA pumps per nm^2                 */
  double   p[CS_NUM_RYR_PARAMS]; /* reference parameter vector           */
};

/* The buffer that binds a species s in one domain d (zero when none
 * does), about the resting state and scaled by the reaction step dt_r:
 * dkon = dt_r k_on, dkoff = dt_r k_off, dka = dt_r k_on (tot - B0) with B0
 * the resting value of s, dj0 = dt_r times the rate at rest (zero for
 * equilibrium initial values).  Padded to 8 values. */
template <typename T>
struct RateTab {
  T dkon, dkoff, dka, dj0, pad[4];
};

template <typename T>
struct DevParams {
  T coef[MAXS][MAXD * MAXD];
  /* Fields hold deviations from the resting state init[domain][species]:
   * dinit[s][d*MAXD+dn] = init[dn][s] - init[d][s] enters the stencil
   * across a domain interface; the buffer tables are taken about rest
   * (slot b, domain d): a = tot - B0, f0 = resting free species, j0 = rate
   * at rest */
  T kon[MAXS][MAXD], koff[MAXS][MAXD], ra[MAXS][MAXD], rf0[MAXS][MAXD], rj0[MAXS][MAXD];
  T dinit[MAXS][MAXD * MAXD];
  double init64[MAXD][MAXS];
  T dt_react;
  uint8_t bsp0[MAXS][MAXD], bsp1[MAXS][MAXD];
  uint32_t nb;
  uint32_t fixed_mask[MAXD];
  uint32_t dirichlet;
  uint32_t nflux, ndiscrete;
  /* species-major form of the buffer reactions: rt[d][s] holds the rates of
   * the buffer binding species s in domain d; rbound marks the species some
   * buffer binds; rfast is set when every buffer of every domain binds
   * species 0 and no species is bound by two buffers of one domain */
  RateTab<T> rt[MAXD][MAXS];
  T rc0[MAXD];   /* resting value of species 0 per domain */
  uint32_t rbound, rfast;
  double dt, h;
  FluxParams flux[MAXF];
};

/* Grid.  Fields are stored [z][species][y][x]: the ns species planes of
 * one z lie next to each other, so a pass over the fields is one read and
 * one write stream instead of one per species. */
struct Geo {
  int nx, ny, nz;
  uint32_t ns;
  uint64_t plane, nv;
};

/* element index of species s at voxel v = z*plane + (y*nx + x) */
__host__ __device__ inline uint64_t elem(const Geo& g, uint32_t s, uint64_t v)
{
  const uint64_t z = v / g.plane;
  return (z * g.ns + s) * g.plane + (v - z * g.plane);
}

/* Cross-lane moves within one 64-wide wavefront (DPP wave_shr:1 and
 * wave_shl:1).  from_left: lane i receives lane i-1's value and lane 0
 * keeps `edge`; from_right: lane i receives lane i+1's value and lane 63
 * keeps `edge`.  All 64 lanes must be active. */
__device__ inline float dpp_from_left(float v, float edge)
{
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(edge), __float_as_int(v), 0x138, 0xF, 0xF, false));
}
__device__ inline float dpp_from_right(float v, float edge)
{
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(edge), __float_as_int(v), 0x130, 0xF, 0xF, false));
}
__device__ inline double dpp_from_left(double v, double edge)
{
  const uint64_t a = (uint64_t)__double_as_longlong(v), e = (uint64_t)__double_as_longlong(edge);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp((int)(uint32_t)e, (int)(uint32_t)a, 0x138, 0xF, 0xF, false);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp((int)(uint32_t)(e >> 32), (int)(uint32_t)(a >> 32), 0x138, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
__device__ inline double dpp_from_right(double v, double edge)
{
  const uint64_t a = (uint64_t)__double_as_longlong(v), e = (uint64_t)__double_as_longlong(edge);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp((int)(uint32_t)e, (int)(uint32_t)a, 0x130, 0xF, 0xF, false);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp((int)(uint32_t)(e >> 32), (int)(uint32_t)(a >> 32), 0x130, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

/* The same moves with 0 for the lane that has no source (bound_ctrl). */
__device__ inline float dpp_from_left_zero(float v)
{
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xF, 0xF, true));
}
__device__ inline float dpp_from_right_zero(float v)
{
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xF, 0xF, true));
}
__device__ inline double dpp_from_left_zero(double v) { return dpp_from_left(v, 0.0); }
__device__ inline double dpp_from_right_zero(double v) { return dpp_from_right(v, 0.0); }

/* Tile handled by hardware block b.  Blocks are dispatched to the 8 XCDs
 * round-robin (b % 8).  xcd 0 keeps that order; xcd 1 gives XCD k the
 * contiguous tiles [k*per, (k+1)*per); xcd = G >= 2 hands out groups of G
 * consecutive tiles round-robin, so each group shares one XCD's L2 while
 * the groups spread over all XCDs.  The launch covers a multiple of 8
 * (xcd 1) or 8*G (xcd G) blocks, so the map is a permutation; tiles past
 * nblocks exit. */
__device__ inline int xcd_tile(int b, int nblocks, int xcd)
{
  if (xcd == 0) return b;
  if (xcd == 1) {
    const int per = (nblocks + 7) / 8;
    return (b % 8) * per + b / 8;
  }
  const int k = b % 8, j = b / 8;
  return ((j / xcd) * 8 + k) * xcd + j % xcd;
}

/* Launch size for nblocks tiles under xcd_tile's map. */
inline int xcd_launch(int nblocks, int xcd)
{
  if (xcd == 0) return nblocks;
  const int m = xcd == 1 ? 8 : 8 * xcd;
  return (nblocks + m - 1) / m * m;
}

/* Per-block copies of the tables every voxel update reads. */
template <typename T, int NS>
struct BlockTables {
  T coef[NS][MAXD * MAXD];
  T dinit[NS][MAXD * MAXD];
  T kon[MAXS][MAXD], koff[MAXS][MAXD], ra[MAXS][MAXD], rf0[MAXS][MAXD], rj0[MAXS][MAXD];
  RateTab<T> rt[MAXD][NS];
  T rc0[MAXD];
  uint8_t b0[MAXS][MAXD], b1[MAXS][MAXD];
  uint32_t fixed[MAXD];

  __device__ void load(const DevParams<T>* __restrict__ P, int tid, int nthr)
  {
    for (int i = tid; i < NS * MAXD * MAXD; i += nthr) {
      coef[i / (MAXD * MAXD)][i % (MAXD * MAXD)] = P->coef[i / (MAXD * MAXD)][i % (MAXD * MAXD)];
      dinit[i / (MAXD * MAXD)][i % (MAXD * MAXD)] = P->dinit[i / (MAXD * MAXD)][i % (MAXD * MAXD)];
    }
    for (int i = tid; i < MAXS * MAXD; i += nthr) {
      const int b = i / MAXD, d = i % MAXD;
      kon[b][d] = P->kon[b][d];
      koff[b][d] = P->koff[b][d];
      ra[b][d] = P->ra[b][d];
      rf0[b][d] = P->rf0[b][d];
      rj0[b][d] = P->rj0[b][d];
      b0[b][d] = P->bsp0[b][d];
      b1[b][d] = P->bsp1[b][d];
    }
    for (int i = tid; i < MAXD * NS; i += nthr) rt[i / NS][i % NS] = P->rt[i / NS][i % NS];
    if (tid < MAXD) {
      fixed[tid] = P->fixed_mask[tid];
      rc0[tid] = P->rc0[tid];
    }
  }
};

/* Buffer reactions in the reference's order (reaction_SIMD,
 * kernel_generic.cpp:397-509, called per buffer from species.c:1689-1692):
 * for every buffer slot b of the voxel's domain, J = k_on (tot - B) F -
 * k_off B, then B += dt J and F -= dt J, where a species that did not
 * diffuse this step is overwritten from u instead of accumulated.  Species
 * indices vary per slot and domain, so they are selected per lane. */
template <typename T, int NS>
__device__ inline void react_general(T nw[NS], const T uc[NS], const BlockTables<T, NS>& S, int d, int nb,
                                     uint32_t dmask, T dtr)
{
  for (int b = 0; b < nb; b++) {
    const int b0 = S.b0[b][d], b1 = S.b1[b][d];
    T B0 = uc[0], B1 = uc[0];
#pragma unroll
    for (int s = 0; s < NS; s++) {
      if (s == b0) B0 = uc[s];
      if (s == b1) B1 = uc[s];
    }
    /* kon (tot - B) F - koff B about rest, B and F deviations */
    const T J = S.rj0[b][d] + S.kon[b][d] * (S.ra[b][d] * B1 - B0 * (S.rf0[b][d] + B1)) - S.koff[b][d] * B0;
    const T inc = dtr * J;
#pragma unroll
    for (int s = 0; s < NS; s++)
      if (s == b0) nw[s] = (((dmask >> s) & 1u) ? nw[s] : uc[s]) + inc;
#pragma unroll
    for (int s = 0; s < NS; s++)
      if (s == b1) nw[s] = (((dmask >> s) & 1u) ? nw[s] : uc[s]) - inc;
  }
}

/* The same reactions in species-major form, for models in which every
 * buffer binds species 0 and each bound species belongs to one buffer of
 * its domain, on steps where species 0 diffuses: then no overwrite can
 * discard an increment (a bound species is written once, and nw = uc for
 * one that did not diffuse), slots a domain lacks are exact no-ops, and
 * the result equals react_general's up to the order in which species 0
 * sums its increments (species order here, slot order there).  Species
 * indices are compile-time, so no per-lane selection is needed. */
template <typename T, int NS>
__device__ inline void react_species_major(T nw[NS], const T uc[NS], const BlockTables<T, NS>& S, int d,
                                           uint32_t rbound)
{
  /* dt_r J = dt_r (j0 + k_on (a F - B (C0 + F)) - k_off B) about rest, with
   * deviations F of species 0 and B of the buffer: three fused
   * multiply-adds per buffer, the absolute species 0 shared */
  const T f0 = S.rc0[d] + uc[0];
  T f = nw[0];
#pragma unroll
  for (int s = 1; s < NS; s++) {
    if (!((rbound >> s) & 1u)) continue;
    const RateTab<T> r = S.rt[d][s];
    const T inc = (r.dka * uc[0] + r.dj0) - uc[s] * (r.dkon * f0 + r.dkoff);
    nw[s] += inc;
    f -= inc;
  }
  nw[0] = f;
}

template <typename T>
__device__ inline void store_next(T* p, T v, bool nt)
{
  if (nt) __builtin_nontemporal_store(v, p);
  else *p = v;
}

/* Raw buffer access: a descriptor (base and byte size, uniform across the
 * wavefront), a per-lane byte offset and a uniform byte offset, e.g. the
 * plane.  Nothing is left for the compiler to turn into per-lane 64-bit
 * address arithmetic, and an access beyond the descriptor's size reads 0
 * and writes nothing instead of faulting.  Word 3 is the raw-buffer
 * format word of the gfx9 family (32-bit data format). */
typedef __amdgpu_buffer_rsrc_t Rsrc;

__device__ inline Rsrc make_rsrc(const void* base, uint32_t bytes)
{
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)bytes, 0x00020000);
}

template <typename T>
__device__ inline T bload(Rsrc r, uint32_t voff, uint32_t soff);
template <>
__device__ inline float bload<float>(Rsrc r, uint32_t voff, uint32_t soff)
{
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
template <>
__device__ inline double bload<double>(Rsrc r, uint32_t voff, uint32_t soff)
{
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
}
__device__ inline uint32_t bload_u8(Rsrc r, uint32_t voff, uint32_t soff)
{
  return __builtin_amdgcn_raw_buffer_load_b8(r, (int)voff, (int)soff, 0);
}

/* aux 2 sets the non-temporal bit (nt) */
template <typename T>
__device__ inline void bstore(Rsrc r, uint32_t voff, uint32_t soff, T v, bool nt);
template <>
__device__ inline void bstore<float>(Rsrc r, uint32_t voff, uint32_t soff, float v, bool nt)
{
  if (nt) __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, (int)voff, (int)soff, 2);
  else __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, (int)voff, (int)soff, 0);
}
template <>
__device__ inline void bstore<double>(Rsrc r, uint32_t voff, uint32_t soff, double v, bool nt)
{
  typedef unsigned int u32x2 __attribute__((__vector_size__(2 * sizeof(unsigned int))));
  const u32x2 b = __builtin_bit_cast(u32x2, v);
  if (nt) __builtin_amdgcn_raw_buffer_store_b64(b, r, (int)voff, (int)soff, 2);
  else __builtin_amdgcn_raw_buffer_store_b64(b, r, (int)voff, (int)soff, 0);
}

/* One-step pass over the columns of one block: lane (x, y) marches
 * planes [z0, z1) of grid g with a register ring of planes z-1, z, z+1
 * (and z+2 with PF 2) and writes every species of the next state: the
 * 7-point stencil of the species in dmask (species.c:1501-1650,
 * kernel_generic.cpp:512-600), buffer reactions (react 1: slot order,
 * 2: species-major; reaction_SIMD, kernel_generic.cpp:397-509), the
 * Dirichlet faces and the fixed domain species (species.c:1037-1077).
 * u, un and ids point at the grid's first z-group; the lane's (x, y) must
 * lie inside the grid.  S holds the block's copy of P's tables. */
template <typename T, int NS, int PF, bool NT>
__device__ __forceinline__ void stencil_columns(const T* __restrict__ u, T* __restrict__ un,
                                                const uint8_t* __restrict__ ids,
                                                const DevParams<T>* __restrict__ P,
                                                const BlockTables<T, NS>& S, const Geo& g, uint32_t dmask,
                                                uint32_t react, const int x, const int y, const int z0,
                                                const int z1)
{
  const int nb = react == 1 ? (int)P->nb : 0;
  const uint32_t rbound = P->rbound;
  const T dtr = P->dt_react;
  const bool dirichlet = P->dirichlet != 0;
  /* clamped neighbour offsets: a missing neighbour reads the voxel
   * itself, so its (u_n - u) term vanishes -- the no-flux outer faces.
   * A lane's column is fixed while it marches z, so its offsets within a
   * plane are computed once (32-bit byte offsets; the host checks
   * plane * sizeof(T) < 2^31); the loop does no per-lane address
   * arithmetic. */
  const uint32_t ES = (uint32_t)sizeof(T);
  const uint32_t col = (uint32_t)y * (uint32_t)g.nx + (uint32_t)x;
  const uint32_t bc = col * ES;
  const uint32_t bxm = (x > 0 ? col - 1 : col) * ES, bxp = (x + 1 < g.nx ? col + 1 : col) * ES;
  const uint32_t bym = (y > 0 ? col - g.nx : col) * ES, byp = (y + 1 < g.ny ? col + g.nx : col) * ES;
  const bool xyface = x == 0 || y == 0 || x + 1 == g.nx || y + 1 == g.ny;
  const int ztop_plane = g.nz - 1;
  const size_t plane = g.plane;
  /* x neighbours come from the adjacent lane (DPP): lanes hold consecutive
   * cells of a row in both tilings.  Lane 0 loads its x-1 and lane 63 its
   * x+1, which lie outside the wavefront; at a row end the neighbour is
   * the voxel itself.  Lanes that left early (past the row or the plane)
   * are never read, since their left neighbour sits at x = nx-1. */
  const int lane = threadIdx.x;
  const bool has_xm = x > 0, has_xp = x + 1 < g.nx;
  const bool edge_lane = (lane == 0 && has_xm) || (lane == BX - 1 && has_xp);
  const uint32_t bedge = lane == 0 ? bxm : bxp;

  /* descriptors over the z-groups this block touches: [zb, ze) of u and
   * of the ids, [z0, z1) of the next state; species s of group z sits at
   * byte (z - base) * gbytes + s * pbytes (the host keeps (zchunk + 3)
   * groups below 2^32 bytes) */
  const int zb = max(z0 - 1, 0), ze = min(g.nz, z1 + 2);
  const uint32_t pbytes = (uint32_t)plane * ES;
  const uint32_t gbytes = (uint32_t)NS * pbytes;
  const Rsrc rin = make_rsrc(u + (size_t)zb * NS * plane, (uint32_t)(ze - zb) * gbytes);
  const Rsrc rout = make_rsrc(un + (size_t)z0 * NS * plane, (uint32_t)(z1 - z0) * gbytes);
  const Rsrc rid = make_rsrc(ids + (size_t)zb * plane, (uint32_t)(ze - zb) * (uint32_t)plane);

  /* One plane: um, uc, up hold planes z-1, z, z+1 (indices clamped, so a
   * missing z neighbour equals uc).  With PF 1 up is loaded here; with
   * PF 2 it was loaded one plane earlier and uq (plane z+2) is loaded here,
   * after this plane's neighbours, so the wait for the neighbours (mostly
   * cache hits: plane z arrived a plane ago) leaves uq in flight. */
  auto plane_step = [&](const int z, const T (&um)[NS], const T (&uc)[NS], T (&up)[NS], T (&uq)[NS]) {
    const uint32_t sz = (uint32_t)(z - zb) * gbytes;           /* group z in rin */
    const uint32_t szi = (uint32_t)(z - zb) * (uint32_t)plane; /* plane z in rid */
    const bool ztop = z < ztop_plane;
    if (PF == 1) {
      const uint32_t sp = (uint32_t)(min(z + 1, ztop_plane) - zb) * gbytes;
#pragma unroll
      for (int s = 0; s < NS; s++) up[s] = bload<T>(rin, bc, sp + s * pbytes);
    }
    /* the in-plane neighbours of every diffusing species are loaded before
     * any of them is used, so one wait covers them all */
    const uint32_t id = bload_u8(rid, col, szi);
    T nxm[NS], nxp[NS], nym[NS], nyp[NS], nxe[NS];
#pragma unroll
    for (int s = 0; s < NS; s++)
      if ((dmask >> s) & 1u) {
        nym[s] = bload<T>(rin, bym, sz + s * pbytes);
        nyp[s] = bload<T>(rin, byp, sz + s * pbytes);
        nxe[s] = uc[s];
        if (edge_lane) nxe[s] = bload<T>(rin, bedge, sz + s * pbytes);
      }
    if (PF == 2) {
      asm volatile("" ::: "memory");   /* keep the plane-ahead loads last */
      const uint32_t sq = (uint32_t)(min(z + 2, ztop_plane) - zb) * gbytes;
#pragma unroll
      for (int s = 0; s < NS; s++) uq[s] = bload<T>(rin, bc, sq + s * pbytes);
    } else {
      /* the next plane is complete before the first store, so only stores
       * are outstanding between planes */
#pragma unroll
      for (int s = 0; s < NS; s++) asm volatile("" ::"v"(up[s]));
    }
#pragma unroll
    for (int s = 0; s < NS; s++)
      if ((dmask >> s) & 1u) {
        const T l = dpp_from_left(uc[s], nxe[s]), r = dpp_from_right(uc[s], nxe[s]);
        nxm[s] = has_xm ? l : uc[s];
        nxp[s] = has_xp ? r : uc[s];
      }
    const int d = id & 7;
    const bool uni = (id & 0x80) != 0;
    T nw[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) nw[s] = uc[s];
    if (__ballot(!uni) == 0) {
      /* every lane's six neighbours share its domain (the common case,
       * decided once per wavefront) */
#pragma unroll
      for (int s = 0; s < NS; s++)
        if ((dmask >> s) & 1u) {
          const T c = uc[s];
          const T exm = nxm[s] - c, exp_ = nxp[s] - c;
          const T eym = nym[s] - c, eyp = nyp[s] - c;
          const T ezm = um[s] - c, ezp = up[s] - c;
          nw[s] = c + S.coef[s][d * MAXD + d] * ((ezm + ezp) + (eym + eyp) + (exm + exp_));
        }
    } else {
      int dxm = d, dxp = d, dym = d, dyp = d, dzm = d, dzp = d;
      if (!uni) {
        dxm = bload_u8(rid, bxm / ES, szi) & 7;
        dxp = bload_u8(rid, bxp / ES, szi) & 7;
        dym = bload_u8(rid, bym / ES, szi) & 7;
        dyp = bload_u8(rid, byp / ES, szi) & 7;
        dzm = z > 0 ? bload_u8(rid, col, szi - (uint32_t)plane) & 7 : d;
        dzp = ztop ? bload_u8(rid, col, szi + (uint32_t)plane) & 7 : d;
      }
#pragma unroll
      for (int s = 0; s < NS; s++)
        if ((dmask >> s) & 1u) {
          const T c = uc[s];
          const T exm = nxm[s] - c, exp_ = nxp[s] - c;
          const T eym = nym[s] - c, eyp = nyp[s] - c;
          const T ezm = um[s] - c, ezp = up[s] - c;
          const T* C = S.coef[s] + d * MAXD;
          if (uni) {
            nw[s] = c + C[d] * ((ezm + ezp) + (eym + eyp) + (exm + exp_));
          } else {
            /* deviations of different domains differ by their resting values */
            const T* D = S.dinit[s] + d * MAXD;
            nw[s] = c + ((C[dzm] * (ezm + D[dzm]) + C[dzp] * (ezp + D[dzp])) +
                         (C[dym] * (eym + D[dym]) + C[dyp] * (eyp + D[dyp])) +
                         (C[dxm] * (exm + D[dxm]) + C[dxp] * (exp_ + D[dxp])));
          }
        }
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
    const uint32_t so = (uint32_t)(z - z0) * gbytes;
#pragma unroll
    for (int s = 0; s < NS; s++) bstore<T>(rout, bc, so + s * pbytes, nw[s], NT);
  };

  /* register ring of PF + 2 planes, rotated by renaming: the loop body
   * holds every position of the ring, so no values are copied; planes past
   * z1 in the last pass are skipped by uniform branches */
  T ra[NS], rb[NS], rc[NS], rd[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) {
    ra[s] = bload<T>(rin, bc, (uint32_t)(max(z0 - 1, 0) - zb) * gbytes + s * pbytes);
    rb[s] = bload<T>(rin, bc, (uint32_t)(z0 - zb) * gbytes + s * pbytes);
    if (PF == 2) rc[s] = bload<T>(rin, bc, (uint32_t)(min(z0 + 1, ztop_plane) - zb) * gbytes + s * pbytes);
  }
  /* drain the prologue loads, so the loop header's wait does not also
   * cover the previous plane's stores (stores count in vmcnt on gfx9) */
  __builtin_amdgcn_s_waitcnt(0);
  if (PF == 1) {
    for (int z = z0; z < z1; z += 3) {
      plane_step(z, ra, rb, rc, rd);
      if (z + 1 < z1) plane_step(z + 1, rb, rc, ra, rd);
      if (z + 2 < z1) plane_step(z + 2, rc, ra, rb, rd);
    }
  } else {
    for (int z = z0; z < z1; z += 4) {
      plane_step(z, ra, rb, rc, rd);
      if (z + 1 < z1) plane_step(z + 1, rb, rc, rd, ra);
      if (z + 2 < z1) plane_step(z + 2, rc, rd, ra, rb);
      if (z + 3 < z1) plane_step(z + 3, rd, ra, rb, rc);
    }
  }
}


/* State of channel ch of discrete flux F after the gating step at time
 * index ti (evaluate_RyR_stochastically, boundaryfluxes.c:295-359): a
 * channel with a forced closing time (-C) closes once that time falls
 * within the gating interval; a stochastic one opens (closes) when its
 * uniform is at most dt_s times the opening (closing) rate at the absolute
 * Ca concentration c0 of its cleft voxel. */
__device__ inline uint8_t gate_one(const FluxParams& F, uint8_t st, double c0, uint64_t ti, double t, double dts,
                                   uint32_t seed, uint32_t instance, const double* __restrict__ draw, uint32_t ch)
{
  const double* p = F.p;
  if (!F.stochastic) {
    if (p[8] <= t + dts) st = 0;
  } else {
    const double kc = cs_rate_close(c0, p);
    const double ko = cs_rate_open(c0, p);
    /* the caller's uniform for this channel, or the Philox stream */
    const double r = draw ? *draw : cs_gate_uniform(seed, instance, ti, F.group, ch);
    if (st == 0) { if (r <= dts * ko) st = 1; }
    else if (st == 1) { if (r <= dts * kc) st = 0; }
  }
  return st;
}

/* Flux of one boundary pair (species.c:1704-1807) from the stored
 * deviations vf, vt at its from and to voxels and their resting values
 * r0, r1: the RyR two-voxel relaxation (ryr_flux, boundaryfluxes.c:87-141)
 * or the SERCA pump (serca_flux, boundaryfluxes.c:12-85). */
__device__ inline double flux_j(const FluxParams& F, double vf, double vt, double r0, double r1, double dt,
                                double h)
{
  if (F.kind == CS_FLUX_RYR) return ((vf - vt) + (r0 - r1)) * F.half_1mK0;
  return cs_serca_flux_d(dt, h, vf + r0, vt + r1, F.density);
}

/* marks a missing row, slot or channel state in the flat records */
constexpr uint32_t NO_ROW = 0xFFFFFFFFu;

/* Flat records of the boundary fluxes for k_bridge, resolved on the host
 * so that a channel, a row or an entry is one load: element indices of
 * the field (species s at voxel v), level-1 slots, resting values. */
struct BChan {          /* a channel of a discrete flux, in gating order */
  uint64_t elem_to;     /* species 0 at the channel's cleft (to) voxel    */
  double rest_to;       /* resting species 0 there                        */
  uint32_t slot_to;     /* level-1 slot of that voxel                     */
  uint32_t flux, ch;    /* flux index, channel index within the flux      */
  uint32_t pad;
};
struct BRow {           /* a k_flux row                                   */
  uint64_t elem;        /* the value the row adds to                      */
  uint32_t start, end;  /* its entries                                    */
  uint32_t skip, pad;   /* 1: species fixed in the domain, or species 0 on
                           a Dirichlet face (the row takes nothing)       */
};
struct BEnt {           /* a pair of a row, in k_flux's order             */
  uint64_t elem_from, elem_to;
  double rest_from, rest_to;
  uint32_t slot_from, slot_to;
  uint32_t state;       /* the channel's state index, or NO_ROW           */
  uint32_t flux_sign;   /* flux | (row at the to voxel) << 8              */
};

inline uint32_t ref_to_dev(uint32_t i, const uint32_t n[3])
{
  const uint32_t z = i % n[2], y = (i / n[2]) % n[1], x = i / (n[2] * n[1]);
  return (z * n[1] + y) * n[0] + x;
}

/* Device tables of problem p (ndiscrete: its number of discrete fluxes). */
template <typename T>
void fill_params(DevParams<T>& dp, const cs_problem* p, uint32_t ndiscrete)
{
  memset(&dp, 0, sizeof dp);
  for (int s = 0; s < MAXS; s++)
    for (int i = 0; i < MAXD * MAXD; i++) dp.coef[s][i] = (T)p->coef[s][i];
  /* resting state: the device stores u - init[domain][species]; the
   * expansions about it are formed in double and rounded once */
  for (int d = 0; d < MAXD; d++)
    for (int s = 0; s < MAXS; s++) {
      dp.init64[d][s] = p->init[d][s];
      if (p->fixed[d][s]) dp.fixed_mask[d] |= 1u << s;
    }
  for (int s = 0; s < MAXS; s++)
    for (int d = 0; d < MAXD; d++)
      for (int dn = 0; dn < MAXD; dn++) dp.dinit[s][d * MAXD + dn] = (T)(p->init[dn][s] - p->init[d][s]);
  for (int b = 0; b < MAXS; b++)
    for (int d = 0; d < MAXD; d++) {
      const double kon = p->k_on[b][d], koff = p->k_off[b][d];
      const double B0 = p->init[d][p->bsp0[b][d] % MAXS], F0 = p->init[d][p->bsp1[b][d] % MAXS];
      const double a = p->tot[b][d] - B0;
      dp.kon[b][d] = (T)kon;
      dp.koff[b][d] = (T)koff;
      dp.ra[b][d] = (T)a;
      dp.rf0[b][d] = (T)F0;
      dp.rj0[b][d] = (T)(kon * a * F0 - koff * B0);
      dp.bsp0[b][d] = p->bsp0[b][d];
      dp.bsp1[b][d] = p->bsp1[b][d];
    }
  /* species-major table: usable when every buffer slot in use binds
   * species 0 (its free species) and no domain binds a species twice */
  bool fast = p->num_buffers > 0;
  for (int d = 0; d < MAXD; d++) {
    uint32_t seen = 0;
    for (int b = 0; b < MAXS; b++) {
      const bool used = p->bsp0[b][d] || p->bsp1[b][d] || p->k_on[b][d] != 0 || p->k_off[b][d] != 0 ||
                        p->tot[b][d] != 0;
      if (!used) continue;
      const int b0 = p->bsp0[b][d];
      if (b >= (int)p->num_buffers || p->bsp1[b][d] != 0 || b0 == 0 || ((seen >> b0) & 1u)) {
        fast = false;
        continue;
      }
      seen |= 1u << b0;
      const double kon = p->k_on[b][d], koff = p->k_off[b][d];
      const double B0 = p->init[d][b0], C0 = p->init[d][0], a = p->tot[b][d] - B0;
      const double dtr = p->dt_react;
      dp.rt[d][b0].dkon = (T)(dtr * kon);
      dp.rt[d][b0].dkoff = (T)(dtr * koff);
      dp.rt[d][b0].dka = (T)(dtr * kon * a);
      dp.rt[d][b0].dj0 = (T)(dtr * (kon * a * C0 - koff * B0));
      dp.rbound |= 1u << b0;
    }
  }
  dp.rfast = fast ? 1u : 0u;
  for (int d = 0; d < MAXD; d++) dp.rc0[d] = (T)p->init[d][0];
  dp.dt_react = (T)p->dt_react;
  dp.nb = p->num_buffers;
  dp.dirichlet = p->force_dirichlet;
  dp.nflux = p->num_fluxes;
  dp.ndiscrete = ndiscrete;
  dp.dt = p->dt;
  dp.h = p->h;
  uint32_t poff = 0, soff = 0, uoff = 0, group = 0;
  for (uint32_t f = 0; f < p->num_fluxes; f++) {
    const cs_flux_desc& F = p->fluxes[f];
    FluxParams& o = dp.flux[f];
    o.kind = F.kind;
    o.species = F.species;
    o.discrete = F.discrete;
    o.inner = F.inner;
    o.npairs = F.num_pairs;
    o.pair_off = poff;
    o.state_off = soff;
    poff += F.num_pairs;
    if (F.discrete) {
      soff += F.num_pairs;
      o.group = group++;
      o.stochastic = !(F.params[8] > 0);
      o.uniform_off = uoff;
      if (o.stochastic) uoff += F.num_pairs;
    }
    memcpy(o.p, F.params, sizeof o.p);
    o.half_1mK0 = F.kind == CS_FLUX_RYR ? (1 - cs_ryr_k0(p->dt, p->h, F.params[6])) / 2 : 0.0;
    o.density = F.kind == CS_FLUX_SERCA ? F.params[1] : 0.0;
  }
}

}  // namespace

#endif /* CS_DEVICE_H */
