/*
 * oracle.h -- CPU FP64 restatement of the calcium_sparks time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the GPU path is
 * compared against (tests/, __graft_entry__.smoke(), bench.py's CPU
 * baseline leg); the product never links or calls it.
 *
 * It re-derives everything from the raw inputs (cs_model.h) independently
 * of the product's builder, and steps the fields in the reference layout
 * u[species][x][y][z] (z fastest) with the reference's u1/du/update-flag
 * sequence (calcium_sparks.c:190-228).  Each function cites the reference
 * lines it restates.
 *
 * Gating uniforms: CS_RNG_MT draws one mt_ldrand per stochastic channel in
 * the reference's order; CS_RNG_PHILOX draws Philox4x32-10 with key
 * (seed, instance) and counter (time_ind low, time_ind high, gating group,
 * channel), where the gating group is the ordinal of the discrete boundary
 * among the discrete boundaries carrying a flux model, in file order, and
 * the uniform is ((x0 << 21) | (x1 >> 11)) * 2^-53.
 */
#ifndef CS_ORACLE_H
#define CS_ORACLE_H

#include <stddef.h>
#include <stdint.h>
#include "cs_model.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_sim oracle_sim;

typedef struct oracle_info {
  uint64_t n[3];
  uint32_t subdivisions;
  double   h, dt, dV;
  uint32_t num_species, num_domains;
  uint32_t substeps[CS_MAX_SPECIES];                 /* per species, 0 = none */
  double   coef[CS_MAX_SPECIES][CS_MAX_DOMAINS * CS_MAX_DOMAINS];
  uint32_t num_all_buffers, reaction_substep, stochastic_substep;
  uint32_t num_discrete;                             /* discrete boundaries */
  uint32_t discrete_size[CS_MAX_BOUNDARIES];
  int32_t  discrete_gated[CS_MAX_BOUNDARIES];        /* gating group or -1  */
  uint32_t seed;
} oracle_info;

int  oracle_create(const cs_geometry* g, const cs_species_model* m,
                   const cs_options* opt, oracle_sim** out, char* err, size_t err_len);
void oracle_free(oracle_sim* s);
void oracle_get_info(const oracle_sim* s, oracle_info* info);

/* One iteration of the reference loop body (diffusion, stochastic events,
 * boundary fluxes, reaction, apply du, fixed values) for time_ind. */
int  oracle_step(oracle_sim* s, uint64_t time_ind);
int  oracle_steps(oracle_sim* s, uint64_t first_time_ind, uint64_t nsteps);

void oracle_get_species(const oracle_sim* s, uint32_t species, double* out);
void oracle_set_species(oracle_sim* s, uint32_t species, const double* in);

/* min / max / sum(u*dV) per (domain, species), index d*ns+s
 * (Species_output_scalar_data, species.c:2806-2863). */
void oracle_stats(const oracle_sim* s, double* mn, double* mx, double* tot);

/* Linescan of one species (species.c:1960-2035): out[N[axis]]. */
void oracle_linescan(const oracle_sim* s, uint32_t species, uint32_t axis, const int32_t offsets[3],
                     uint32_t num_domains, const double conv[3], double* out);

/* Channel states of discrete boundary k (file order among discrete). */
int  oracle_get_states(const oracle_sim* s, uint32_t k, uint8_t* out);
int  oracle_set_states(oracle_sim* s, uint32_t k, const uint8_t* in);
/* channels of discrete boundary k ever open since creation */
int  oracle_get_recruited(const oracle_sim* s, uint32_t k, uint8_t* out);

/* Reference scalar functions, restated (boundaryfluxes.c:12-141). */
double oracle_ryr_flux(double dt, double h, double u0, double u1, const double* p);
double oracle_serca_flux(double dt, double h, double u0, double u1, const double* p);

/* Generators, restated from their published algorithms. */
void   oracle_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void   oracle_mt_seed(uint32_t seed);       /* global stream like mtwist    */
double oracle_mt_ldrand(void);
uint32_t oracle_mt_u32(void);

/* Dense stencil of one species on an arbitrary grid (reference interior
 * formula applied everywhere with no-flux faces); used to pin the
 * stencil against closed forms.  du = u + sum_n C[d][dn]*(u_n-u). */
void oracle_stencil(const double* u, double* du, const double* coef5x5,
                    const uint8_t* ids, uint64_t nx, uint64_t ny, uint64_t nz);

#ifdef __cplusplus
}
#endif
#endif
