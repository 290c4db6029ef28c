/*
 * cs_host.h -- host side of the calcium_sparks path (libcshost.so, plain C).
 *
 *  - cs_setup_build: derives the device problem (cs_problem) from the raw
 *    inputs the way Geometry_construct + Species_construct +
 *    BoundaryFluxes_construct + BoundaryFluxes_init_stochastic_boundaries
 *    do (reference geometry.c:16-356, species.c:95-884, :938-1035,
 *    boundaryfluxes.c:361-605): refinement, dt, per-species substeps,
 *    5x5 coefficient tables, reaction tables, boundary voxel pairs, initial
 *    channel states.
 *  - cs_mt_*: the Mersenne Twister (MT19937 with 32-bit seeding) that the
 *    reference draws its gating uniforms from (mt_seed32new / mt_ldrand,
 *    reference mtwist.c:407-439, mtwist.h:635-677).
 *  - cs_loop_*: the reference's time loop bookkeeping (calcium_sparks.c:
 *    167-168, :190-228; abort rule species.c:3384-3488).
 */
#ifndef CS_HOST_H
#define CS_HOST_H

#include <stddef.h>
#include <stdint.h>
#include "cs_model.h"
#include "csparks.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Mersenne Twister ---------------------------------------------------*/
typedef struct cs_mt {
  uint32_t mt[624];
  uint32_t idx;
} cs_mt;

void     cs_mt_seed(cs_mt* g, uint32_t seed);   /* init_genrand(seed)       */
uint32_t cs_mt_u32(cs_mt* g);                   /* genrand_int32            */
double   cs_mt_ldrand(cs_mt* g);                /* (u32<<32 | u32) * 2^-64  */

/* ---- setup -------------------------------------------------------------*/
typedef struct cs_discrete_info {
  char     name[CS_NAME_LEN];    /* boundary name                          */
  uint32_t boundary;             /* index in the geometry file             */
  int32_t  flux;                 /* index into problem.fluxes, -1 if none  */
  uint32_t size;                 /* number of channels                     */
} cs_discrete_info;

typedef struct cs_setup {
  cs_problem problem;            /* arrays owned by the setup              */
  uint32_t subdivisions;
  double   h_geom;
  uint64_t N[3];                 /* refined global sizes (= problem.n)     */
  uint32_t num_domains, num_species;
  char     domain_names[CS_MAX_DOMAINS][CS_NAME_LEN];
  char     species_names[CS_MAX_SPECIES][CS_NAME_LEN];
  double   volumes[CS_MAX_DOMAINS];       /* geometry volume per domain    */
  int64_t  domain_voxels[CS_MAX_DOMAINS]; /* geometry voxels per domain    */
  uint32_t num_discrete;
  cs_discrete_info discrete[CS_MAX_BOUNDARIES];
  uint32_t num_diffusive_species;         /* num_all_diffusive             */
  uint8_t  all_diffusive[CS_MAX_SPECIES]; /* diffusive index -> species    */
  uint32_t has_ryr_boundary, has_p_ryr_boundary; /* names "ryr"/"p_ryr"    */
  int32_t  ryr_discrete, p_ryr_discrete;  /* discrete index of those, -1   */
  cs_mt    mt;                   /* generator after the stochastic init    */
  uint32_t seed;                 /* seed used (given or drawn)             */
} cs_setup;

/* Build a setup.  Returns CS_OK or CS_EINVAL with a message in err. */
int  cs_setup_build(const cs_geometry* geom, const cs_species_model* model,
                    const cs_options* opt, cs_setup** out, char* err,
                    size_t err_len);
void cs_setup_free(cs_setup* s);

/* Uniforms for cs_step_uniforms in the reference's draw order, taken from
 * the setup's Mersenne Twister (advancing it).  out may be NULL to query
 * *count only (then the generator does not advance). */
int cs_setup_mt_uniforms(cs_setup* s, uint64_t first_time_ind, uint64_t nsteps,
                         double* out, uint64_t* count);

/* ---- time loop ---------------------------------------------------------*/
/* save_interval = max(1, floor(dt_save/dt))  (calcium_sparks.c:167-168) */
uint64_t cs_save_interval(double dt_save, double dt);

/* Number of iterations of for (t=0; t<tstop; t+=dt) with t accumulated in
 * double exactly like calcium_sparks.c:190. */
uint64_t cs_count_steps(double tstop, double dt);

/* Abort rule of Species_check_open_discrete_boundaries (species.c:3384-
 * 3488), driven by the channel snapshot the reference refreshes at output
 * time.  state->last_closed starts at -1. */
typedef struct cs_abort_state {
  double tabort;
  double last_closed;
  int    all_closed;             /* snapshot: all ryr / p_ryr closed        */
  int    active;                 /* tabort != 0 and discrete boundaries    */
} cs_abort_state;

void cs_abort_init(cs_abort_state* a, double tabort, int have_discrete);
/* Record a fresh channel snapshot (taken at an output). */
void cs_abort_snapshot(cs_abort_state* a, int all_closed);
/* Evaluate the rule at time t (= t+dt of the step just taken); returns 1
 * when the run must stop. */
int  cs_abort_check(cs_abort_state* a, double t);

/* sizeof of cs_geometry (0), cs_species_model (1), cs_options (2),
 * cs_problem (3), cs_setup (4), cs_flux_desc (5); 0 otherwise. */
size_t cs_host_sizeof(int which);

#ifdef __cplusplus
}
#endif
#endif /* CS_HOST_H */
