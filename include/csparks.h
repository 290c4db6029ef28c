/*
 * csparks.h -- C ABI of the MI355X time-stepping library (libcsparks.so).
 *
 * One call of cs_step() advances the resident domain by whole time steps of
 * the reference's explicit scheme (calcium_sparks.c:190-228):
 *
 *   diffusion      Species_step_diffusion          species.c:1501-1650
 *                  kernel_SIMD / kernel_LOTR       kernel_generic.cpp:512-600
 *                  side_SIMD + scalar borders      kernel_generic.cpp:175-229,
 *                                                  species.c:1282-1498
 *   gating         Species_evaluate_stochastic_events species.c:1910-1957
 *                  evaluate_RyR_stochastically     boundaryfluxes.c:295-359
 *   boundary flux  Species_step_boundary_fluxes    species.c:1704-1807
 *                  ryr_flux / serca_flux           boundaryfluxes.c:87-141, :12-85
 *   reaction       Species_step_reaction           species.c:1653-1702
 *                  reaction_SIMD                   kernel_generic.cpp:397-509
 *   apply du       Species_apply_du                species.c:1080-1121
 *   fixed values   Species_apply_fixed_values      species.c:1037-1077
 *
 * The whole state lives on one GPU; there is no ghost exchange.  Host
 * arrays crossing this boundary use the reference layout [n0][n1][n2]
 * (z fastest) and double precision, whatever precision the device uses.
 *
 * Errors: every entry point returns CS_OK (0) or a negative CS_E* code and
 * records a message retrievable with cs_last_error().  Contexts are not
 * re-entrant: one host thread per context.
 */
#ifndef CSPARKS_H
#define CSPARKS_H

#include <stddef.h>
#include <stdint.h>
#include "cs_model.h"

#ifdef __cplusplus
extern "C" {
#endif

#define CS_MAX_FLUXES 4
#define CS_NUM_RYR_PARAMS 10

enum {
  CS_OK = 0,
  CS_EINVAL = -1,     /* bad argument / inconsistent descriptor          */
  CS_EDEVICE = -2,    /* HIP runtime error                               */
  CS_ENOMEM = -3,     /* device or host allocation failed                */
  CS_ERANGE = -4      /* index or size out of range                      */
};

/* Flux models of the reference's plugin table (basefluxinfo.h:34-46). */
enum { CS_FLUX_RYR = 0, CS_FLUX_SERCA = 1 };

/* One boundary flux.  pairs[] holds (from, to) linear voxel indices
 * x*n1*n2 + y*n2 + z of the refined grid (reference species.c:938-1035).
 * params[] keeps the reference's parameter vector order:
 *   RyR  : Kd_close, k_max_close, n_open, k_max_open, k_min_open, n_close,
 *          i_unitary, k_min_close, t_close, Kd_open (boundaryfluxes.c:447-456)
 *   SERCA: v_cyt_to_A_sr, density, scale            (boundaryfluxes.c:536-538)
 * Discrete fluxes are channels gated every stochastic substep; a positive
 * t_close (params[8]) closes them deterministically at t_close without
 * drawing random numbers. */
typedef struct cs_flux_desc {
  uint32_t kind;
  uint32_t species;          /* species slot moved across the boundary   */
  uint32_t substep;          /* fires when time_ind % substep == 0       */
  uint32_t inner;            /* 1: also subtract J at the from voxel     */
  uint32_t discrete;         /* 1: channel pairs with open/closed state  */
  uint32_t num_pairs;
  const uint32_t* pairs;     /* [num_pairs][2]                           */
  double   params[CS_NUM_RYR_PARAMS];
  const uint8_t* init_states;/* [num_pairs] 0/1, discrete only           */
} cs_flux_desc;

/* Everything one time step needs, already derived from the input files
 * (dt, substeps, coefficient tables, refined ids, boundary voxels). */
typedef struct cs_problem {
  uint32_t n[3];             /* refined voxels per axis                  */
  uint32_t num_domains, num_species;
  uint32_t precision;        /* CS_PREC_F64 / CS_PREC_F32                */
  uint32_t rng;              /* CS_RNG_PHILOX / CS_RNG_MT                */
  uint64_t seed;             /* Philox key word 0                        */
  uint64_t instance;         /* Philox key word 1                        */
  double   h, dt, dV;
  const uint8_t* domain_ids; /* [n0][n1][n2], z fastest                  */
  double   init[CS_MAX_DOMAINS][CS_MAX_SPECIES];
  /* diffusion: substeps[s] == 0 marks a non-diffusive species; coef[s] is
   * the 5x5 table dt*k_s*sigma/h^2 of species.c:1547-1560 */
  uint32_t substeps[CS_MAX_SPECIES];
  double   coef[CS_MAX_SPECIES][CS_MAX_DOMAINS * CS_MAX_DOMAINS];
  /* reaction tables, transposed [buffer][domain] like species.c:249-253 */
  uint32_t num_buffers, reaction_substep;
  double   dt_react;
  uint8_t  bsp0[CS_MAX_SPECIES][CS_MAX_DOMAINS];
  uint8_t  bsp1[CS_MAX_SPECIES][CS_MAX_DOMAINS];
  double   k_on[CS_MAX_SPECIES][CS_MAX_DOMAINS];
  double   k_off[CS_MAX_SPECIES][CS_MAX_DOMAINS];
  double   tot[CS_MAX_SPECIES][CS_MAX_DOMAINS];
  uint32_t num_fluxes;
  cs_flux_desc fluxes[CS_MAX_FLUXES];  /* discrete ones in gating order  */
  uint32_t stochastic_substep;
  uint32_t force_dirichlet;  /* -f: species 0 on the outer faces := init */
  uint8_t  fixed[CS_MAX_DOMAINS][CS_MAX_SPECIES]; /* fixed domain species */
} cs_problem;

typedef struct cs_context cs_context;

/* Replaces Species_construct + Species_init_values + Species_init_du
 * (species.c:95-884, :887-935, :22-32): allocates the device state, uploads
 * the problem and fills every voxel with init[domain][species]. */
int cs_create(const cs_problem* problem, cs_context** ctx);
int cs_destroy(cs_context* ctx);

/* Device ordinal the library runs on, and its name. */
int cs_device_info(int* device, char* name, size_t name_len);

/* Reset the fields to init[domain][species] and the channels to the
 * problem's init_states (Species_init_values, species.c:887-935). */
int cs_reset(cs_context* ctx);

/* Field transfer in the reference layout, double precision. */
int cs_upload_species(cs_context* ctx, uint32_t species, const double* src);
int cs_download_species(cs_context* ctx, uint32_t species, double* dst);

/* Advance nsteps steps with time indices first_time_ind .. +nsteps-1
 * (the reference's time_ind starts at 1).  Asynchronous on the context's
 * stream.  In CS_RNG_MT mode use cs_step_uniforms instead. */
int cs_step(cs_context* ctx, uint64_t first_time_ind, uint64_t nsteps);

/* Same as cs_step, with the gating uniforms supplied by the caller: for
 * every gating step in the range (time_ind % stochastic_substep == 0), in
 * order, one draw per channel of every discrete flux whose t_close <= 0,
 * fluxes in descriptor order.  count must equal that total. */
int cs_step_uniforms(cs_context* ctx, uint64_t first_time_ind, uint64_t nsteps,
                     const double* uniforms, uint64_t count);

/* Number of uniforms cs_step_uniforms needs for a step range. */
int cs_uniforms_needed(const cs_context* ctx, uint64_t first_time_ind,
                       uint64_t nsteps, uint64_t* count);

int cs_synchronize(cs_context* ctx);

/* Per (domain, species) min, max and sum(u*dV) of the current fields,
 * arrays of num_domains*num_species, index d*num_species+s.  Max starts
 * from FLT_MIN and min from FLT_MAX like species.c:2808-2814; sums are
 * accumulated in double (Species_output_scalar_data, species.c:2785-2863). */
int cs_reduce_stats(cs_context* ctx, double* min_values, double* max_values,
                    double* total_values);

/* Weighted projection of one species on a reference axis (0 x, 1 y, 2 z)
 * for the linescan of Species_compute_convolution_with_gaussian_function_
 * locally (species.c:1960-2035): proj[k] = sum over the voxels whose
 * coordinate on `axis` is k and whose domain id d has bit d set in
 * domain_mask, of u * w_a * w_b, with w_a, w_b the weights of the two other
 * axes at the voxel (wx, wy, wz have n[0], n[1], n[2] entries; the one of
 * `axis` is not read and may be NULL).  The reference's 3-D Gaussian is
 * the product of one Gaussian per axis, so its linescan is the 1-D
 * convolution of this projection along the axis.  FP64 sums in a fixed
 * order; the field stays on the device. */
int cs_linescan(cs_context* ctx, uint32_t species, uint32_t axis, const double* wx,
                const double* wy, const double* wz, uint32_t domain_mask, double* proj);

/* cs_linescan that allocates nothing: no hipMalloc, hipFree or other
 * allocating call, for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid it).  No
 * counterpart in the reference.  The weights, the partial sums and the
 * projection live in a fixed device array of CS_LINESCAN_NOALLOC_MAX_DOUBLES
 * doubles (8 MiB) per GPU, held until the process ends; a call needs
 * nx+ny+nz, plus nz*ceil(ny/64)*nx (axis 0) or ny*nz (axes 1, 2), plus
 * n[axis] doubles, each rounded up to 32, and fails with CS_ERANGE beyond
 * that.  Every call on one GPU shares the array: calls that overlap on
 * one GPU (other host threads, other contexts' streams) race.  Results
 * equal cs_linescan's up to rounding of the final sums. */
#define CS_LINESCAN_NOALLOC_MAX_DOUBLES (1u << 20)
int cs_linescan_noalloc(cs_context* ctx, uint32_t species, uint32_t axis, const double* wx,
                        const double* wy, const double* wz, uint32_t domain_mask, double* proj);

/* Channel states of discrete flux f (0 closed, 1 open). */
int cs_get_channel_states(cs_context* ctx, uint32_t flux, uint8_t* states);
int cs_set_channel_states(cs_context* ctx, uint32_t flux, const uint8_t* states);

/* Channels of discrete flux f that were open at the start or after any
 * gating step since cs_create / cs_reset (states set by
 * cs_set_channel_states count too): the RyRs a spark recruited.  No
 * reference counterpart; the reference reports only the current states
 * (discrete_<name> datasets, species.c:3050-3080). */
int cs_get_channel_recruited(cs_context* ctx, uint32_t flux, uint8_t* recruited);

/* Time nsteps steps on the context stream with HIP events: total wall
 * time of the range and the mean duration of the stencil/reaction passes
 * (ms; a fused pass covers two steps).  Uses Philox gating. */
int cs_time_steps(cs_context* ctx, uint64_t first_time_ind, uint64_t nsteps,
                  double* total_ms, double* stencil_kernel_ms);

/* cs_time_steps that also reports the number of passes and how many of
 * them were fused two-step passes. */
int cs_time_passes(cs_context* ctx, uint64_t first_time_ind, uint64_t nsteps, double* total_ms,
                   double* pass_ms, uint64_t* passes, uint64_t* fused);

/* 1 when the problem admits fused two-step passes, else 0 with the
 * reason in why (if given).  They need at most 7 species, every boundary
 * flux acting on species 0, and a z-group of every species of a plane
 * small enough that the pass's 32-bit buffer offsets cover at least one
 * plane plus its four neighbour planes.  Up to 16384 flux rows (species,
 * voxel) and channels one k_bridge block gates and sums them between
 * passes; beyond that each of its phases is a launch of many blocks.
 * No reference counterpart. */
int cs_fused_available(const cs_context* ctx, char* why, size_t why_len);

/* Launch shape of the fused stencil/reaction kernel: blocks of 64 x rows
 * threads march zchunk planes; xcd_remap orders tiles over the 8 XCDs so
 * that neighbouring tiles share an L2 for their halo rows (0: hardware
 * order, 1: one contiguous slab per XCD, G >= 2: groups of G consecutive
 * tiles per XCD, round-robin); prefetch is
 * the number of planes loaded ahead (1 or 2; 1 with lds_tile);
 * nontemporal streams the next
 * state past the caches; lds_tile stages each plane in LDS so the x/y
 * neighbours are read from LDS instead of memory; linear hands each
 * wavefront 64 consecutive (x,y) columns of the plane across row ends, so
 * no lane idles when nx is not a multiple of 64; yblock > 0 selects the
 * row-blocked kernel, in which every lane owns yblock consecutive rows (2
 * or 4), a block is `rows` wavefronts side by side along x (1 to 8) and x
 * neighbours move between lanes by DPP.  fuse = 1 runs each pair of
 * consecutive steps whose flux species (and species 0, with channels)
 * diffuse on both steps as one two-step pass (k_step2): a block computes
 * the first step on its tile plus a one-voxel ring and the second on the
 * tile, so the intermediate state stays on chip and the fields cross HBM
 * once per two steps; fuse_zchunk is its planes per block and fuse_xcd
 * its tile order (as xcd_remap).  The other steps, and every step with
 * fuse = 0, run one at a time.  cs_create sets fuse = 1 for FP32 fields
 * when the problem admits fused passes and their launch has at least
 * half as many blocks as the device has CUs, else 0.  fuse_fixed = 1 (the default) runs the
 * pairs of step masks of the reference's five-slot species layout (Ca,
 * dye, CMDN-CSQN, ATP, TRPN) in FP32 with kernels compiled for those
 * masks; 0 runs every pass with the masks read at run time (same
 * results to FP32 rounding).  No reference counterpart: the reference's
 * OpenMP/AVX-512 loop has no launch shape (kernel_generic.cpp:561-600). */
typedef struct cs_launch_config {
  uint32_t rows, zchunk, xcd_remap, prefetch, nontemporal, lds_tile, linear, yblock;
  uint32_t fuse, fuse_zchunk, fuse_xcd, fuse_fixed;
} cs_launch_config;

int cs_get_launch_config(const cs_context* ctx, cs_launch_config* cfg);
int cs_set_launch_config(cs_context* ctx, const cs_launch_config* cfg);

/* Time every candidate launch shape on the context's grid and keep the
 * fastest.  The steps from first_time_ind over one period of the substep
 * pattern are grouped by (diffusing species, reactions); each group is
 * launched reps times per candidate and weighted by its share of the
 * period.  When the problem admits fused passes, the fusable step pairs
 * of two periods are timed the same way for each fused shape (without
 * gating), and fusing is kept when a pass beats two single steps.
 * Launches read the current state and write the spare buffer, so the
 * state is unchanged; best_ms is the winner's weighted mean kernel time
 * per step. */
int cs_autotune(cs_context* ctx, uint64_t first_time_ind, uint32_t reps,
                cs_launch_config* best, double* best_ms);

/* Copy `bytes` (rounded down to 16) between two scratch device buffers
 * reps times with a plain copy kernel (k_calibration_copy); used to scale
 * memory-traffic counters.  No reference counterpart. */
int cs_calibration_copy(uint64_t bytes, uint32_t reps, double* ms_per_copy);

/* ---- ensembles -----------------------------------------------------------
 *
 * An ensemble steps many independent realisations of one species model
 * together on the device.  The reference runs one realisation per
 * process: the seed is set once (calcium_sparks.c:113-116), the loop
 * (calcium_sparks.c:190-228) advances that one domain, and ensembles are
 * separate jobs (Run_scripts/G1/nonP/run_G1_nonP_0_sr900_ryr55.sh:61).
 * Every realisation keeps its own grid size, domain ids, boundary pairs,
 * initial channel states and Philox key (seed, instance); everything else
 * in its cs_problem (species, dt, tables, flux models, substeps,
 * precision, generator) must equal the first problem's.  A realisation
 * of an ensemble takes the same steps, in the same arithmetic, as a
 * cs_context built from its cs_problem.
 *
 * Per-realisation arrays are packed in realisation order: field transfers
 * and channel queries name the realisation; cs_batch_reduce_stats fills
 * [count][num_domains][num_species]. */
typedef struct cs_batch cs_batch;

int cs_batch_create(const cs_problem* problems, uint32_t count, cs_batch** out);
int cs_batch_destroy(cs_batch* b);
/* Fields back to rest, channels to their initial states, all realisations
 * active (Species_init_values per realisation, species.c:887-935). */
int cs_batch_reset(cs_batch* b);
/* Realisations, voxels over all of them, and device bytes of the fields
 * (two stacks) plus domain ids. */
int cs_batch_info(const cs_batch* b, uint32_t* count, uint64_t* voxels, uint64_t* field_bytes);
int cs_batch_upload_species(cs_batch* b, uint32_t realisation, uint32_t species, const double* src);
int cs_batch_download_species(cs_batch* b, uint32_t realisation, uint32_t species, double* dst);
/* cs_step / cs_step_uniforms for every active realisation.  In CS_RNG_MT
 * mode the uniforms of each gating step are the realisations' draws (as
 * cs_step_uniforms orders them) one realisation after the other. */
int cs_batch_step(cs_batch* b, uint64_t first_time_ind, uint64_t nsteps);
int cs_batch_step_uniforms(cs_batch* b, uint64_t first_time_ind, uint64_t nsteps,
                           const double* uniforms, uint64_t count);
int cs_batch_uniforms_needed(const cs_batch* b, uint64_t first_time_ind, uint64_t nsteps,
                             uint64_t* count);
int cs_batch_synchronize(cs_batch* b);
/* cs_reduce_stats of every realisation: [count][num_domains][num_species]. */
int cs_batch_reduce_stats(cs_batch* b, double* min_values, double* max_values, double* total_values);
/* cs_linescan of every realisation at once.  weights holds, per
 * realisation, its n[0] + n[1] + n[2] axis weights (x, then y, then z);
 * proj receives, per realisation, n[axis] values. */
int cs_batch_linescan(cs_batch* b, uint32_t species, uint32_t axis, const double* weights,
                      uint32_t domain_mask, double* proj);
int cs_batch_get_channel_states(cs_batch* b, uint32_t realisation, uint32_t flux, uint8_t* states);
int cs_batch_get_channel_recruited(cs_batch* b, uint32_t realisation, uint32_t flux, uint8_t* recruited);
/* Per realisation: channels open after its latest gating step, the time
 * index of the latest gating step that left a channel open (0: none
 * since cs_batch_reset), and channels recruited so far.  Any pointer may
 * be NULL.  This is what the -X rule (species.c:3384-3488) needs to end a
 * realisation's spark. */
int cs_batch_channel_summary(cs_batch* b, uint32_t* nopen, uint64_t* last_open_time_ind,
                             uint32_t* nrecruited);
/* Keep stepping only the realisations with active[i] != 0; the others
 * keep their state.  The stencil pass drops their tiles. */
int cs_batch_set_active(cs_batch* b, const uint8_t* active);

/* Where an ensemble's gating takes its channel states from:
 *   CS_GATE_OWN      gating as cs_step does (the default);
 *   CS_GATE_REPLAY   no gating: the states change only by
 *                    cs_batch_copy_channel_states (schedule replay);
 *   CS_GATE_CLOSURE  gating as cs_step does, on the concentration the
 *                    closure network predicts for each channel mouth
 *                    instead of the cleft voxel's (Philox only).
 * The network must outlive the ensemble's use of it.  No reference
 * counterpart: the reference always gates on the cleft voxel. */
enum { CS_GATE_OWN = 0, CS_GATE_REPLAY = 1, CS_GATE_CLOSURE = 2 };
struct cs_mlp;
int cs_batch_set_gating(cs_batch* b, uint32_t mode, const struct cs_mlp* net);
/* dst's channel states := src's (same realisations and channels). */
int cs_batch_copy_channel_states(cs_batch* dst, const cs_batch* src);
/* Channels over all realisations, in realisation order. */
int cs_batch_channel_count(const cs_batch* b, uint32_t* nchan);
/* The closure's raw inputs for every channel, [nchan][30]: absolute Ca
 * of the 3x3x3 voxels around its cleft voxel (index 9 (dz+1) + 3 (dy+1)
 * + dx+1, clamped at the faces), absolute Ca of its SR voxel, open
 * channels within Chebyshev distance 1 and 2 (itself included). */
int cs_batch_closure_features(cs_batch* b, float* features);
/* Absolute Ca at every channel's cleft voxel, [nchan]. */
int cs_batch_channel_ca(cs_batch* b, double* ca);
/* Absolute Ca at every channel's SR (from) voxel, [nchan]. */
int cs_batch_channel_sr_ca(cs_batch* b, double* ca);
/* State of every channel (1 open), [nchan], in channel order. */
int cs_batch_channel_open(cs_batch* b, uint8_t* open);
/* The closure's latest prediction for every channel mouth, [nchan]. */
int cs_batch_closure_predictions(cs_batch* b, double* ca);
/* Flux closure, used with CS_GATE_CLOSURE: a second network, on the same
 * inputs, predicts for every channel the ratio of the release the fine
 * grid makes through it to the release of the coarse grid's two-voxel
 * formula (ryr_flux, boundaryfluxes.c:87-141), and every channel's flux is
 * scaled by it.  Every step then predicts both networks from the state the
 * next step's fluxes and gating read.  NULL turns it off.  The network
 * must outlive the ensemble's use of it.  No reference counterpart. */
int cs_batch_set_flux_closure(cs_batch* b, const struct cs_mlp* net);
/* The flux closure's latest ratios, [nchan]. */
int cs_batch_flux_predictions(cs_batch* b, double* ratio);
/* Per realisation: channel evaluations whose normalised closure inputs
 * left the training range since cs_batch_reset (the range monitor). */
int cs_batch_monitor(cs_batch* b, uint32_t* flags);

/* Launch shape of the ensemble's stencil pass: blocks of 64 x rows
 * threads, each tile 64*rows consecutive plane columns of one realisation
 * by zchunk planes; prefetch and nontemporal as in cs_launch_config.
 * persistent = 1 runs one persistent block per resident slot, taking tiles
 * from eight per-XCD work queues; 0 launches one block per tile, the tiles
 * numbered in groups per XCD, and leaves the scheduling to the hardware
 * (same results, bit for bit).  No reference counterpart. */
typedef struct cs_batch_launch {
  uint32_t rows, zchunk, prefetch, nontemporal, persistent;
} cs_batch_launch;
int cs_batch_get_launch(const cs_batch* b, cs_batch_launch* cfg);
int cs_batch_set_launch(cs_batch* b, const cs_batch_launch* cfg);
/* Time every candidate shape on one period of the step pattern from
 * first_time_ind (state unchanged) and keep the fastest. */
int cs_batch_autotune(cs_batch* b, uint64_t first_time_ind, uint32_t reps, cs_batch_launch* best,
                      double* best_ms);
/* on != 0: cs_batch_step replays whole periods of the step pattern (the
 * least common multiple of every substep, doubled if odd) as captured HIP
 * graphs, with Philox gating.  No reference counterpart. */
int cs_batch_set_graph(cs_batch* b, uint32_t on);
/* Steps of one captured period (the least common multiple of every
 * substep, doubled if odd): step counts that are multiples of it replay
 * graphs only. */
int cs_batch_step_period(const cs_batch* b, uint64_t* period);
/* Time nsteps steps: total ms, and the mean stencil pass (0 with graphs). */
int cs_batch_time_steps(cs_batch* b, uint64_t first_time_ind, uint64_t nsteps, double* total_ms,
                        double* stencil_ms);

/* ---- learned closure ----------------------------------------------------
 *
 * A sub-grid closure for the coarse-grid method, with no reference
 * counterpart: the reference gates each channel on the concentration of
 * its cleft voxel (evaluate_RyR_stochastically, boundaryfluxes.c:295-359),
 * which at 12 nm dilutes the nanodomain a channel sees at 3 nm.  A
 * 30-64-64-1 perceptron with GELU units predicts the fine-grid calcium at
 * the channel mouth from coarse-grid features; it runs on the matrix
 * cores in bfloat16 with FP32 accumulation.
 *
 * Parameters are packed FP32 (CS_MLP_NUM_PARAMS): W1 [32][64] (row =
 * input; rows 30, 31 multiply zero padding), b1 [64], W2 [64][64] (row =
 * input), b2 [64], w3 [64], b3 [1].  Inputs x [n][30] are normalised
 * features; the output y is the normalised logarithm of the
 * concentration (cs_mlp_set_normalisation). */
#define CS_MLP_NUM_PARAMS 6337
typedef struct cs_mlp cs_mlp;   /* declared with the ensembles above */

int cs_mlp_create(const float* params, cs_mlp** out);
int cs_mlp_destroy(cs_mlp* m);
int cs_mlp_num_params(void);
int cs_mlp_get_params(cs_mlp* m, float* params);
int cs_mlp_set_params(cs_mlp* m, const float* params);
/* Feature normalisation the ensemble applies before the network: input k
 * is (f_k - in_mean[k]) / in_std[k] with f_k = ln(c) for the 28
 * concentrations and the count itself for the two open-channel counts;
 * the prediction is c = exp(y out_std + out_mean).  A normalised input
 * outside [in_lo - margin, in_hi + margin] trips the range monitor. */
int cs_mlp_set_normalisation(cs_mlp* m, const float* in_mean, const float* in_std, const float* in_lo,
                             const float* in_hi, float out_mean, float out_std, float margin);
/* y[n] for normalised inputs x[n][30] (host arrays). */
int cs_mlp_forward(cs_mlp* m, const float* x, uint64_t n, float* y);
/* Gradient of mean((y - t)^2) over n samples, and that mean. */
int cs_mlp_grad(cs_mlp* m, const float* x, const float* t, uint64_t n, float* grad, double* loss);
/* Training set (normalised inputs and targets), kept on the device;
 * minibatches walk a shuffled order drawn from shuffle_seed, reshuffled
 * every epoch. */
int cs_mlp_set_data(cs_mlp* m, const float* x, const float* t, uint64_t n, uint64_t shuffle_seed);
/* `steps` Adam steps (beta 0.9, 0.999, eps 1e-8) of `batch` samples at
 * learning rate lr; last_loss is the mean squared error of the last batch
 * before its update. */
int cs_mlp_train(cs_mlp* m, uint32_t steps, uint32_t batch, float lr, double* last_loss);

/* Scalar flux functions of the plugin table, evaluated on the host with
 * the same code the device runs (ryr_flux boundaryfluxes.c:87-141,
 * serca_flux :12-85). */
double cs_ryr_flux(double dt, double h, double u0, double u1, const double* params);
double cs_serca_flux(double dt, double h, double u0, double u1, const double* params);

/* Philox4x32-10 block (Salmon et al., SC'11), exposed for known-answer tests. */
void cs_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

const char* cs_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* CSPARKS_H */
