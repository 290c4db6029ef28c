/*
 * cs_model.h -- in-memory form of the calcium_sparks inputs.
 *
 * The three structs below hold exactly what the reference reads from its
 * two HDF5 input files and its command line, so that every consumer (the
 * problem builder, the CPU oracle, the HDF5 reader/writer and the Python
 * bindings) shares one plain-C description of a run:
 *
 *   cs_geometry       <- geometry file   (reference geometry.c:92-141,
 *                                          :487-574, :576-1011)
 *   cs_species_model  <- species file    (reference species.c:103-258,
 *                                          :346-362, :780;
 *                                          boundaryfluxes.c:361-542)
 *   cs_options        <- argp options    (reference options.c:13-393,
 *                                          defaults options.c:502-542)
 *
 * Layout conventions follow the reference: voxel arrays are [N0][N1][N2]
 * with the last (z) index fastest, boundary pair lists are uint16 [n][6]
 * = (x0,y0,z0,x1,y1,z1) in geometry voxels.  Concentrations are in uM,
 * time in ms, lengths in nm, diffusion constants ("sigma") in nm^2/ms.
 */
#ifndef CS_MODEL_H
#define CS_MODEL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CS_MAX_DOMAINS    5   /* reference max_num_domains, species.c:16   */
#define CS_MAX_SPECIES    8
#define CS_MAX_BOUNDARIES 8
#define CS_MAX_NAMED      8   /* entries of -O / -C lists                  */
#define CS_MAX_OPEN      64   /* indices per -O entry                      */
#define CS_NAME_LEN      32

/* Boundary types, reference geometry.h boundary_t */
enum { CS_BOUNDARY_DENSITY = 0, CS_BOUNDARY_DISCRETE = 1 };

typedef struct cs_geometry {
  uint64_t N[3];                 /* geometry voxels per axis               */
  double   h;                    /* geometry resolution, attr /h (nm)      */
  double   global_size[3];       /* attr /global_size                      */
  uint32_t num_domains;
  char     domain_names[CS_MAX_DOMAINS][CS_NAME_LEN];
  uint8_t  domain_indices[CS_MAX_DOMAINS];
  uint8_t  domain_connections[CS_MAX_DOMAINS * CS_MAX_DOMAINS]; /* [nd][nd] */
  const uint8_t* voxels;         /* [N0][N1][N2] domain ids, z fastest     */
  uint32_t num_boundaries;
  char     boundary_names[CS_MAX_BOUNDARIES][CS_NAME_LEN];
  uint32_t boundary_types[CS_MAX_BOUNDARIES];
  uint64_t boundary_sizes[CS_MAX_BOUNDARIES];
  const uint16_t* boundaries[CS_MAX_BOUNDARIES]; /* [size][6]              */
} cs_geometry;

/* RyR / pRyR parameters (species file groups /ryr and /p_ryr). */
typedef struct cs_channel_model {
  double Kd_open, k_min_open, k_max_open, n_open;
  double Kd_close, k_min_close, k_max_close, n_close;
  double i_unitary;              /* pA                                     */
  char   boundary[CS_NAME_LEN];
  char   species[CS_NAME_LEN];
} cs_channel_model;

/* SERCA parameters (species file group /serca). */
typedef struct cs_serca_model {
  double v_cyt_to_A_sr, density, scale;
  char   boundary[CS_NAME_LEN];
  char   species[CS_NAME_LEN];   /* the reference reads this from /ryr     */
} cs_serca_model;

typedef struct cs_species_model {
  uint32_t num_domains, num_species;
  char     domain_names[CS_MAX_DOMAINS][CS_NAME_LEN];
  char     species_names[CS_MAX_SPECIES][CS_NAME_LEN];
  /* per domain d: */
  uint32_t num_diffusive[CS_MAX_DOMAINS];
  uint8_t  diffusive[CS_MAX_DOMAINS][CS_MAX_SPECIES];
  double   sigma[CS_MAX_DOMAINS][CS_MAX_SPECIES];   /* indexed like diffusive */
  double   init[CS_MAX_DOMAINS][CS_MAX_SPECIES];
  uint32_t num_buffers[CS_MAX_DOMAINS];
  double   k_on[CS_MAX_DOMAINS][CS_MAX_SPECIES];    /* [d][buffer]          */
  double   k_off[CS_MAX_DOMAINS][CS_MAX_SPECIES];
  double   tot[CS_MAX_DOMAINS][CS_MAX_SPECIES];
  uint8_t  buffer_species[CS_MAX_DOMAINS][CS_MAX_SPECIES][2]; /* (bound, free) */
  uint32_t use_ryr, use_p_ryr, use_serca;
  cs_channel_model ryr, p_ryr;
  cs_serca_model   serca;
  double   conv[3];              /* /convolution_constants                 */
  uint32_t num_fixed;            /* /num_fixed_domain_species              */
  uint8_t  fixed[2 * CS_MAX_DOMAINS * CS_MAX_SPECIES]; /* (domain, species) pairs */
} cs_species_model;

typedef struct cs_open_entry {
  char     name[CS_NAME_LEN];    /* boundary name given to -O              */
  uint32_t count;
  int32_t  idx[CS_MAX_OPEN];
} cs_open_entry;

typedef struct cs_close_entry {
  char   name[CS_NAME_LEN];      /* boundary name given to -C              */
  double time;
} cs_close_entry;

/* Precision of the stored fields and stencil arithmetic. */
enum { CS_PREC_F64 = 0, CS_PREC_F32 = 1 };

/* Source of the gating uniforms.
 *  CS_RNG_PHILOX : counter-based Philox4x32-10 keyed by (seed, instance),
 *                  counter (gating step, channel); used on the device.
 *  CS_RNG_MT     : the reference's draw order from one Mersenne Twister
 *                  (MT19937, 32-bit seeding) consumed channel by channel;
 *                  draws are produced on the host and streamed to the
 *                  device. */
enum { CS_RNG_PHILOX = 0, CS_RNG_MT = 1 };

typedef struct cs_options {
  double   tstop;                /* -t, default 10 ms                      */
  double   h;                    /* -h, default 6 nm                       */
  double   dt;                   /* -d, <=0: from the stability limit      */
  double   dt_save;              /* -o, default 0.5 ms                     */
  double   dt_update_react;      /* -T, default -1                         */
  double   dt_update_stoch;      /* -D, default -1                         */
  double   tabort;               /* -X, 0 = inactive                       */
  int32_t  force_dirichlet;      /* -f                                     */
  int32_t  seed_given;           /* -S given                               */
  uint32_t seed;
  uint32_t num_open;
  cs_open_entry  open[CS_MAX_NAMED];     /* -O name idx.. name idx..       */
  uint32_t num_t_close;
  cs_close_entry t_close[CS_MAX_NAMED];  /* -C name T ..                   */
  uint32_t precision;            /* CS_PREC_*                              */
  uint32_t rng;                  /* CS_RNG_*                               */
} cs_options;

/* Fill *opt with the reference defaults (options.c:502-542); tabort gets a
 * defined default of 0 (inactive) where the reference leaves it unset. */
void cs_options_defaults(cs_options* opt);

#ifdef __cplusplus
}
#endif
#endif /* CS_MODEL_H */
