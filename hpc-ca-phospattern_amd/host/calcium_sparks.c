/*
 * calcium_sparks -- drop-in command-line program for the MI355X path.
 *
 * Same options, parsing rules, input files, output file and stdout tables
 * as the reference program (calcium_sparks.c:19-45, options.c:13-601,
 * species.c:2253-2551, :2785-3092); the time loop runs on the GPU through
 * the cs_* C ABI, synchronising with the host only at outputs.  One GPU
 * holds the whole domain, so -p / --mpi_dist are accepted and checked
 * against a single process.
 *
 * Extra options (long only): --precision f64|f32, --rng philox|mt,
 * --autotune (time the kernel launch shapes before the run).
 */
#include <argp.h>
#include <float.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "cs_host.h"
#include "cs_io.h"
#include "csparks.h"

enum { OPT_ALL_DATA = 777, OPT_MPI_DIST = 778, OPT_PRECISION = 779, OPT_RNG = 780, OPT_AUTOTUNE = 781 };

typedef struct {
  int axis;
  char* species;
  int num_domains;
  char** domains;
  int offsets[3];
} linescan_arg;

typedef struct {
  cs_options opt;
  char* casename;
  char* geometry_file;
  char* model_file;
  int silent, verbose, autotune;
  char** save_species;
  int num_save_species;
  char** all_data_species;
  int all_data_num_species;
  int num_ax_points[3];
  double* ax_points[3];
  linescan_arg* linescan;
  int split_plane;
  int mpi_dist[3], mpi_dist_given;
} cli_args;

static void die(const char* fmt, ...) __attribute__((noreturn, format(printf, 1, 2)));
static void die(const char* fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vfprintf(stderr, fmt, ap);
  va_end(ap);
  exit(1);
}

const char* argp_program_version = "calcium_spark 0.1 (MI355X)";
static char args_doc[] = "calcium_spark -- A finite volume program modeling a calcium spark simulation";

static struct argp_option options[] = {
  {"casename",        'c', "CASENAME", 0, "Sets the name of the run to CASENAME. Default (\"casename\")", 0},
  {"dt",              'd', "TIME", 0, "Forced minimal timestep. If negative time step will be deducted from diffusion constants. Default(-1.)", 0},
  {"force_dirichlet", 'f', 0, 0, "Force to use Dirichlet boundary condition.", 0},
  {"geometry_file",   'g', "FILE", 0, "Read the geometry from FILE. Exprects a .h5 file.", 0},
  {"resolution",      'h', "RES", 0, "Run simulation with RES nm as the spatial resolution. Default(6.)", 0},
  {"species_file",    'm', "FILE", 0, "Read the species model from FILE. Expects a .h5 file.", 0},
  {"dt_save",         'o', "TIME", 0, "Save solution for each dt_save. Default(.5)", 0},
  {"split_processes", 'p', "PLANE", 0, "Split processes in given directions {XY, YZ, XZ}. Default(XY).", 0},
  {"quiet",           'q', 0, 0, "Don't produce any output", 0},
  {"save_species",    's', "SPECIES", 0, "A list of species that should be saved in 2D sheets at given z-coordinates.", 0},
  {"tstop",           't', "TIME", 0, "Run simulation for TIME number of ms. Default(10.)", 0},
  {"verbose",         'v', 0, 0, "Produce verbose output", 0},
  {"save_x_coords",   'x', "COORD", 0, "X coordinates for the 2D sheet that will be saved.", 0},
  {"save_y_coords",   'y', "COORD", 0, "Y coordinates for the 2D sheet that will be saved.", 0},
  {"save_z_coords",   'z', "COORD", 0, "Z coordinates for the 2D sheet that will be saved.", 0},
  {"linescan",        'l', "DIR_SPCECIES_COORD", 0, "Computes convolution of the species with 3-D Gaussian function.", 0},
  {"t_close",         'C', "TIME", 0, "Time for closing currents. If negative, they are closed stochastically. Default (-1)", 0},
  {"dt_update_stoch", 'D', "TIME", 0, "Time step for updating the stochastics. If negative it will be the same as global timestep. Default (-1)", 0},
  {"open",            'O', "OPEN_STATES", 0, "Initialize given discrete fluxes to open. Otherwise they are randomly initialized.", 0},
  {"seed",            'S', "SEED", 0, "Provide a seed for the random generator. If not given the seed will be random each run.", 0},
  {"dt_update_react", 'T', "TIME", 0, "Time step for updating the reactions. If negative it will be the same as global timestep. Default (-1)", 0},
  {"abort",           'X', "TYPE TIME", 0, "Abort simulations after TIME ms after the last RyR channel has been closed. By default it is inactive.", 0},
  {"all_data",        OPT_ALL_DATA, "SPECIES", 0, "Save species values of all points (3D)", 0},
  {"mpi_dist",        OPT_MPI_DIST, "MPI_DIST", 0, "An mpi distribution for processor splitting in XYZ directions. It overwrites the -p option. By default it is switched off.", 0},
  {"precision",       OPT_PRECISION, "f64|f32", 0, "Field storage and stencil precision on the GPU. Default f64.", 0},
  {"rng",             OPT_RNG, "philox|mt", 0, "Gating uniforms: counter-based Philox (default) or the reference's Mersenne Twister stream.", 0},
  {"autotune",        OPT_AUTOTUNE, 0, 0, "Time the kernel launch shapes on this grid before the run.", 0},
  {0}
};

static int is_nonnumeric(const char* s) { return atoi(s) == 0 && s[0] != '0'; }

/* parse_opt follows options.c:13-393, including its multi-token options */
static error_t parse_opt(int key, char* arg, struct argp_state* state)
{
  cli_args* a = (cli_args*)state->input;
  int offset, i, index, num_args;
  switch (key) {
    case 'c': a->casename = arg; break;
    case 'd': a->opt.dt = atof(arg); break;
    case 'f': a->opt.force_dirichlet = 1; break;
    case 'g': a->geometry_file = arg; break;
    case 'h': a->opt.h = atof(arg); break;
    case 'm': a->model_file = arg; break;
    case 'o': a->opt.dt_save = atof(arg); break;
    case 'p':
      if (!strcmp(arg, "XY")) a->split_plane = 0;
      else if (!strcmp(arg, "YZ")) a->split_plane = 1;
      else if (!strcmp(arg, "XZ")) a->split_plane = 2;
      else if (!strcmp(arg, "XYZ")) a->split_plane = 3;
      else die("***Error: Expected argument for option -p is one of {XY, XZ, YZ}. Instead got %s\n", arg);
      break;
    case 'q': a->silent = 1; break;
    case 's':
      offset = 0;
      while (state->argv[state->next + offset] != NULL && state->argv[state->next + offset][0] != '-') offset++;
      a->save_species = &state->argv[state->next - 1];
      a->num_save_species = offset + 1;
      state->next += offset;
      break;
    case 't': a->opt.tstop = atof(arg); break;
    case 'v': a->verbose = 1; break;
    case 'x': case 'y': case 'z': {
      int ax = key == 'x' ? 0 : key == 'y' ? 1 : 2;
      offset = 0;
      while (state->argv[state->next + offset] != NULL && state->argv[state->next + offset][0] != '-') offset++;
      a->num_ax_points[ax] = offset + 1;
      free(a->ax_points[ax]);
      a->ax_points[ax] = (double*)malloc(sizeof(double) * a->num_ax_points[ax]);
      for (i = 0; i < a->num_ax_points[ax]; i++) a->ax_points[ax][i] = atof(state->argv[state->next - 1 + i]);
      state->next += offset;
      break;
    }
    case 'l': {
      /* options.c:154-208; the digit test there accepts every '-' token
       * whose second character is >= '0' or <= '9', i.e. any, so -l takes
       * all remaining tokens */
      offset = 0;
      while (state->argv[state->next + offset] != NULL &&
             (state->argv[state->next + offset][0] != '-' ||
              (state->argv[state->next + offset][0] == '-' &&
               (state->argv[state->next + offset][1] >= '0' || state->argv[state->next + offset][1] <= '9'))))
        offset++;
      num_args = offset + 1;
      if (num_args < 4)
        die("*** Error: Incorrect number of provided arguments for --linescan option. Expects at least 4, "
            "butinstead found %d.\n", num_args);
      char** args = &state->argv[state->next - 1];
      if (strlen(args[0]) != 1 || (args[0][0] != 'x' && args[0][0] != 'y' && args[0][0] != 'z'))
        die("*** Error: Expects axis x,y or z as the first argument passed to --linescan (-l). Instead got: %s\n", args[0]);
      linescan_arg* l = (linescan_arg*)calloc(1, sizeof *l);
      l->axis = args[0][0] - 'x';
      l->species = args[1];
      l->num_domains = num_args - 4;
      l->domains = l->num_domains ? &args[2] : NULL;
      const int dir = l->axis;
      const int sgn = dir % 2 == 0 ? 1 : -1;
      l->offsets[((dir + sgn) % 3 + 3) % 3] = atoi(args[2 + l->num_domains]);
      l->offsets[((dir + 2 * sgn) % 3 + 3) % 3] = atoi(args[3 + l->num_domains]);
      a->linescan = l;
      state->next += offset;
      break;
    }
    case 'C': {
      offset = 0;
      num_args = 1;
      while (state->argv[state->next + offset] != NULL && state->argv[state->next + offset][0] != '-') offset++;
      num_args += offset;
      if (num_args < 2 && num_args % 2 != 0)
        die("*** Error: Incorrect number of provided arguments for --t_close option. Expects an even positive "
            "number. Instead found %d provided arguments.\n", num_args);
      a->opt.num_t_close = (uint32_t)(num_args / 2);
      if (a->opt.num_t_close > CS_MAX_NAMED) die("*** Error: too many --t_close entries\n");
      for (index = 0; index <= offset; index++) {
        const char* tok = state->argv[state->next + index - 1];
        if (index % 2 == 0) {
          if (!is_nonnumeric(state->argv[state->next - 1]))
            die("Error: Expected a string on the  %d arguments for --t_close option. Instead got: %s.\n", index + 1, tok);
          snprintf(a->opt.t_close[index / 2].name, CS_NAME_LEN, "%s", tok);
        } else {
          double r = atof(tok);
          r += strcmp(tok, "0") == 0;
          r += strcmp(tok, "0.0") == 0;
          if (!r) die("Error: Expected a REAL number on the %d arguments for --t_close option. Instead got: %s.\n", index + 1, tok);
          a->opt.t_close[index / 2].time = atof(tok);
        }
      }
      state->next += offset;
      break;
    }
    case 'D': a->opt.dt_update_stoch = atof(arg); break;
    case 'O': {
      int current = -1, num_states = 0, read_number;
      offset = 0;
      if (!is_nonnumeric(state->argv[state->next - 1]))
        die("*** Error: Expected a string as the first argument for --open option. Instead got: %s.\n", state->argv[state->next - 1]);
      int length = 0;
      while (state->argv[state->next + offset] != NULL && state->argv[state->next + offset][0] != '-') {
        if (is_nonnumeric(state->argv[state->next + offset])) length++;
        offset++;
      }
      length++;
      if (offset == 0 || (atoi(state->argv[state->next + offset - 1]) == 0 && strcmp(state->argv[state->next + offset - 1], "0") != 0))
        die("*** Error: Expected an integer as the last argument for --force_discrete_boundaries_open option.Instead got: %s.\n",
            state->argv[state->next + offset - 1]);
      if (length > CS_MAX_NAMED) die("*** Error: too many --open entries\n");
      a->opt.num_open = (uint32_t)length;
      for (index = 0; index <= offset; index++) {
        if (state->argv[state->next + index] == NULL) read_number = 0;
        else {
          read_number = atoi(state->argv[state->next + index]);
          read_number += strcmp(state->argv[state->next + index], "0") == 0;
        }
        if (read_number) num_states++;
        else {
          cs_open_entry* e = &a->opt.open[++current];
          snprintf(e->name, CS_NAME_LEN, "%s", state->argv[state->next + index - num_states - 1]);
          if (num_states > CS_MAX_OPEN) die("*** Error: too many open indices\n");
          e->count = (uint32_t)num_states;
          for (int id = 0; id < num_states; id++) {
            e->idx[id] = atoi(state->argv[state->next + index - num_states + id]);
            if (e->idx[id] < 0)
              die("***Error: Expected a positive number for the open indices given by --force_discrete_boundaries_open option. For %s got %d.\n",
                  e->name, e->idx[id]);
            if (id > 0 && e->idx[id] <= e->idx[id - 1])
              die("***Error: Expected a monotonicly increasing numbers for the open states indices given by "
                  "--force_discrete_boundaries_open option. For %s got %d and %d.\n", e->name, e->idx[id - 1], e->idx[id]);
          }
          num_states = 0;
        }
      }
      state->next += offset;
      break;
    }
    case 'S': a->opt.seed = (uint32_t)atoi(arg); a->opt.seed_given = 1; break;
    case 'T': a->opt.dt_update_react = atof(arg); break;
    case 'X': a->opt.tabort = atof(arg); break;
    case OPT_ALL_DATA:
      offset = 0;
      while (state->argv[state->next + offset] != NULL && state->argv[state->next + offset][0] != '-') offset++;
      a->all_data_species = &state->argv[state->next - 1];
      a->all_data_num_species = offset + 1;
      state->next += offset;
      break;
    case OPT_MPI_DIST:
      for (offset = 0; offset < 3; offset++) {
        if (state->argv[state->next - 1 + offset] == NULL || state->argv[state->next - 1 + offset][0] == '-')
          die("*** Error: Excpected three positive integers after the --mpi_dist option. Found only %d\n", offset);
        a->mpi_dist[offset] = atoi(state->argv[state->next - 1 + offset]);
        if (a->mpi_dist[offset] < 1)
          die("*** Error: Encountered a wrong number at position %d for the --mpi_dist option. Expected an integer > 0.\n", offset + 1);
      }
      a->mpi_dist_given = 1;
      state->next += 2;
      break;
    case OPT_PRECISION:
      if (!strcmp(arg, "f64")) a->opt.precision = CS_PREC_F64;
      else if (!strcmp(arg, "f32")) a->opt.precision = CS_PREC_F32;
      else die("*** Error: --precision expects f64 or f32, got %s\n", arg);
      break;
    case OPT_RNG:
      if (!strcmp(arg, "philox")) a->opt.rng = CS_RNG_PHILOX;
      else if (!strcmp(arg, "mt")) a->opt.rng = CS_RNG_MT;
      else die("*** Error: --rng expects philox or mt, got %s\n", arg);
      break;
    case OPT_AUTOTUNE: a->autotune = 1; break;
    case ARGP_KEY_ARG: argp_usage(state); break;
    default: return ARGP_ERR_UNKNOWN;
  }
  return 0;
}

static struct argp argp = {options, parse_opt, 0, args_doc, 0, 0, 0};

/* arguments_output, options.c:395-500 */
static void print_arguments(const cli_args* a)
{
  const char coord[3] = {'x', 'y', 'z'};
  printf("\nArguments:\n");
  printf("-----------------------------------------------------------------------------\n");
  printf("  casename:          \"%s\"\n", a->casename);
  printf("  geometry_file:     \"%s\"\n", a->geometry_file);
  printf("  species_file:      \"%s\"\n", a->model_file);
  printf("  tstop:             %.6f ms\n", a->opt.tstop);
  printf("  h:                 %.1f nm\n", a->opt.h);
  printf("  dt:                %.3g\n", a->opt.dt);
  printf("  dt_save:           %.3g\n", a->opt.dt_save);
  printf("  dt_update_stoch:   %.3g\n", a->opt.dt_update_stoch);
  printf("  dt_update_react:   %.3g\n", a->opt.dt_update_react);
  printf("  t_close:           %s", a->opt.num_t_close ? "" : "-");
  for (uint32_t i = 0; i < a->opt.num_t_close; i++) {
    printf("%s %.3g", a->opt.t_close[i].name, a->opt.t_close[i].time);
    if (i < a->opt.num_t_close - 1) printf(", ");
  }
  printf("\n");
  printf("  linescan:          %s", a->linescan ? "" : "-\n");
  if (a->linescan) {
    const int dir = a->linescan->axis, sgn = dir % 2 == 0 ? 1 : -1;
    const int o1 = ((dir + sgn) % 3 + 3) % 3, o2 = ((dir + 2 * sgn) % 3 + 3) % 3;
    printf("%c, %s, %d, %d\n", coord[dir], a->linescan->species, a->linescan->offsets[o1], a->linescan->offsets[o2]);
  }
  printf("  abort simulations:%s", a->opt.tabort != 0 ? "" : "-\n");
  if (a->opt.tabort != 0) printf("%.3g\n", a->opt.tabort);
  printf("  save_2D_species:   %s", a->num_save_species ? "" : "-");
  for (int i = 0; i < a->num_save_species; i++) {
    printf("%s", a->save_species[i]);
    if (i < a->num_save_species - 1) printf(", ");
  }
  printf("\n");
  if (a->num_save_species) {
    if (a->num_ax_points[0] + a->num_ax_points[1] + a->num_ax_points[2] == 0) printf("   z-coords (nm):   center\n");
    else
      for (int ax = 0; ax < 3; ax++)
        if (a->num_ax_points[ax]) {
          printf("    %c-coords (nm):   ", coord[ax]);
          for (int i = 0; i < a->num_ax_points[ax]; i++) {
            printf("%.2f", a->ax_points[ax][i]);
            if (i < a->num_ax_points[ax] - 1) printf(", ");
          }
          printf("\n");
        }
  }
  printf("  save_all_data:     %s", a->all_data_num_species ? "" : "-");
  for (int i = 0; i < a->all_data_num_species; i++) {
    printf("%s", a->all_data_species[i]);
    if (i < a->all_data_num_species - 1) printf(", ");
  }
  printf("\n");
  printf("  open states:       %s", a->opt.num_open ? "" : "-");
  for (uint32_t i = 0; i < a->opt.num_open; i++) {
    if (i > 0) printf("                   ");
    printf("%s: ", a->opt.open[i].name);
    for (uint32_t j = 0; j < a->opt.open[i].count; j++) {
      printf("%d", a->opt.open[i].idx[j]);
      if (j < a->opt.open[i].count - 1) printf(", ");
    }
    printf("\n");
  }
  printf("\n\n");
}

static double now(void)
{
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return t.tv_sec + t.tv_nsec * 1e-9;
}

static int species_index(const cs_setup* s, const char* name)
{
  for (uint32_t i = 0; i < s->num_species; i++)
    if (!strcmp(s->species_names[i], name)) return (int)i;
  return -1;
}

static void all_species_list(const cs_setup* s, char* buf, size_t len)
{
  buf[0] = 0;
  for (uint32_t i = 0; i < s->num_species; i++) {
    size_t l = strlen(buf);
    snprintf(buf + l, len - l, i == 0 ? "\"%s\"" : ", \"%s\"", s->species_names[i]);
  }
}

/* gauss, utils.c:209-213 */
static double gauss(double x, double sigma)
{
  const double inv_sqrt = 1.0 / sqrt(sigma);
  return M_SQRT1_2 * M_2_SQRTPI * inv_sqrt / 2 * exp(-x * x / 2 / sigma);
}

typedef struct {
  cs_setup* setup;
  cs_context* ctx;
  cli_args* args;
  const cs_geometry* geom;
  const cs_species_model* model;
  int* save_idx;
  int* all_idx;
  int sheets_per_axis[3];
  int* global_indices[3];
  int linescan_species;
  int linescan_domain_ids[CS_MAX_DOMAINS];
  double* field;            /* host copy of one species, reference layout */
  int32_t* states;
  uint8_t* states8;
  int* snapshot_closed;     /* channel snapshot for the abort rule         */
  double t_write;
} run_ctx;

static void cs_try(int rc, const char* what)
{
  if (rc != CS_OK) die("*** ERROR: %s failed: %s\n", what, cs_last_error());
}

/* Species_output_data (species.c:2253-2343) + the tables of
 * Species_output_scalar_data (species.c:2945-3025) */
static int write_output(run_ctx* R, uint32_t output_ind, double t, uint64_t time_ind, cs_abort_state* ab)
{
  double t0 = now();
  cs_setup* S = R->setup;
  cli_args* a = R->args;
  const uint32_t nd = S->num_domains, ns = S->num_species;
  double mn[CS_MAX_DOMAINS * CS_MAX_SPECIES], mx[CS_MAX_DOMAINS * CS_MAX_SPECIES], tot[CS_MAX_DOMAINS * CS_MAX_SPECIES];
  double avg[CS_MAX_DOMAINS * CS_MAX_SPECIES];
  cs_try(cs_reduce_stats(R->ctx, mn, mx, tot), "cs_reduce_stats");
  for (uint32_t i = 0; i < nd * ns; i++) avg[i] = tot[i] / S->volumes[i / ns];
  char fname[512], err[512];
  snprintf(fname, sizeof fname, "%s.h5", a->casename);
  cs_h5out* o = cs_out_open(fname, output_ind, t, time_ind, S->problem.h, S->N, a->geometry_file, a->model_file, err, sizeof err);
  if (!o) die("*** ERROR: %s\n", err);
  if (cs_out_scalars(o, nd, ns, S->domain_names, S->species_names, mn, mx, avg)) die("*** ERROR: writing scalars failed\n");
  /* channel states in the order of the discrete boundaries */
  int all_closed_ryr = 1, all_closed_pryr = 1;
  for (uint32_t k = 0; k < S->num_discrete; k++) {
    const cs_discrete_info* di = &S->discrete[k];
    memset(R->states, 0, sizeof(int32_t) * (di->size ? di->size : 1));
    if (di->flux >= 0) {
      cs_try(cs_get_channel_states(R->ctx, (uint32_t)di->flux, R->states8), "cs_get_channel_states");
      for (uint32_t i = 0; i < di->size; i++) R->states[i] = R->states8[i];
    }
    int open = 0;
    for (uint32_t i = 0; i < di->size; i++) open += R->states[i];
    if ((int)k == S->ryr_discrete && open) all_closed_ryr = 0;
    if ((int)k == S->p_ryr_discrete && open) all_closed_pryr = 0;
    if (cs_out_discrete(o, di->name, R->states, di->size)) die("*** ERROR: writing discrete states failed\n");
    R->snapshot_closed[k] = open == 0;
  }
  /* Species_check_open_discrete_boundaries looks at ryr and/or p_ryr */
  int all_closed = 0;
  if (S->has_ryr_boundary && !S->has_p_ryr_boundary) all_closed = all_closed_ryr;
  else if (!S->has_ryr_boundary && S->has_p_ryr_boundary) all_closed = all_closed_pryr;
  else if (S->has_ryr_boundary && S->has_p_ryr_boundary) all_closed = all_closed_ryr && all_closed_pryr;
  cs_abort_snapshot(ab, all_closed);

  if (!a->silent) {
    double conservation = 0.0;
    for (uint32_t i = 0; i < nd * ns; i++) conservation += tot[i];
    printf("\n%u: time: %.2f ms, time_index: %lu, conservation: %g\n", output_ind, t, (unsigned long)time_ind, conservation);
    printf("-----------------------------------------------------------------------------\n");
    const char* titles[3] = {"Average:", "Min:", "Max:"};
    const double* vals[3] = {avg, mn, mx};
    for (int k = 0; k < 3; k++) {
      printf("%8s", titles[k]);
      for (uint32_t s = 0; s < ns; s++) printf("%13s", S->species_names[s]);
      printf("\n");
      for (uint32_t d = 0; d < nd; d++) {
        printf("%7s:", S->domain_names[d]);
        for (uint32_t s = 0; s < ns; s++) printf("%13.2e", vals[k][d * ns + s]);
        printf("\n");
      }
      printf("\n");
    }
    if (S->num_discrete > 0) {
      printf("%8s", "Channels:\n");
      for (uint32_t k = 0; k < S->num_discrete; k++) {
        const cs_discrete_info* di = &S->discrete[k];
        memset(R->states, 0, sizeof(int32_t) * (di->size ? di->size : 1));
        if (di->flux >= 0) {
          cs_try(cs_get_channel_states(R->ctx, (uint32_t)di->flux, R->states8), "cs_get_channel_states");
          for (uint32_t i = 0; i < di->size; i++) R->states[i] = R->states8[i];
        }
        printf("%7s: ", di->name);
        int sum = 0;
        /* the reference compares with the size of boundary k, not of the
         * k-th discrete boundary (species.c:3014) */
        const long wrap = (long)R->geom->boundary_sizes[k] - 1;
        for (uint32_t bi = 0; bi < di->size; bi++) {
          sum += R->states[bi];
          printf("%u:%d", bi, R->states[bi]);
          if ((long)bi < wrap) {
            printf(", ");
            if (bi % 14 == 0 && bi != 0) printf("\n         ");
          }
        }
        printf(" | %d/%d\n", sum, (int)di->size);
      }
      printf("\n");
    }
  }

  /* sheets (species.c:2346-2484) */
  const uint64_t* N = S->N;
  if (a->num_save_species) {
    for (int i = 0; i < a->num_save_species; i++) {
      int si = R->save_idx[i];
      cs_try(cs_download_species(R->ctx, (uint32_t)si, R->field), "cs_download_species");
      for (int ax = 0; ax < 3; ax++) {
        if (!R->sheets_per_axis[ax]) continue;
        const int lx = (ax + 1) % 3, ly = (ax + 2) % 3;
        const uint64_t d0 = N[lx], d1 = N[ly];
        double* sheet = (double*)malloc(sizeof(double) * d0 * d1);
        for (int k = 0; k < R->sheets_per_axis[ax]; k++) {
          const uint64_t idx = (uint64_t)R->global_indices[ax][k];
          for (uint64_t p = 0; p < d0; p++)
            for (uint64_t q = 0; q < d1; q++) {
              uint64_t c[3];
              c[ax] = idx;
              c[lx] = p;
              c[ly] = q;
              sheet[p * d1 + q] = R->field[(c[0] * N[1] + c[1]) * N[2] + c[2]];
            }
          char dsn[128];
          const char axn[3] = {'x', 'y', 'z'};
          snprintf(dsn, sizeof dsn, "%s_%c_sheet_%d", S->species_names[si], axn[ax], R->global_indices[ax][k]);
          uint64_t dims[2] = {d0, d1};
          if (cs_out_dataset(o, dsn, sheet, 2, dims, 0)) die("*** ERROR: writing sheet failed\n");
        }
        free(sheet);
      }
    }
  }
  /* all_data (species.c:2487-2551) */
  for (int i = 0; i < a->all_data_num_species; i++) {
    int si = R->all_idx[i];
    cs_try(cs_download_species(R->ctx, (uint32_t)si, R->field), "cs_download_species");
    char dsn[128];
    snprintf(dsn, sizeof dsn, "%s_all_data", S->species_names[si]);
    if (cs_out_dataset(o, dsn, R->field, 3, N, 1)) die("*** ERROR: writing all_data failed\n");
  }
  /* linescan (species.c:1960-2102): the 3-D Gaussian factorises, so the
   * field is first projected on the scan axis with the two transverse
   * Gaussians (on the device), then convolved along the axis here */
  if (a->linescan) {
    linescan_arg* l = a->linescan;
    const double h = S->problem.h;
    const int dir = l->axis;
    double off[3];
    for (int ax = 0; ax < 3; ax++) off[ax] = (double)((int)(N[ax] / 2) + l->offsets[ax]) * h;
    double* g[3];
    for (int ax = 0; ax < 3; ax++) {
      g[ax] = (double*)malloc(sizeof(double) * N[ax]);
      for (uint64_t k = 0; k < N[ax]; k++) g[ax][k] = gauss(off[ax] - (double)k * h, R->model->conv[ax]);
    }
    double* proj = (double*)calloc(N[dir], sizeof(double));
    /* the reference keeps the voxels whose domain id is below the number
     * of listed domains (species.c:2015-2024) */
    const uint32_t mask = !l->num_domains ? 0xFFFFFFFFu
                          : l->num_domains >= 8 ? 0xFFu : ((1u << l->num_domains) - 1u);
    cs_try(cs_linescan(R->ctx, (uint32_t)R->linescan_species, (uint32_t)dir, g[0], g[1], g[2], mask, proj),
           "cs_linescan");
    double* line = (double*)calloc(N[dir], sizeof(double));
    for (uint64_t p = 0; p < N[dir]; p++) {
      double sum = 0;
      for (uint64_t k = 0; k < N[dir]; k++) sum += proj[k] * gauss((double)p * h - (double)k * h, R->model->conv[dir]);
      line[p] = sum;
    }
    char dom[64] = "", dsn[160];
    for (int k = 0; k < l->num_domains; k++) {
      size_t len = strlen(dom);
      snprintf(dom + len, sizeof dom - len, "%d_", R->linescan_domain_ids[k]);
    }
    const char axn[3] = {'x', 'y', 'z'};
    const int sgn = dir % 2 == 0 ? 1 : -1;
    const int o1 = ((dir + sgn) % 3 + 3) % 3, o2 = ((dir + 2 * sgn) % 3 + 3) % 3;
    snprintf(dsn, sizeof dsn, "linescan_%s_%s%c_%d_%c_%d", l->species, dom, axn[o1], l->offsets[o1], axn[o2], l->offsets[o2]);
    uint64_t dims[1] = {N[dir]};
    if (cs_out_dataset(o, dsn, line, 1, dims, 0)) die("*** ERROR: writing linescan failed\n");
    for (int ax = 0; ax < 3; ax++) free(g[ax]);
    free(proj);
    free(line);
  }
  cs_out_close(o);
  R->t_write += now() - t0;
  return 0;
}

typedef struct {
  cs_context* ctx;
  cs_setup* setup;
  uint64_t first, pending;
  double* uni;
  uint64_t uni_cap;
} flush_state;

/* issue the pending steps first .. first+pending-1 */
static void flush_steps(flush_state* F, uint64_t last_time_ind)
{
  if (!F->pending) return;
  if (F->setup->problem.rng == CS_RNG_MT) {
    uint64_t need = 0;
    cs_setup_mt_uniforms(F->setup, F->first, F->pending, NULL, &need);
    if (need > F->uni_cap) {
      free(F->uni);
      F->uni = (double*)malloc(sizeof(double) * need);
      F->uni_cap = need;
    }
    cs_setup_mt_uniforms(F->setup, F->first, F->pending, F->uni, &need);
    cs_try(cs_step_uniforms(F->ctx, F->first, F->pending, F->uni, need), "cs_step_uniforms");
  } else {
    cs_try(cs_step(F->ctx, F->first, F->pending), "cs_step");
  }
  F->first = last_time_ind + 1;
  F->pending = 0;
}

int main(int argc, char** argv)
{
  double t_start = now();
  cli_args a;
  memset(&a, 0, sizeof a);
  cs_options_defaults(&a.opt);
  a.casename = "casename";
  argp_parse(&argp, argc, argv, 0, 0, &a);

  /* one process: -p and --mpi_dist must describe a single rank
   * (calcium_sparks.c:77-98) */
  if (a.mpi_dist_given && a.mpi_dist[0] * a.mpi_dist[1] * a.mpi_dist[2] != 1)
    die("*** Error: --mpi_dist option was used. Expected the number of processes to be equal to %d. Instead got 1.\n",
        a.mpi_dist[0] * a.mpi_dist[1] * a.mpi_dist[2]);
  if (!a.geometry_file) die("*** ERROR: Expected a geometry file passed as argument: --geometry_file or -g\n");
  if (!a.model_file) die("*** ERROR: Expected a species file passed as argument: --species_file or -m\n");
  a.silent = a.verbose == 0 ? a.silent : 0;
  if (!a.silent) print_arguments(&a);

  cs_geometry geom;
  cs_species_model model;
  char err[512];
  if (cs_io_read_geometry(a.geometry_file, &geom, err, sizeof err)) die("%s\n", err);
  if (cs_io_read_species(a.model_file, &model, err, sizeof err)) die("%s\n", err);
  cs_setup* S = NULL;
  if (cs_setup_build(&geom, &model, &a.opt, &S, err, sizeof err) != CS_OK) die("%s\n", err);

  run_ctx R;
  memset(&R, 0, sizeof R);
  R.setup = S;
  R.args = &a;
  R.geom = &geom;
  R.model = &model;
  char names[512];
  /* save species (species.c:364-415) */
  R.save_idx = (int*)calloc(a.num_save_species + 1, sizeof(int));
  for (int i = 0; i < a.num_save_species; i++) {
    R.save_idx[i] = species_index(S, a.save_species[i]);
    if (R.save_idx[i] < 0) {
      all_species_list(S, names, sizeof names);
      die("*** ERROR: Expected given save species to be present in the model loaded from file. \n\"%s\" cannot be found in [%s].\n",
          a.save_species[i], names);
    }
  }
  R.all_idx = (int*)calloc(a.all_data_num_species + 1, sizeof(int));
  for (int i = 0; i < a.all_data_num_species; i++) {
    R.all_idx[i] = species_index(S, a.all_data_species[i]);
    if (R.all_idx[i] < 0) {
      all_species_list(S, names, sizeof names);
      die("*** ERROR: Expected given save species to be present in the model loaded from file. \n\"%s\" cannot be found in [%s].\n",
          a.all_data_species[i], names);
    }
  }
  /* sheet coordinates (species.c:792-836) */
  if (a.num_save_species) {
    if (a.num_ax_points[0] + a.num_ax_points[1] + a.num_ax_points[2] == 0) {
      a.num_ax_points[2] = 1;
      a.ax_points[2] = (double*)malloc(sizeof(double));
      a.ax_points[2][0] = (double)(S->N[2] / 2 - S->subdivisions / 2) * S->problem.h;
    }
    for (int ax = 0; ax < 3; ax++) {
      R.sheets_per_axis[ax] = a.num_ax_points[ax];
      R.global_indices[ax] = (int*)calloc(a.num_ax_points[ax] + 1, sizeof(int));
      for (int i = 0; i < a.num_ax_points[ax]; i++) {
        if (a.ax_points[ax][i] <= 0 || a.ax_points[ax][i] >= S->N[ax] * S->problem.h)
          die("*** ERROR: Expected given z coordinate for save species to be within: (0,%f), got: %f.\n",
              S->N[ax] * S->problem.h, a.ax_points[ax][i]);
        R.global_indices[ax][i] = (int)floor(a.ax_points[ax][i] / S->problem.h);
      }
    }
  }
  if (a.linescan) {
    R.linescan_species = species_index(S, a.linescan->species);
    if (R.linescan_species < 0) die("*** ERROR: \"%s\" is not a valid species name.\n", a.linescan->species);
    for (int i = 0; i < a.linescan->num_domains && i < CS_MAX_DOMAINS; i++) {
      int found = -1;
      for (uint32_t d = 0; d < S->num_domains; d++)
        if (!strcmp(S->domain_names[d], a.linescan->domains[i])) found = (int)d;
      if (found < 0) die("*** ERROR: \"%s\" is not a valid domain name.\n", a.linescan->domains[i]);
      R.linescan_domain_ids[i] = found;
    }
    for (int ax = 0; ax < 3; ax++) {
      const int lo = (int)(S->N[ax] / 2), hi = (int)S->N[ax] - lo - 1;
      if (ax != a.linescan->axis && (a.linescan->offsets[ax] < -lo || a.linescan->offsets[ax] > hi))
        die("*** ERROR: Expected given relative positions for convolution to be within: [%d, %d], got: %d.\n",
            -lo, hi, a.linescan->offsets[ax]);
    }
  }

  cs_context* ctx = NULL;
  cs_try(cs_create(&S->problem, &ctx), "cs_create");
  R.ctx = ctx;
  const uint64_t nv = S->N[0] * S->N[1] * S->N[2];
  R.field = (double*)malloc(sizeof(double) * nv);
  uint32_t maxch = 1;
  for (uint32_t k = 0; k < S->num_discrete; k++) if (S->discrete[k].size > maxch) maxch = S->discrete[k].size;
  R.states = (int32_t*)calloc(maxch, sizeof(int32_t));
  R.states8 = (uint8_t*)calloc(maxch, 1);
  R.snapshot_closed = (int*)calloc(S->num_discrete + 1, sizeof(int));

  const double dt = S->problem.dt;
  const uint64_t save_interval = cs_save_interval(a.opt.dt_save, dt);
  if (a.verbose) {
    /* Species_output_time_stepping_information, species.c:3210-3242 */
    printf("\nTime stepping:\n");
    printf("-----------------------------------------------------------------------------\n");
    printf("%13s %10.2g ms\n%13s%10.3g ms\n%13s%10zu\n%13s%10zu\n\n", "tstop", a.opt.tstop, "dt", dt, "num dt",
           (size_t)(a.opt.tstop / dt), "save interval", (size_t)save_interval);
    printf("%13s%13s%10s\n", "Diffusion", "dt [ms]", "substeps");
    for (uint32_t k = 0; k < S->num_diffusive_species; k++) {
      uint32_t si = S->all_diffusive[k];
      printf("%13s%13.3g%10d\n", S->species_names[si], dt * S->problem.substeps[si], (int)S->problem.substeps[si]);
    }
    printf("\n%13s%13s%10s\n", "", "dt [ms]", "substeps");
    printf("%13s%13.3g%10d\n", "Discr bound", S->problem.stochastic_substep * dt, (int)S->problem.stochastic_substep);
    printf("\n%13s%13s%10s\n", "", "dt [ms]", "substeps");
    printf("%13s%13.3g%10d\n", "Reaction", S->problem.reaction_substep * dt, (int)S->problem.reaction_substep);
    char dev[256];
    int devid = 0;
    if (cs_device_info(&devid, dev, sizeof dev) == CS_OK) printf("\nDevice: %d %s\n", devid, dev);
  }
  if (a.autotune) {
    cs_launch_config cfg;
    double ms = 0;
    cs_try(cs_autotune(ctx, 1, 1, &cfg, &ms), "cs_autotune");
    if (a.verbose)
      printf("Launch shape: rows %u zchunk %u xcd %u prefetch %u nontemporal %u lds %u (%.1f us)\n", cfg.rows,
             cfg.zchunk, cfg.xcd_remap, cfg.prefetch, cfg.nontemporal, cfg.lds_tile, ms * 1e3);
  }

  cs_abort_state ab;
  cs_abort_init(&ab, a.opt.tabort, S->num_discrete > 0);
  uint32_t output_ind = 0;
  write_output(&R, output_ind, 0.0, 0, &ab);

  /* the reference loop (calcium_sparks.c:190-228); device steps are
   * issued in runs and the host waits only at outputs */
  double t_loop = now();
  flush_state F = {ctx, S, 1, 0, NULL, 0};
  double t;
  uint64_t time_ind;
  for (t = 0.0, time_ind = 1; t < a.opt.tstop; t += dt, time_ind++) {
    F.pending++;
    if (time_ind % save_interval == 0) {
      flush_steps(&F, time_ind);
      cs_try(cs_synchronize(ctx), "cs_synchronize");
      write_output(&R, ++output_ind, t + dt, time_ind, &ab);
    }
    if (cs_abort_check(&ab, t + dt)) break;
  }
  flush_steps(&F, time_ind);
  cs_try(cs_synchronize(ctx), "cs_synchronize");
  t_loop = now() - t_loop;
  if (a.verbose) {
    printf("\n%30s %10s\n", "Timings", "seconds");
    printf("---------------------------------------------------------------\n");
    printf("%30s|%10.3f\n", "Write", R.t_write);
    printf("%30s|%10.3f\n", "Main Loop time", t_loop);
    printf("%30s|%10.3f\n", "Simulation time", now() - t_start);
  }
  cs_destroy(ctx);
  cs_setup_free(S);
  cs_io_free_geometry(&geom);
  free(R.field);
  free(R.states);
  free(R.states8);
  free(F.uni);
  return 0;
}
