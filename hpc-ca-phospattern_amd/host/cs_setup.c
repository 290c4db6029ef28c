/*
 * Problem builder: raw calcium_sparks inputs -> cs_problem for the device.
 *
 * Follows the reference's construction code step by step, so the derived
 * numbers (dt, substeps, tables, voxel pairs, initial channel states) are
 * the ones the reference would use; each block cites what it restates.
 */
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "cs_host.h"

#define DT_SCALE 0.95   /* reference types.h:46 */

static int fail(char* err, size_t len, const char* fmt, ...)
{
  if (err && len) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(err, len, fmt, ap);
    va_end(ap);
  }
  return CS_EINVAL;
}

void cs_options_defaults(cs_options* o)
{
  memset(o, 0, sizeof(*o));
  o->tstop = 10.0;
  o->h = 6.0;
  o->dt = -1.0;
  o->dt_save = 0.5;
  o->dt_update_react = -1.0;
  o->dt_update_stoch = -1.0;
  o->tabort = 0.0;
  o->precision = CS_PREC_F64;
  o->rng = CS_RNG_PHILOX;
}

static int find_name(const char (*names)[CS_NAME_LEN], uint32_t n, const char* name)
{
  for (uint32_t i = 0; i < n; i++)
    if (strncmp(names[i], name, CS_NAME_LEN) == 0) return (int)i;
  return -1;
}

static const cs_open_entry* find_open(const cs_options* o, const char* name)
{
  /* arguments_get_open_states_info, options.c:582-591 */
  for (uint32_t i = 0; i < o->num_open; i++)
    if (strncmp(o->open[i].name, name, CS_NAME_LEN) == 0) return &o->open[i];
  return NULL;
}

static double find_t_close(const cs_options* o, const char* name)
{
  /* arguments_get_t_close, options.c:593-601 */
  for (uint32_t i = 0; i < o->num_t_close; i++)
    if (strncmp(o->t_close[i].name, name, CS_NAME_LEN) == 0) return o->t_close[i].time;
  return -1.0;
}

static uint32_t random_seed(void)
{
  uint32_t s = 0;
  FILE* f = fopen("/dev/urandom", "rb");
  if (f) {
    if (fread(&s, sizeof s, 1, f) != 1) s = 0;
    fclose(f);
  }
  if (!s) s = (uint32_t)time(NULL) ^ (uint32_t)clock();
  return s;
}

/* Geometry_get_boundary_dir, geometry.c:1492-1530 */
static int boundary_dir(const uint16_t* p, int outer)
{
  for (int dim = 0; dim < 3; dim++)
    if (p[dim] != p[dim + 3]) return p[dim] < p[dim + 3] ? dim * 2 : dim * 2 + 1;
  return outer ? 0 : -1;
}

/* Expand geometry pairs to refined voxel pairs (species.c:938-1035). */
static int expand_pairs(const cs_geometry* g, uint32_t b, uint32_t sub, const uint32_t n[3],
                        int discrete, int outer, uint32_t** out, uint32_t* count,
                        char* err, size_t elen)
{
  const uint16_t* P = g->boundaries[b];
  uint64_t m = g->boundary_sizes[b];
  uint32_t sb = discrete ? 1 : sub;
  uint32_t* pairs = (uint32_t*)malloc(sizeof(uint32_t) * 2 * (size_t)(m * sb * sb ? m * sb * sb : 1));
  if (!pairs) return fail(err, elen, "*** ERROR: out of memory");
  size_t k = 0;
  for (uint64_t j = 0; j < m; j++) {
    const uint16_t* p = P + 6 * j;
    for (int c = 0; c < 6; c++)
      if (p[c] >= g->N[c % 3]) {
        free(pairs);
        return fail(err, elen, "*** ERROR: boundary %s[%llu] lies outside the geometry",
                    g->boundary_names[b], (unsigned long long)j);
      }
    int dir2 = boundary_dir(p, outer);
    if (dir2 < 0) {
      free(pairs);
      return fail(err, elen, "*** ERROR: Could not find direction of %s[%llu].",
                  g->boundary_names[b], (unsigned long long)j);
    }
    int dim = dir2 / 2, lr = dir2 % 2;
    int odim = dim != 0 ? 0 : 1;
    int idim = dim != 2 ? 2 : 1;
    uint32_t from_ind = p[dim] * sub + (lr == 0 ? sub - 1 : 0);
    uint32_t to_ind = p[dim + 3] * sub + (lr == 0 ? 0 : sub - 1);
    for (uint32_t oi = 0; oi < sb; oi++) {
      uint32_t o_ind = sub * p[odim] + oi + (discrete ? sub / 2 : 0);
      for (uint32_t ii = 0; ii < sb; ii++) {
        uint32_t i_ind = sub * p[idim] + ii + (discrete ? sub / 2 : 0);
        uint32_t f[3], t[3];
        f[dim] = from_ind; f[odim] = o_ind; f[idim] = i_ind;
        t[dim] = to_ind;   t[odim] = o_ind; t[idim] = i_ind;
        pairs[2 * k + 0] = (f[0] * n[1] + f[1]) * n[2] + f[2];
        pairs[2 * k + 1] = (t[0] * n[1] + t[1]) * n[2] + t[2];
        k++;
      }
    }
  }
  *out = pairs;
  *count = (uint32_t)k;
  return CS_OK;
}

/* Position of a boundary (geometry.c:642-662): outer when both voxels of
 * a pair coincide; all pairs must agree. */
static int boundary_outer(const cs_geometry* g, uint32_t b, int* outer, char* err, size_t elen)
{
  const uint16_t* P = g->boundaries[b];
  for (uint64_t j = 0; j < g->boundary_sizes[b]; j++) {
    const uint16_t* p = P + 6 * j;
    int eq = p[0] == p[3] && p[1] == p[4] && p[2] == p[5];
    if (j == 0) *outer = eq;
    else if (*outer != eq)
      return fail(err, elen, "*** ERROR: boundary %s mixes inner and outer pairs",
                  g->boundary_names[b]);
  }
  if (g->boundary_sizes[b] == 0) *outer = 0;
  return CS_OK;
}

static void channel_params(const cs_channel_model* m, double t_close, double* p)
{
  /* parameter vector order of boundaryfluxes.c:447-456 */
  p[0] = m->Kd_close;  p[1] = m->k_max_close; p[2] = m->n_open;
  p[3] = m->k_max_open; p[4] = m->k_min_open; p[5] = m->n_close;
  p[6] = m->i_unitary; p[7] = m->k_min_close; p[8] = t_close;
  p[9] = m->Kd_open;
}

/* init_RyR_model_states_stochastically, boundaryfluxes.c:229-272 */
static void init_states_stochastic(const double* p, uint32_t n, const double* cto,
                                   cs_mt* mt, uint8_t* s)
{
  for (uint32_t i = 0; i < n; i++) {
    double c0 = cto[i];
    double ko = fmax(fmin(pow(c0 / p[9], p[2]), p[3]), p[4]);
    double kc = fmax(fmin(pow(c0 / p[0], p[5]), p[1]), p[7]);
    double r = cs_mt_ldrand(mt);
    s[i] = (r <= kc / (ko + kc)) ? 0 : 1;
  }
}

void cs_setup_free(cs_setup* s)
{
  if (!s) return;
  free((void*)s->problem.domain_ids);
  for (uint32_t f = 0; f < CS_MAX_FLUXES; f++) {
    free((void*)s->problem.fluxes[f].pairs);
    free((void*)s->problem.fluxes[f].init_states);
  }
  free(s);
}

int cs_setup_build(const cs_geometry* g, const cs_species_model* m,
                   const cs_options* opt, cs_setup** out, char* err, size_t elen)
{
  *out = NULL;
  if (err && elen) err[0] = 0;
  if (g->num_domains == 0 || g->num_domains > CS_MAX_DOMAINS)
    return fail(err, elen, "*** ERROR: unsupported number of domains %u", g->num_domains);
  if (m->num_species == 0 || m->num_species > CS_MAX_SPECIES)
    return fail(err, elen, "*** ERROR: unsupported number of species %u", m->num_species);
  if (g->num_boundaries > CS_MAX_BOUNDARIES)
    return fail(err, elen, "*** ERROR: unsupported number of boundaries %u", g->num_boundaries);
  /* species.c:114-116, :176-180 */
  if (m->num_domains != g->num_domains)
    return fail(err, elen, "*** ERROR: Expected the same number of domains in geometry "
                "and model files. %d!=%d", g->num_domains, m->num_domains);
  for (uint32_t d = 0; d < g->num_domains; d++)
    if (strncmp(g->domain_names[d], m->domain_names[d], CS_NAME_LEN) != 0)
      return fail(err, elen, "*** ERROR: Expected the domain names to be the same in "
                  "geometry and model files. %s!=%s", g->domain_names[d], m->domain_names[d]);
  if (opt->h <= 0) return fail(err, elen, "*** ERROR: resolution must be positive");

  cs_setup* s = (cs_setup*)calloc(1, sizeof(cs_setup));
  if (!s) return fail(err, elen, "*** ERROR: out of memory");
  cs_problem* P = &s->problem;
  const uint32_t nd = g->num_domains, ns = m->num_species;
  s->num_domains = nd;
  s->num_species = ns;
  memcpy(s->domain_names, g->domain_names, sizeof s->domain_names);
  memcpy(s->species_names, m->species_names, sizeof s->species_names);

  /* geometry.c:104-105 */
  double subd = fmax(1.0, floor(g->h / opt->h));
  if (subd > 255) { cs_setup_free(s); return fail(err, elen, "*** ERROR: too many subdivisions"); }
  uint32_t sub = (uint32_t)subd;
  double h = g->h / sub;
  s->subdivisions = sub;
  s->h_geom = g->h;
  uint32_t n[3];
  for (int a = 0; a < 3; a++) {
    uint64_t v = g->N[a] * sub;
    if (g->N[a] == 0 || v > 0xFFFFFFFFull) { cs_setup_free(s); return fail(err, elen, "*** ERROR: bad geometry size"); }
    n[a] = (uint32_t)v;
    s->N[a] = v;
    P->n[a] = n[a];
  }
  uint64_t nvox = (uint64_t)n[0] * n[1] * n[2];
  if (nvox > 0xFFFFFFFFull) { cs_setup_free(s); return fail(err, elen, "*** ERROR: more than 2^32 voxels"); }

  /* geometry voxel counts and volumes (geometry.c:146-232) */
  const uint64_t gvox = g->N[0] * g->N[1] * g->N[2];
  for (uint64_t i = 0; i < gvox; i++) {
    uint8_t d = g->voxels[i];
    if (d >= nd) { cs_setup_free(s); return fail(err, elen, "*** ERROR: voxel domain id %u out of range", d); }
    s->domain_voxels[d]++;
  }
  double dVg = g->h * g->h * g->h;
  for (uint32_t d = 0; d < nd; d++) s->volumes[d] = s->domain_voxels[d] * dVg;

  /* refined domain ids (species.c:585-596) */
  uint8_t* ids = (uint8_t*)malloc(nvox);
  if (!ids) { cs_setup_free(s); return fail(err, elen, "*** ERROR: out of memory"); }
  for (uint32_t x = 0; x < n[0]; x++)
    for (uint32_t y = 0; y < n[1]; y++) {
      const uint8_t* src = g->voxels + ((uint64_t)(x / sub) * g->N[1] + y / sub) * g->N[2];
      uint8_t* dst = ids + ((uint64_t)x * n[1] + y) * n[2];
      for (uint32_t z = 0; z < n[2]; z++) dst[z] = src[z / sub];
    }
  P->domain_ids = ids;
  P->num_domains = nd;
  P->num_species = ns;
  P->h = h;
  P->dV = h * h * h;              /* species.c:155 */
  P->precision = opt->precision;
  P->rng = opt->rng;

  /* diffusive species and sigma (species.c:159-209) */
  double max_sigma = 0.0;
  double max_sigma_species[CS_MAX_SPECIES] = {0};
  uint8_t is_diff[CS_MAX_SPECIES] = {0};
  for (uint32_t d = 0; d < nd; d++) {
    if (m->num_diffusive[d] > ns) { cs_setup_free(s); return fail(err, elen, "*** ERROR: bad num_diffusive"); }
    for (uint32_t k = 0; k < m->num_diffusive[d]; k++) {
      uint32_t sp = m->diffusive[d][k];
      if (sp >= ns) { cs_setup_free(s); return fail(err, elen, "*** ERROR: bad diffusive species index"); }
      is_diff[sp] = 1;
      max_sigma = fmax(max_sigma, m->sigma[d][k]);
      max_sigma_species[sp] = fmax(m->sigma[d][k], max_sigma_species[sp]);
    }
  }
  /* species.c:265 (expression order kept so dt is bit-identical) */
  double dt = opt->dt > 0 ? opt->dt : DT_SCALE * (1. / 6) * h * h / max_sigma;
  if (!(dt > 0) || !isfinite(dt)) { cs_setup_free(s); return fail(err, elen, "*** ERROR: cannot derive a time step"); }
  P->dt = dt;

  /* per-species substeps (species.c:281-294) */
  uint32_t ndiff = 0;
  uint32_t substeps_dsi[CS_MAX_SPECIES] = {0};
  for (uint32_t sp = 0; sp < ns; sp++) {
    if (!is_diff[sp]) continue;
    double sdt = fabs(max_sigma - max_sigma_species[sp]) < 1e-12 ? dt
                 : DT_SCALE * (1. / 6) * h * h / max_sigma_species[sp];
    double k = fmax(1., floor(sdt / dt));
    if (k > 255) { cs_setup_free(s); return fail(err, elen, "*** ERROR: species %s needs more than 255 substeps", m->species_names[sp]); }
    substeps_dsi[ndiff] = (uint32_t)k;
    s->all_diffusive[ndiff] = (uint8_t)sp;
    P->substeps[sp] = (uint32_t)k;
    ndiff++;
  }
  s->num_diffusive_species = ndiff;

  /* coefficient tables (species.c:1547-1560); sigma is indexed by the
   * global diffusive index dsi, as the reference does */
  for (uint32_t dsi = 0; dsi < ndiff; dsi++) {
    uint32_t sp = s->all_diffusive[dsi];
    double* C = P->coef[sp];
    double kk = substeps_dsi[dsi];
    for (uint32_t di = 0; di < nd; di++) {
      if (!(dsi < m->num_diffusive[di] && m->sigma[di][dsi] != 0)) continue;
      for (uint32_t dn = 0; dn < di; dn++)
        if (g->domain_connections[di * nd + dn] && dsi < m->num_diffusive[dn] && m->sigma[dn][dsi] != 0)
          C[di * CS_MAX_DOMAINS + dn] = C[dn * CS_MAX_DOMAINS + di] =
            dt * 0.5 * (m->sigma[di][dsi] + m->sigma[dn][dsi]) / (h * h) * kk;
      C[di * CS_MAX_DOMAINS + di] = dt * m->sigma[di][dsi] / (h * h) * kk;
    }
  }

  /* buffers (species.c:211-257, :298-308) */
  uint8_t is_buf[CS_MAX_SPECIES] = {0};
  for (uint32_t d = 0; d < nd; d++) {
    if (m->num_buffers[d] > ns) { cs_setup_free(s); return fail(err, elen, "*** ERROR: bad num_buffers"); }
    for (uint32_t b = 0; b < m->num_buffers[d]; b++) {
      uint8_t b0 = m->buffer_species[d][b][0], b1 = m->buffer_species[d][b][1];
      if (b0 >= ns || b1 >= ns) { cs_setup_free(s); return fail(err, elen, "*** ERROR: bad buffer species"); }
      P->bsp0[b][d] = b0;
      P->bsp1[b][d] = b1;
      P->k_on[b][d] = m->k_on[d][b];
      P->k_off[b][d] = m->k_off[d][b];
      P->tot[b][d] = m->tot[d][b];
      is_buf[b0] = 1;
    }
  }
  uint32_t nbuf = 0;
  for (uint32_t sp = 0; sp < ns; sp++) nbuf += is_buf[sp];
  P->num_buffers = nbuf;
  /* species.c:784-787 */
  P->stochastic_substep = (uint32_t)fmax(1., floor(opt->dt_update_stoch / dt));
  P->reaction_substep = (uint32_t)fmax(1., floor(opt->dt_update_react / dt));
  P->dt_react = dt * P->reaction_substep;   /* species.c:1666 */

  for (uint32_t d = 0; d < nd; d++)
    for (uint32_t sp = 0; sp < ns; sp++) P->init[d][sp] = m->init[d][sp];

  /* fixed domain species (species.c:345-358, :1037-1044) */
  for (uint32_t i = 0; i < m->num_fixed; i++) {
    uint8_t d = m->fixed[2 * i], sp = m->fixed[2 * i + 1];
    if (d >= nd || sp >= ns) { cs_setup_free(s); return fail(err, elen, "*** ERROR: bad fixed domain species"); }
    P->fixed[d][sp] = 1;
  }
  P->force_dirichlet = opt->force_dirichlet ? 1 : 0;

  /* discrete boundaries in file order (geometry.c:256-262) */
  for (uint32_t b = 0; b < g->num_boundaries; b++) {
    if (g->boundary_types[b] != CS_BOUNDARY_DISCRETE) continue;
    cs_discrete_info* di = &s->discrete[s->num_discrete++];
    strncpy(di->name, g->boundary_names[b], CS_NAME_LEN - 1);
    di->boundary = b;
    di->flux = -1;
    di->size = (uint32_t)g->boundary_sizes[b];
  }
  int ryr_b = find_name(g->boundary_names, g->num_boundaries, "ryr");
  int pryr_b = find_name(g->boundary_names, g->num_boundaries, "p_ryr");
  s->has_ryr_boundary = ryr_b >= 0;
  s->has_p_ryr_boundary = pryr_b >= 0;
  s->ryr_discrete = s->p_ryr_discrete = -1;
  for (uint32_t k = 0; k < s->num_discrete; k++) {
    if ((int)s->discrete[k].boundary == ryr_b) s->ryr_discrete = (int)k;
    if ((int)s->discrete[k].boundary == pryr_b) s->p_ryr_discrete = (int)k;
  }

  /* -O indices must lie inside their boundary (calcium_sparks.c:134-149) */
  for (uint32_t i = 0; i < opt->num_open; i++) {
    int b = find_name(g->boundary_names, g->num_boundaries, opt->open[i].name);
    if (b < 0) { cs_setup_free(s); return fail(err, elen, "*** ERROR: \"%s\" is not a valid boundary name.", opt->open[i].name); }
    int32_t mx = 0;
    for (uint32_t j = 0; j < opt->open[i].count; j++) if (opt->open[i].idx[j] > mx) mx = opt->open[i].idx[j];
    if ((uint64_t)mx >= g->boundary_sizes[b]) {
      cs_setup_free(s);
      return fail(err, elen, "*** ERROR: Expected max %s indices given by option "
                  "--force_discrete_boundaries_open: %d to be lower than the number of %s boundaries: %d",
                  opt->open[i].name, mx, opt->open[i].name, (int)g->boundary_sizes[b]);
    }
  }

  /* flux models (boundaryfluxes.c:361-542): ryr, p_ryr, serca; the first
   * model registered on a boundary wins (BoundaryFluxes_get_flux_info) */
  typedef struct { int kind; const cs_channel_model* ch; const char* lit; const char* bname; const char* spname; } model_ref;
  model_ref mods[3];
  int nm = 0;
  if (m->use_ryr)   { mods[nm].kind = CS_FLUX_RYR; mods[nm].ch = &m->ryr; mods[nm].lit = "ryr"; mods[nm].bname = m->ryr.boundary; mods[nm].spname = m->ryr.species; nm++; }
  if (m->use_p_ryr) { mods[nm].kind = CS_FLUX_RYR; mods[nm].ch = &m->p_ryr; mods[nm].lit = "p_ryr"; mods[nm].bname = m->p_ryr.boundary; mods[nm].spname = m->p_ryr.species; nm++; }
  if (m->use_serca) { mods[nm].kind = CS_FLUX_SERCA; mods[nm].ch = NULL; mods[nm].lit = NULL; mods[nm].bname = m->serca.boundary; mods[nm].spname = m->serca.species; nm++; }
  int mod_boundary[3];
  int boundary_model[CS_MAX_BOUNDARIES];
  for (uint32_t b = 0; b < CS_MAX_BOUNDARIES; b++) boundary_model[b] = -1;
  for (int i = 0; i < nm; i++) {
    int b = find_name(g->boundary_names, g->num_boundaries, mods[i].bname);
    if (b < 0) { cs_setup_free(s); return fail(err, elen, "*** ERROR: \"%s\" is not a valid boundary name.", mods[i].bname); }
    mod_boundary[i] = b;
    if (boundary_model[b] < 0) boundary_model[b] = i;
    if (mods[i].lit && find_name(g->boundary_names, g->num_boundaries, mods[i].lit) < 0) {
      cs_setup_free(s);
      return fail(err, elen, "*** ERROR: \"%s\" is not a valid boundary name.", mods[i].lit);
    }
  }
  /* order: discrete boundaries in gating order, then the rest */
  uint32_t order[CS_MAX_BOUNDARIES], norder = 0;
  for (uint32_t k = 0; k < s->num_discrete; k++) {
    uint32_t b = s->discrete[k].boundary;
    if (boundary_model[b] >= 0) order[norder++] = b;
  }
  for (uint32_t b = 0; b < g->num_boundaries; b++)
    if (boundary_model[b] >= 0 && g->boundary_types[b] != CS_BOUNDARY_DISCRETE) order[norder++] = b;
  if (norder > CS_MAX_FLUXES) { cs_setup_free(s); return fail(err, elen, "*** ERROR: too many boundary fluxes"); }

  /* seed (calcium_sparks.c:113-116) */
  s->seed = opt->seed_given ? opt->seed : random_seed();
  cs_mt_seed(&s->mt, s->seed);
  P->seed = s->seed;
  P->instance = 0;

  int flux_of_model[3] = {-1, -1, -1};
  for (uint32_t f = 0; f < norder; f++) {
    uint32_t b = order[f];
    int mi = boundary_model[b];
    flux_of_model[mi] = (int)f;
    cs_flux_desc* F = &P->fluxes[f];
    int sp = find_name(m->species_names, ns, mods[mi].spname);
    if (sp < 0) { cs_setup_free(s); return fail(err, elen, "*** ERROR: \"%s\" is not a valid species name.", mods[mi].spname); }
    int dsi = -1;
    for (uint32_t k = 0; k < ndiff; k++) if (s->all_diffusive[k] == (uint8_t)sp) dsi = (int)k;
    if (dsi < 0) { cs_setup_free(s); return fail(err, elen, "*** ERROR: \"%s\" is not a valid diffusive species name.", mods[mi].spname); }
    int outer = 0;
    if (boundary_outer(g, b, &outer, err, elen) != CS_OK) { cs_setup_free(s); return CS_EINVAL; }
    int discrete = g->boundary_types[b] == CS_BOUNDARY_DISCRETE;
    if (discrete && mods[mi].kind != CS_FLUX_RYR) {
      cs_setup_free(s);
      return fail(err, elen, "*** ERROR: discrete boundary %s needs a gated flux model", g->boundary_names[b]);
    }
    uint32_t* pairs = NULL;
    uint32_t np = 0;
    if (expand_pairs(g, b, sub, n, discrete, outer, &pairs, &np, err, elen) != CS_OK) { cs_setup_free(s); return CS_EINVAL; }
    F->kind = (uint32_t)mods[mi].kind;
    F->species = (uint32_t)sp;
    F->substep = substeps_dsi[dsi];
    F->inner = outer ? 0 : 1;
    F->discrete = (uint32_t)discrete;
    F->num_pairs = np;
    F->pairs = pairs;
    if (mods[mi].kind == CS_FLUX_RYR) {
      channel_params(mods[mi].ch, find_t_close(opt, mods[mi].lit), F->params);
    } else {
      F->params[0] = m->serca.v_cyt_to_A_sr;
      F->params[1] = m->serca.density;
      F->params[2] = m->serca.scale;
    }
    if (discrete) {
      uint8_t* st = (uint8_t*)calloc(np ? np : 1, 1);
      if (!st) { cs_setup_free(s); return fail(err, elen, "*** ERROR: out of memory"); }
      F->init_states = st;
      for (uint32_t k = 0; k < s->num_discrete; k++)
        if (s->discrete[k].boundary == b) s->discrete[k].flux = (int32_t)f;
    }
  }
  P->num_fluxes = norder;

  /* initial channel states (boundaryfluxes.c:544-605) */
  if (m->use_ryr && s->num_discrete > 0) {
    const cs_open_entry* info = find_open(opt, m->ryr.boundary);
    /* Ca (species slot 0) at the to-voxels of the first model's boundary */
    int fb = flux_of_model[0];
    double* cto = NULL;
    uint32_t ncto = 0;
    if (fb >= 0 && P->fluxes[fb].discrete) {
      const cs_flux_desc* F = &P->fluxes[fb];
      ncto = F->num_pairs;
      cto = (double*)malloc(sizeof(double) * (ncto ? ncto : 1));
      for (uint32_t i = 0; i < ncto; i++) cto[i] = P->init[ids[F->pairs[2 * i + 1]]][0];
    }
    /* inside this branch mods[0] is the ryr model and mods[1] the p_ryr one */
    for (int which = 0; which < 2; which++) {
      if (which == 1 && !m->use_p_ryr) break;
      int f = flux_of_model[which];
      const cs_open_entry* oi = which == 0 ? info : find_open(opt, m->p_ryr.boundary);
      if (f < 0 || !P->fluxes[f].discrete) continue;
      cs_flux_desc* F = &P->fluxes[f];
      uint8_t* st = (uint8_t*)F->init_states;
      if (oi) {
        memset(st, 0, F->num_pairs);
        for (uint32_t j = 0; j < oi->count; j++) {
          if (oi->idx[j] < 0 || (uint32_t)oi->idx[j] >= F->num_pairs) { free(cto); cs_setup_free(s); return fail(err, elen, "*** ERROR: open index out of range"); }
          st[oi->idx[j]] = 1;
        }
      } else {
        if (!cto || F->num_pairs > ncto) {
          free(cto);
          cs_setup_free(s);
          return fail(err, elen, "*** ERROR: stochastic initialisation of %s reads the ryr "
                      "boundary, which has fewer channels", which ? "p_ryr" : "ryr");
        }
        init_states_stochastic(F->params, F->num_pairs, cto, &s->mt, st);
      }
    }
    free(cto);
  }
  *out = s;
  return CS_OK;
}

int cs_setup_mt_uniforms(cs_setup* s, uint64_t first, uint64_t nsteps, double* out, uint64_t* count)
{
  const cs_problem* P = &s->problem;
  uint64_t per = 0;
  for (uint32_t f = 0; f < P->num_fluxes; f++)
    if (P->fluxes[f].discrete && !(P->fluxes[f].params[8] > 0)) per += P->fluxes[f].num_pairs;
  uint64_t D = P->stochastic_substep, gates = 0;
  for (uint64_t t = first; t < first + nsteps; t++) if (t % D == 0) gates++;
  if (s->num_discrete == 0) gates = 0;
  *count = gates * per;
  if (!out) return CS_OK;
  for (uint64_t i = 0; i < *count; i++) out[i] = cs_mt_ldrand(&s->mt);
  return CS_OK;
}

/* Struct sizes, so foreign-language mirrors can check their layout. */
size_t cs_host_sizeof(int which)
{
  switch (which) {
    case 0: return sizeof(cs_geometry);
    case 1: return sizeof(cs_species_model);
    case 2: return sizeof(cs_options);
    case 3: return sizeof(cs_problem);
    case 4: return sizeof(cs_setup);
    case 5: return sizeof(cs_flux_desc);
    default: return 0;
  }
}
