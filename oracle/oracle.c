/*
 * oracle.c -- CPU FP64 restatement of the calcium_sparks step.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Written from the reference's
 * semantics, not from the product code: the setup is re-derived here and
 * the stepping keeps the reference's u1 / du / update_species sequence.
 */
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include "oracle.h"

#define ND_MAX CS_MAX_DOMAINS
#define NS_MAX CS_MAX_SPECIES

typedef struct obnd {
  int      model;          /* -1 none, 0 ryr, 1 p_ryr, 2 serca            */
  int      discrete, inner;
  uint32_t npairs;
  uint64_t* from;          /* refined linear indices                      */
  uint64_t* to;
  uint8_t* open;           /* discrete only                               */
  uint8_t* ever;           /* open at any gating step so far (discrete)   */
  int      gate_group;     /* ordinal among gated discrete boundaries     */
  double   params[10];
  int      dsi;            /* diffusive index of the flux species         */
} obnd;

struct oracle_sim {
  uint64_t n[3], nv;
  uint32_t sub, nd, ns, nb_geom;
  double h, dt, dV;
  uint8_t* ids;
  double* u1[NS_MAX];
  double* du[NS_MAX];
  uint8_t upd[NS_MAX];
  double init[ND_MAX][NS_MAX];
  /* diffusion */
  uint32_t ndiff;
  uint32_t alldiff[NS_MAX];
  uint32_t k[NS_MAX];               /* by diffusive index */
  double   coef[NS_MAX][25];        /* by diffusive index, stride 5 */
  /* reaction (transposed [bi*nd+di]) */
  uint32_t nbuf, R, D;
  uint8_t  bsp0T[NS_MAX * ND_MAX], bsp1T[NS_MAX * ND_MAX];
  double   konT[NS_MAX * ND_MAX], koffT[NS_MAX * ND_MAX], totT[NS_MAX * ND_MAX];
  uint8_t  bsp0[ND_MAX][NS_MAX], bsp1[ND_MAX][NS_MAX];
  /* boundaries */
  uint32_t nbnd;
  obnd     bnd[CS_MAX_BOUNDARIES];
  uint32_t ndisc;
  uint32_t disc_b[CS_MAX_BOUNDARIES];
  /* misc */
  int dirichlet;
  uint8_t fixed[ND_MAX][NS_MAX];
  int rng;
  uint32_t seed;
  struct mt_state* mt;       /* this run's Mersenne Twister             */
};

/* ------------------------------------------------------------------ RNG */
/* MT19937, Matsumoto & Nishimura (1998), 2002 seeding; mt_ldrand combines
 * two consecutive outputs as in mtwist.h:635-677.  The reference keeps one
 * generator per process (mtwist's default state), which is one per
 * realisation; here each oracle run owns its generator (ostate), so runs
 * that step interleaved in one process do not share a stream.  The
 * oracle_mt_* exports use a process-wide one, for known-answer tests. */
typedef struct mt_state {
  uint32_t mt[624];
  int mti;
} mt_state;

static void mt_seed(mt_state* g, uint32_t seed)
{
  g->mt[0] = seed;
  for (int i = 1; i < 624; i++) g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
  g->mti = 624;
}

static uint32_t mt_u32(mt_state* g)
{
  static const uint32_t mag[2] = {0u, 0x9908b0dfu};
  uint32_t y;
  if (g->mti >= 624) {
    if (g->mti == 625) mt_seed(g, 5489u);
    int kk;
    for (kk = 0; kk < 624 - 397; kk++) {
      y = (g->mt[kk] & 0x80000000u) | (g->mt[kk + 1] & 0x7fffffffu);
      g->mt[kk] = g->mt[kk + 397] ^ (y >> 1) ^ mag[y & 1u];
    }
    for (; kk < 623; kk++) {
      y = (g->mt[kk] & 0x80000000u) | (g->mt[kk + 1] & 0x7fffffffu);
      g->mt[kk] = g->mt[kk + (397 - 624)] ^ (y >> 1) ^ mag[y & 1u];
    }
    y = (g->mt[623] & 0x80000000u) | (g->mt[0] & 0x7fffffffu);
    g->mt[623] = g->mt[396] ^ (y >> 1) ^ mag[y & 1u];
    g->mti = 0;
  }
  y = g->mt[g->mti++];
  y ^= (y >> 11);
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= (y >> 18);
  return y;
}

static double mt_ldrand(mt_state* g)
{
  uint64_t a = mt_u32(g);
  uint64_t b = mt_u32(g);
  return (double)((a << 32) | b) / 18446744073709551616.0;
}

static mt_state g_mt = {{0}, 625};

void oracle_mt_seed(uint32_t seed) { mt_seed(&g_mt, seed); }
uint32_t oracle_mt_u32(void) { return mt_u32(&g_mt); }
double oracle_mt_ldrand(void) { return mt_ldrand(&g_mt); }

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11). */
void oracle_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; r++) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static double philox_uniform(uint32_t seed, uint64_t ti, uint32_t group, uint32_t ch)
{
  uint32_t c[4] = {(uint32_t)ti, (uint32_t)(ti >> 32), group, ch};
  uint32_t k[2] = {seed, 0u};
  uint32_t o[4];
  oracle_philox(c, k, o);
  uint64_t m = ((uint64_t)o[0] << 21) | (o[1] >> 11);
  return (double)m * 0x1p-53;
}

/* ----------------------------------------------------- flux functions */
/* serca_flux, boundaryfluxes.c:12-85 */
double oracle_serca_flux(double dt, double h, double u0, double u1, const double* params)
{
  double Ca_sr = u1 / 1000, Ca_i = u0 / 1000;
  double k_p1 = 25900, k_p2 = 2540, k_p3 = 20.5, k_m1 = 16, k_m2 = 67200, k_m3 = 149;
  double kdcai = 0.9, kdcasr = 2.24, kdh1 = 1.09e-5, kdhi = 3.54e-3, kdhsr = 1.05e-8, kdh = 7.24e-5;
  double H_i = 1e-4, ATP = 5, ADP = 36.3e-3, P_i = 1;
  double T_Cai = Ca_i / kdcai, T_Casr = Ca_sr / kdcasr;
  double T_Cai2 = T_Cai * T_Cai, T_Casr2 = T_Casr * T_Casr;
  double T_H1 = H_i / kdh1, T_Hi = (H_i * H_i) / kdhi, T_Hsr = (H_i * H_i) / kdhsr, T_H = H_i / kdh;
  double a_p1 = k_p1 * ATP;
  double a_p2 = k_p2 * T_Cai2 / (T_Cai2 + T_Cai2 * T_Hi + T_Hi * (1 + T_H1));
  double a_p3 = k_p3 * T_Hsr / (T_Casr2 * T_H + T_H + T_Hsr * (1 + T_H));
  double a_m1 = k_m1 * T_Hi / (T_Cai2 + T_Cai2 * T_Hi + T_Hi * (1 + T_H1));
  double a_m2 = k_m2 * ADP * T_Casr2 * T_H / (T_Casr2 * T_H + T_H + T_Hsr * (1 + T_H));
  double a_m3 = k_m3 * P_i;
  double s1 = a_p2 * a_p3 + a_m1 * a_p3 + a_m1 * a_m2;
  double s2 = a_p1 * a_p3 + a_m2 * a_p1 + a_m2 * a_m3;
  double s3 = a_p1 * a_p2 + a_m3 * a_m1 + a_m3 * a_p2;
  double v_cycle = (a_p1 * a_p2 * a_p3 - a_m1 * a_m2 * a_m3) / (s1 + s2 + s3);
  double density = params[1];
  double Q = 1e-3 / 6e17;
  double S = v_cycle * Q;
  double I = density * S;
  double A = h * h;
  double J = A * I;
  double V = 1e-24 * h * h * h;
  return dt * (J / V);
}

/* ryr_flux, boundaryfluxes.c:87-141: exact two-voxel relaxation */
double oracle_ryr_flux(double dt, double h, double u0, double u1, const double* params)
{
  double i_c = params[6];
  double K_0 = exp(-2 * dt * 1e15 / (2 * 96485.) * i_c / (1300. - 0.1) / (h * h * h));
  return (u0 - u1) * (1 - K_0) / 2;
}

/* ------------------------------------------------------------- helpers */
static int ofail(char* err, size_t len, const char* fmt, ...)
{
  if (err && len) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(err, len, fmt, ap);
    va_end(ap);
  }
  return -1;
}

static int name_index(const char (*names)[CS_NAME_LEN], uint32_t n, const char* s)
{
  for (uint32_t i = 0; i < n; i++) if (!strncmp(names[i], s, CS_NAME_LEN)) return (int)i;
  return -1;
}

void oracle_free(oracle_sim* s)
{
  if (!s) return;
  free(s->ids);
  for (uint32_t i = 0; i < NS_MAX; i++) { free(s->u1[i]); free(s->du[i]); }
  for (uint32_t b = 0; b < CS_MAX_BOUNDARIES; b++) {
    free(s->bnd[b].from); free(s->bnd[b].to); free(s->bnd[b].open); free(s->bnd[b].ever);
  }
  free(s->mt);
  free(s);
}

#define IDX(x, y, z) (((uint64_t)(x) * s->n[1] + (y)) * s->n[2] + (z))

/* ---------------------------------------------------------------- setup */
int oracle_create(const cs_geometry* g, const cs_species_model* m, const cs_options* opt,
                  oracle_sim** out, char* err, size_t elen)
{
  *out = NULL;
  if (m->num_domains != g->num_domains || g->num_domains > ND_MAX || m->num_species > NS_MAX)
    return ofail(err, elen, "domain/species count mismatch");
  oracle_sim* s = (oracle_sim*)calloc(1, sizeof(oracle_sim));
  s->nd = g->num_domains;
  s->ns = m->num_species;
  /* geometry.c:104-105 */
  s->sub = (uint32_t)fmax(1., floor(g->h / opt->h));
  s->h = g->h / s->sub;
  for (int a = 0; a < 3; a++) s->n[a] = g->N[a] * s->sub;
  s->nv = s->n[0] * s->n[1] * s->n[2];
  s->dV = s->h * s->h * s->h;
  s->ids = (uint8_t*)malloc(s->nv);
  for (uint64_t x = 0; x < s->n[0]; x++)
    for (uint64_t y = 0; y < s->n[1]; y++)
      for (uint64_t z = 0; z < s->n[2]; z++)
        s->ids[IDX(x, y, z)] = g->voxels[((x / s->sub) * g->N[1] + y / s->sub) * g->N[2] + z / s->sub];

  /* species.c:159-209, :265 */
  double maxs = 0, maxs_sp[NS_MAX] = {0};
  int isdiff[NS_MAX] = {0};
  for (uint32_t d = 0; d < s->nd; d++)
    for (uint32_t i = 0; i < m->num_diffusive[d]; i++) {
      uint32_t sp = m->diffusive[d][i];
      isdiff[sp] = 1;
      maxs = fmax(maxs, m->sigma[d][i]);
      maxs_sp[sp] = fmax(m->sigma[d][i], maxs_sp[sp]);
    }
  s->dt = opt->dt > 0 ? opt->dt : 0.95 * (1. / 6) * s->h * s->h / maxs;
  /* species.c:281-294 */
  s->ndiff = 0;
  for (uint32_t sp = 0; sp < s->ns; sp++) {
    if (!isdiff[sp]) continue;
    double sdt = fabs(maxs - maxs_sp[sp]) < 1e-12 ? s->dt : 0.95 * (1. / 6) * s->h * s->h / maxs_sp[sp];
    s->k[s->ndiff] = (uint32_t)fmax(1., floor(sdt / s->dt));
    s->alldiff[s->ndiff] = sp;
    s->ndiff++;
  }
  /* species.c:1547-1560 */
  for (uint32_t dsi = 0; dsi < s->ndiff; dsi++) {
    double* C = s->coef[dsi];
    for (uint32_t di = 0; di < s->nd; di++)
      if (dsi < m->num_diffusive[di] && m->sigma[di][dsi] != 0) {
        for (uint32_t dn = 0; dn < di; dn++)
          if (g->domain_connections[di * s->nd + dn] && dsi < m->num_diffusive[dn] && m->sigma[dn][dsi] != 0)
            C[di * 5 + dn] = C[dn * 5 + di] =
              s->dt * 0.5 * (m->sigma[di][dsi] + m->sigma[dn][dsi]) / (s->h * s->h) * s->k[dsi];
        C[di * 5 + di] = s->dt * m->sigma[di][dsi] / (s->h * s->h) * s->k[dsi];
      }
  }
  /* species.c:232-257, :298-308 */
  int isbuf[NS_MAX] = {0};
  for (uint32_t d = 0; d < s->nd; d++)
    for (uint32_t b = 0; b < m->num_buffers[d]; b++) {
      uint32_t t = b * s->nd + d;
      s->bsp0T[t] = m->buffer_species[d][b][0];
      s->bsp1T[t] = m->buffer_species[d][b][1];
      s->konT[t] = m->k_on[d][b];
      s->koffT[t] = m->k_off[d][b];
      s->totT[t] = m->tot[d][b];
      s->bsp0[d][b] = m->buffer_species[d][b][0];
      s->bsp1[d][b] = m->buffer_species[d][b][1];
      isbuf[m->buffer_species[d][b][0]] = 1;
    }
  for (uint32_t sp = 0; sp < s->ns; sp++) s->nbuf += isbuf[sp];
  s->D = (uint32_t)fmax(1., floor(opt->dt_update_stoch / s->dt));
  s->R = (uint32_t)fmax(1., floor(opt->dt_update_react / s->dt));
  memcpy(s->init, m->init, sizeof s->init);
  s->dirichlet = opt->force_dirichlet;
  for (uint32_t i = 0; i < m->num_fixed; i++) s->fixed[m->fixed[2 * i]][m->fixed[2 * i + 1]] = 1;
  s->rng = (int)opt->rng;

  /* fields: Species_init_values + Species_init_du (species.c:887-935, :22-32) */
  for (uint32_t sp = 0; sp < s->ns; sp++) {
    s->u1[sp] = (double*)malloc(sizeof(double) * s->nv);
    s->du[sp] = (double*)malloc(sizeof(double) * s->nv);
    for (uint64_t v = 0; v < s->nv; v++) s->u1[sp][v] = s->init[s->ids[v]][sp];
    memcpy(s->du[sp], s->u1[sp], sizeof(double) * s->nv);
  }

  /* boundaries (geometry.c:576-1011, species.c:938-1035) */
  s->nb_geom = g->num_boundaries;
  int gate = 0;
  const char* mnames[3] = {m->ryr.boundary, m->p_ryr.boundary, m->serca.boundary};
  int muse[3] = {(int)m->use_ryr, (int)m->use_p_ryr, (int)m->use_serca};
  for (uint32_t b = 0; b < g->num_boundaries; b++) {
    obnd* B = &s->bnd[b];
    const uint16_t* P = g->boundaries[b];
    uint64_t nb = g->boundary_sizes[b];
    B->discrete = g->boundary_types[b] == CS_BOUNDARY_DISCRETE;
    B->model = -1;
    for (int k = 0; k < 3; k++)
      if (muse[k] && !strncmp(mnames[k], g->boundary_names[b], CS_NAME_LEN)) { B->model = k; break; }
    int outer = nb > 0 && P[0] == P[3] && P[1] == P[4] && P[2] == P[5];
    B->inner = !outer;
    uint32_t sb = B->discrete ? 1 : s->sub;
    B->npairs = (uint32_t)(nb * sb * sb);
    B->from = (uint64_t*)malloc(sizeof(uint64_t) * (B->npairs + 1));
    B->to = (uint64_t*)malloc(sizeof(uint64_t) * (B->npairs + 1));
    uint64_t c = 0;
    for (uint64_t j = 0; j < nb; j++) {
      const uint16_t* p = P + 6 * j;
      int dim2 = 0;
      for (int a = 2; a >= 0; a--)
        if (p[a] != p[a + 3]) dim2 = p[a] < p[a + 3] ? 2 * a : 2 * a + 1;
      int dim = dim2 / 2, lr = dim2 % 2;
      int od = dim != 0 ? 0 : 1, id = dim != 2 ? 2 : 1;
      uint64_t fi = (uint64_t)p[dim] * s->sub + (lr == 0 ? s->sub - 1 : 0);
      uint64_t ti = (uint64_t)p[dim + 3] * s->sub + (lr == 0 ? 0 : s->sub - 1);
      for (uint32_t oi = 0; oi < sb; oi++)
        for (uint32_t ii = 0; ii < sb; ii++) {
          uint64_t f[3], t[3];
          f[dim] = fi; t[dim] = ti;
          f[od] = t[od] = (uint64_t)s->sub * p[od] + oi + (B->discrete ? s->sub / 2 : 0);
          f[id] = t[id] = (uint64_t)s->sub * p[id] + ii + (B->discrete ? s->sub / 2 : 0);
          B->from[c] = IDX(f[0], f[1], f[2]);
          B->to[c] = IDX(t[0], t[1], t[2]);
          c++;
        }
    }
    B->gate_group = -1;
    if (B->discrete) {
      s->disc_b[s->ndisc++] = b;
      B->open = (uint8_t*)calloc(B->npairs + 1, 1);
      B->ever = (uint8_t*)calloc(B->npairs + 1, 1);
      if (B->model >= 0) B->gate_group = gate++;
    }
    if (B->model == 0 || B->model == 1) {
      const cs_channel_model* cm = B->model == 0 ? &m->ryr : &m->p_ryr;
      const char* lit = B->model == 0 ? "ryr" : "p_ryr";
      double tc = -1;
      for (uint32_t i = 0; i < opt->num_t_close; i++)
        if (!strncmp(opt->t_close[i].name, lit, CS_NAME_LEN)) { tc = opt->t_close[i].time; break; }
      double* pp = B->params;
      pp[0] = cm->Kd_close; pp[1] = cm->k_max_close; pp[2] = cm->n_open; pp[3] = cm->k_max_open;
      pp[4] = cm->k_min_open; pp[5] = cm->n_close; pp[6] = cm->i_unitary; pp[7] = cm->k_min_close;
      pp[8] = tc; pp[9] = cm->Kd_open;
    } else if (B->model == 2) {
      B->params[0] = m->serca.v_cyt_to_A_sr; B->params[1] = m->serca.density; B->params[2] = m->serca.scale;
    }
    if (B->model >= 0) {
      const char* spn = B->model == 0 ? m->ryr.species : B->model == 1 ? m->p_ryr.species : m->serca.species;
      int sp = name_index(m->species_names, s->ns, spn);
      B->dsi = -1;
      for (uint32_t k = 0; k < s->ndiff; k++) if ((int)s->alldiff[k] == sp) B->dsi = (int)k;
      if (B->dsi < 0) { oracle_free(s); return ofail(err, elen, "flux species not diffusive"); }
    }
  }

  /* seeding and initial channel states (calcium_sparks.c:113-116,
   * boundaryfluxes.c:229-293, :544-605) */
  s->seed = opt->seed_given ? opt->seed : 12345u;
  s->mt = (mt_state*)calloc(1, sizeof(mt_state));
  if (!s->mt) { oracle_free(s); return ofail(err, elen, "out of memory"); }
  mt_seed(s->mt, s->seed);
  if (m->use_ryr && s->ndisc > 0) {
    int rb = -1, pb = -1;
    for (uint32_t b = 0; b < s->nb_geom; b++) {
      if (s->bnd[b].model == 0) rb = (int)b;
      if (s->bnd[b].model == 1) pb = (int)b;
    }
    for (int which = 0; which < 2; which++) {
      int b = which == 0 ? rb : pb;
      if (b < 0 || !s->bnd[b].discrete) continue;
      const char* bname = which == 0 ? m->ryr.boundary : m->p_ryr.boundary;
      const cs_open_entry* info = NULL;
      for (uint32_t i = 0; i < opt->num_open; i++)
        if (!strncmp(opt->open[i].name, bname, CS_NAME_LEN)) { info = &opt->open[i]; break; }
      obnd* B = &s->bnd[b];
      if (info) {
        memset(B->open, 0, B->npairs);
        for (uint32_t j = 0; j < info->count; j++) B->open[info->idx[j]] = 1;
      } else {
        /* values of the ryr boundary (boundaryfluxes.c:578-597) */
        obnd* RB = &s->bnd[rb];
        const double* pp = B->params;
        for (uint32_t i = 0; i < B->npairs && i < RB->npairs; i++) {
          double c0 = s->u1[0][RB->to[i]];
          double ko = fmax(fmin(pow(c0 / pp[9], pp[2]), pp[3]), pp[4]);
          double kc = fmax(fmin(pow(c0 / pp[0], pp[5]), pp[1]), pp[7]);
          double r = mt_ldrand(s->mt);
          B->open[i] = (r <= kc / (ko + kc)) ? 0 : 1;
        }
      }
    }
  }
  for (uint32_t k = 0; k < s->ndisc; k++) {
    obnd* B = &s->bnd[s->disc_b[k]];
    memcpy(B->ever, B->open, B->npairs);
  }
  *out = s;
  return 0;
}

void oracle_get_info(const oracle_sim* s, oracle_info* I)
{
  memset(I, 0, sizeof(*I));
  for (int a = 0; a < 3; a++) I->n[a] = s->n[a];
  I->subdivisions = s->sub;
  I->h = s->h; I->dt = s->dt; I->dV = s->dV;
  I->num_species = s->ns; I->num_domains = s->nd;
  for (uint32_t d = 0; d < s->ndiff; d++) {
    I->substeps[s->alldiff[d]] = s->k[d];
    memcpy(I->coef[s->alldiff[d]], s->coef[d], sizeof(double) * 25);
  }
  I->num_all_buffers = s->nbuf;
  I->reaction_substep = s->R;
  I->stochastic_substep = s->D;
  I->num_discrete = s->ndisc;
  for (uint32_t k = 0; k < s->ndisc; k++) {
    I->discrete_size[k] = s->bnd[s->disc_b[k]].npairs;
    I->discrete_gated[k] = s->bnd[s->disc_b[k]].gate_group;
  }
  I->seed = s->seed;
}

/* ------------------------------------------------------------ stepping */
void oracle_stencil(const double* u, double* du, const double* C, const uint8_t* ids,
                    uint64_t nx, uint64_t ny, uint64_t nz)
{
  const uint64_t sx = ny * nz, sy = nz;
  #pragma omp parallel for collapse(2) schedule(static)
  for (uint64_t x = 0; x < nx; x++)
    for (uint64_t y = 0; y < ny; y++)
      for (uint64_t z = 0; z < nz; z++) {
        uint64_t v = x * sx + y * sy + z;
        double uv = u[v];
        const double* c = C + ids[v] * 5;
        double acc = uv;
        /* scalar interior order of species.c:1583-1607: z, then y, then x;
         * neighbours outside the grid are omitted (no-flux faces) */
        if (z > 0)      acc += c[ids[v - 1]] * (u[v - 1] - uv);
        if (z + 1 < nz) acc += c[ids[v + 1]] * (u[v + 1] - uv);
        if (y > 0)      acc += c[ids[v - sy]] * (u[v - sy] - uv);
        if (y + 1 < ny) acc += c[ids[v + sy]] * (u[v + sy] - uv);
        if (x > 0)      acc += c[ids[v - sx]] * (u[v - sx] - uv);
        if (x + 1 < nx) acc += c[ids[v + sx]] * (u[v + sx] - uv);
        du[v] = acc;
      }
}

/* evaluate_RyR_stochastically, boundaryfluxes.c:295-359 */
static void gate_boundary(oracle_sim* s, obnd* B, uint64_t ti)
{
  const double* p = B->params;
  double t = ti * s->dt, dts = s->dt * s->D;
  for (uint32_t i = 0; i < B->npairs; i++) {
    if (p[8] > 0) {
      if (p[8] <= t + dts) B->open[i] = 0;
      continue;
    }
    double c0 = s->u1[0][B->to[i]];   /* Ca (slot 0) at the to voxel, species.c:2569 */
    double kc = fmax(fmin(pow(c0 / p[0], p[5]), p[1]), p[7]);
    double ko = fmax(fmin(pow(c0 / p[9], p[2]), p[3]), p[4]);
    double r = s->rng == CS_RNG_MT ? mt_ldrand(s->mt)
                                   : philox_uniform(s->seed, ti, (uint32_t)B->gate_group, i);
    if (B->open[i] == 0) { if (r <= dts * ko) B->open[i] = 1; }
    else if (B->open[i] == 1) { if (r <= dts * kc) B->open[i] = 0; }
  }
}

int oracle_step(oracle_sim* s, uint64_t ti)
{
  /* Species_step_diffusion (species.c:1501-1650) */
  for (uint32_t dsi = 0; dsi < s->ndiff; dsi++) {
    if (ti % s->k[dsi] != 0) continue;
    uint32_t sp = s->alldiff[dsi];
    s->upd[sp] = 1;
    oracle_stencil(s->u1[sp], s->du[sp], s->coef[dsi], s->ids, s->n[0], s->n[1], s->n[2]);
  }
  /* Species_evaluate_stochastic_events (species.c:1910-1957) */
  if (s->ndisc > 0 && ti % s->D == 0)
    for (uint32_t k = 0; k < s->ndisc; k++) {
      obnd* B = &s->bnd[s->disc_b[k]];
      if (B->model == 0 || B->model == 1) gate_boundary(s, B, ti);
      for (uint32_t i = 0; i < B->npairs; i++) B->ever[i] |= B->open[i];
    }
  /* Species_step_boundary_fluxes (species.c:1704-1807) */
  for (uint32_t b = 0; b < s->nb_geom; b++) {
    obnd* B = &s->bnd[b];
    if (B->model < 0) continue;
    if (ti % s->k[B->dsi] != 0) continue;
    uint32_t sp = s->alldiff[B->dsi];
    s->upd[sp] = 1;
    double* u = s->u1[sp];
    double* du = s->du[sp];
    for (uint32_t i = 0; i < B->npairs; i++) {
      if (B->discrete && !B->open[i]) continue;
      double J = B->model == 2 ? oracle_serca_flux(s->dt, s->h, u[B->from[i]], u[B->to[i]], B->params)
                               : oracle_ryr_flux(s->dt, s->h, u[B->from[i]], u[B->to[i]], B->params);
      du[B->to[i]] += J;
      if (B->inner) du[B->from[i]] -= J;
    }
  }
  /* Species_step_reaction (species.c:1653-1702), scalar form of
   * reaction() in kernel_generic.cpp:397-470 */
  if (s->nbuf > 0 && ti % s->R == 0) {
    const double dt = s->dt * s->R;
    const uint32_t nd = s->nd;
    #pragma omp parallel for schedule(static)
    for (uint64_t v = 0; v < s->nv; v++) {
      uint32_t di = s->ids[v];
      for (uint32_t bi = 0; bi < s->nbuf; bi++) {
        uint32_t t = bi * nd + di;
        uint32_t b0 = s->bsp0T[t], b1 = s->bsp1T[t];
        double B0 = s->u1[b0][v], B1 = s->u1[b1][v];
        double J = s->konT[t] * (s->totT[t] - B0) * B1 - s->koffT[t] * B0;
        if (s->upd[b0]) s->du[b0][v] += dt * J; else s->du[b0][v] = B0 + dt * J;
        if (s->upd[b1]) s->du[b1][v] -= dt * J; else s->du[b1][v] = B1 - dt * J;
      }
    }
    for (uint32_t di = 0; di < nd; di++)
      for (uint32_t bi = 0; bi < s->nbuf; bi++) {
        s->upd[s->bsp0[di][bi]] = 1;
        s->upd[s->bsp1[di][bi]] = 1;
      }
  }
  /* Species_apply_du (species.c:1080-1121) */
  for (uint32_t sp = 0; sp < s->ns; sp++) {
    if (!s->upd[sp]) continue;
    double* dst = s->u1[sp];
    const double* src = s->du[sp];
    #pragma omp parallel for schedule(static)
    for (uint64_t v = 0; v < s->nv; v++) dst[v] = src[v];
    s->upd[sp] = 0;
  }
  /* Species_apply_fixed_values (species.c:1046-1077) */
  if (s->dirichlet) {
    for (uint64_t x = 0; x < s->n[0]; x++)
      for (uint64_t y = 0; y < s->n[1]; y++)
        for (uint64_t z = 0; z < s->n[2]; z++) {
          if (!(x == 0 || y == 0 || z == 0 || x + 1 == s->n[0] || y + 1 == s->n[1] || z + 1 == s->n[2])) continue;
          uint64_t v = IDX(x, y, z);
          s->u1[0][v] = s->init[s->ids[v]][0];
        }
  }
  for (uint32_t sp = 0; sp < s->ns; sp++) {
    int any = 0;
    for (uint32_t d = 0; d < s->nd; d++) any |= s->fixed[d][sp] != 0;
    if (!any) continue;
    #pragma omp parallel for schedule(static)
    for (uint64_t v = 0; v < s->nv; v++)
      if (s->fixed[s->ids[v]][sp]) s->u1[sp][v] = s->init[s->ids[v]][sp];
  }
  return 0;
}

int oracle_steps(oracle_sim* s, uint64_t first, uint64_t n)
{
  for (uint64_t t = first; t < first + n; t++) oracle_step(s, t);
  return 0;
}

void oracle_get_species(const oracle_sim* s, uint32_t sp, double* out)
{
  memcpy(out, s->u1[sp], sizeof(double) * s->nv);
}

void oracle_set_species(oracle_sim* s, uint32_t sp, const double* in)
{
  memcpy(s->u1[sp], in, sizeof(double) * s->nv);
  memcpy(s->du[sp], in, sizeof(double) * s->nv);
}

/* gauss, utils.c:209-213; the constants are 1/sqrt(2) and 2/sqrt(pi)
 * (math.h's M_SQRT1_2 and M_2_SQRTPI, hidden under -std=c11) */
static double oracle_gauss(double x, double sigma)
{
  const double sqrt1_2 = 0.70710678118654752440, two_sqrtpi = 1.12837916709551257390;
  const double inv_sqrt_sigma = 1.0 / sqrt(sigma);
  return sqrt1_2 * two_sqrtpi * inv_sqrt_sigma / 2 * exp(-x * x / 2 / sigma);
}

/* Linescan, species.c:1960-2035 with the offsets of species.c:856-877:
 * for every index p along `axis`, the sum over all voxels of
 * u * gauss(o_x - x) * gauss(o_y - y) * gauss(o_z - z), where a coordinate
 * is index * h, o = (N/2 + offset) * h on the two other axes and p * h on
 * `axis`.  With num_domains > 0 a voxel counts only if its domain id is
 * below num_domains: the reference compares the id with the loop counter
 * over the listed domains, not with the listed ids (species.c:2015-2024).
 * Evaluated directly, O(N[axis] * voxels). */
void oracle_linescan(const oracle_sim* s, uint32_t sp, uint32_t axis, const int32_t offsets[3],
                     uint32_t num_domains, const double conv[3], double* out)
{
  double off[3];
  for (int ax = 0; ax < 3; ax++) off[ax] = (double)((int64_t)(s->n[ax] / 2) + offsets[ax]) * s->h;
  const double* u = s->u1[sp];
  for (uint64_t p = 0; p < s->n[axis]; p++) {
    off[axis] = (double)p * s->h;
    double sum = 0;
    for (uint64_t x = 0; x < s->n[0]; x++)
      for (uint64_t y = 0; y < s->n[1]; y++)
        for (uint64_t z = 0; z < s->n[2]; z++) {
          const uint64_t v = (x * s->n[1] + y) * s->n[2] + z;
          if (num_domains && s->ids[v] >= num_domains) continue;
          sum += u[v] * oracle_gauss(off[0] - (double)x * s->h, conv[0]) *
                 oracle_gauss(off[1] - (double)y * s->h, conv[1]) *
                 oracle_gauss(off[2] - (double)z * s->h, conv[2]);
        }
    out[p] = sum;
  }
}

void oracle_stats(const oracle_sim* s, double* mn, double* mx, double* tot)
{
  uint32_t nds = s->nd * s->ns;
  for (uint32_t i = 0; i < nds; i++) { mn[i] = FLT_MAX; mx[i] = FLT_MIN; tot[i] = 0; }
  for (uint32_t sp = 0; sp < s->ns; sp++)
    for (uint64_t v = 0; v < s->nv; v++) {
      uint32_t k = s->ids[v] * s->ns + sp;
      double val = s->u1[sp][v];
      mn[k] = val < mn[k] ? val : mn[k];
      mx[k] = val > mx[k] ? val : mx[k];
      tot[k] += val * s->dV;
    }
}

int oracle_get_states(const oracle_sim* s, uint32_t k, uint8_t* out)
{
  if (k >= s->ndisc) return -1;
  const obnd* B = &s->bnd[s->disc_b[k]];
  memcpy(out, B->open, B->npairs);
  return 0;
}

int oracle_set_states(oracle_sim* s, uint32_t k, const uint8_t* in)
{
  if (k >= s->ndisc) return -1;
  obnd* B = &s->bnd[s->disc_b[k]];
  memcpy(B->open, in, B->npairs);
  for (uint32_t i = 0; i < B->npairs; i++) B->ever[i] |= B->open[i];
  return 0;
}

/* Channels of discrete boundary k open at the start or after any gating
 * step so far: the RyRs a spark recruited (a statistic of this project;
 * the reference reports none). */
int oracle_get_recruited(const oracle_sim* s, uint32_t k, uint8_t* out)
{
  if (k >= s->ndisc) return -1;
  const obnd* B = &s->bnd[s->disc_b[k]];
  memcpy(out, B->ever, B->npairs);
  return 0;
}
