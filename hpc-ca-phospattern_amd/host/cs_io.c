/*
 * HDF5 reader/writer for the calcium_sparks file formats.
 *
 * Attribute types follow the shipped files: counts and flags are uint8,
 * parameters float64, names fixed-length strings; the reader converts to
 * the in-memory types with HDF5's own conversion, so scalar and
 * one-element attributes both work.  Missing attributes are errors with
 * the reference's message (utils.c:77-99).
 */
#include <hdf5.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "cs_io.h"

static int ioerr(char* err, size_t len, const char* fmt, ...)
{
  if (err && len) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(err, len, fmt, ap);
    va_end(ap);
  }
  return -1;
}

static void quiet_hdf5(void) { H5Eset_auto2(H5E_DEFAULT, NULL, NULL); }

/* ------------------------------------------------------------ reading */
static int rd(hid_t loc, const char* obj, const char* name, hid_t mtype, void* buf, char* err, size_t elen)
{
  hid_t a = H5Aopen_by_name(loc, obj, name, H5P_DEFAULT, H5P_DEFAULT);
  if (a < 0) return ioerr(err, elen, "Could not open attr \"%s%s\"", obj, name);
  herr_t rc = H5Aread(a, mtype, buf);
  H5Aclose(a);
  if (rc < 0) return ioerr(err, elen, "Could not read attr \"%s%s\"", obj, name);
  return 0;
}

static int rd_count(hid_t loc, const char* obj, const char* name, hsize_t* n)
{
  hid_t a = H5Aopen_by_name(loc, obj, name, H5P_DEFAULT, H5P_DEFAULT);
  if (a < 0) return -1;
  hid_t sp = H5Aget_space(a);
  hssize_t k = H5Sget_simple_extent_npoints(sp);
  H5Sclose(sp);
  H5Aclose(a);
  *n = k < 0 ? 0 : (hsize_t)k;
  return 0;
}

static int rd_str(hid_t loc, const char* obj, const char* name, char* out, size_t outlen, char* err, size_t elen)
{
  hid_t a = H5Aopen_by_name(loc, obj, name, H5P_DEFAULT, H5P_DEFAULT);
  if (a < 0) return ioerr(err, elen, "Could not open attr \"%s%s\"", obj, name);
  hid_t t = H5Aget_type(a);
  int rc = 0;
  if (H5Tget_class(t) != H5T_STRING) {
    rc = ioerr(err, elen, "attr \"%s%s\" is not a string", obj, name);
  } else if (H5Tis_variable_str(t) > 0) {
    char* s = NULL;
    hid_t mt = H5Tcopy(H5T_C_S1);
    H5Tset_size(mt, H5T_VARIABLE);
    if (H5Aread(a, mt, &s) < 0 || !s) rc = ioerr(err, elen, "Could not read attr \"%s%s\"", obj, name);
    else {
      snprintf(out, outlen, "%s", s);
      H5free_memory(s);
    }
    H5Tclose(mt);
  } else {
    size_t sz = H5Tget_size(t);
    char* buf = (char*)calloc(sz + 1, 1);
    if (H5Aread(a, t, buf) < 0) rc = ioerr(err, elen, "Could not read attr \"%s%s\"", obj, name);
    else {
      buf[sz] = 0;
      if (strlen(buf) >= outlen) rc = ioerr(err, elen, "attr \"%s%s\" too long", obj, name);
      else snprintf(out, outlen, "%s", buf);
    }
    free(buf);
  }
  H5Tclose(t);
  H5Aclose(a);
  return rc;
}

void cs_io_free_geometry(cs_geometry* g)
{
  free((void*)g->voxels);
  g->voxels = NULL;
  for (int i = 0; i < CS_MAX_BOUNDARIES; i++) {
    free((void*)g->boundaries[i]);
    g->boundaries[i] = NULL;
  }
}

int cs_io_read_geometry(const char* path, cs_geometry* g, char* err, size_t elen)
{
  quiet_hdf5();
  memset(g, 0, sizeof *g);
  hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) return ioerr(err, elen, "*** ERROR: cannot open geometry file %s", path);
  int rc = -1;
  hid_t dset = -1, sp = -1;
  uint32_t num = 0;
  char name[64];
  if (rd(f, "/", "global_size", H5T_NATIVE_DOUBLE, g->global_size, err, elen)) goto out;
  if (rd(f, "/", "h", H5T_NATIVE_DOUBLE, &g->h, err, elen)) goto out;
  if (rd(f, "domains", "num", H5T_NATIVE_UINT32, &num, err, elen)) goto out;
  if (num == 0 || num > CS_MAX_DOMAINS) { ioerr(err, elen, "*** ERROR: unsupported number of domains %u", num); goto out; }
  g->num_domains = num;
  for (uint32_t i = 0; i < num; i++) {
    snprintf(name, sizeof name, "name_%u", i);
    if (rd_str(f, "domains", name, g->domain_names[i], CS_NAME_LEN, err, elen)) goto out;
  }
  if (rd(f, "domains", "indices", H5T_NATIVE_UINT8, g->domain_indices, err, elen)) goto out;
  /* voxels */
  dset = H5Dopen2(f, "/domains/voxels", H5P_DEFAULT);
  if (dset < 0) { ioerr(err, elen, "*** ERROR: no /domains/voxels"); goto out; }
  sp = H5Dget_space(dset);
  if (H5Sget_simple_extent_ndims(sp) != 3) { ioerr(err, elen, "*** ERROR: voxels must be 3-D"); goto out; }
  hsize_t dims[3];
  H5Sget_simple_extent_dims(sp, dims, NULL);
  for (int a = 0; a < 3; a++) g->N[a] = dims[a];
  uint8_t* vox = (uint8_t*)malloc(dims[0] * dims[1] * dims[2]);
  if (!vox) { ioerr(err, elen, "*** ERROR: out of memory"); goto out; }
  if (H5Dread(dset, H5T_NATIVE_UINT8, H5S_ALL, H5S_ALL, H5P_DEFAULT, vox) < 0) {
    free(vox);
    ioerr(err, elen, "*** ERROR: cannot read voxels");
    goto out;
  }
  g->voxels = vox;
  H5Sclose(sp); sp = -1;
  H5Dclose(dset); dset = -1;
  /* domain connections (geometry.c:549-574) */
  dset = H5Dopen2(f, "/domains/domain_connections", H5P_DEFAULT);
  if (dset < 0) { ioerr(err, elen, "*** ERROR: no /domains/domain_connections"); goto out; }
  sp = H5Dget_space(dset);
  hsize_t cd[2] = {0, 0};
  H5Sget_simple_extent_dims(sp, cd, NULL);
  if (cd[0] != num || cd[1] != num) { ioerr(err, elen, "*** ERROR: domain_connections is not %ux%u", num, num); goto out; }
  if (H5Dread(dset, H5T_NATIVE_UINT8, H5S_ALL, H5S_ALL, H5P_DEFAULT, g->domain_connections) < 0) {
    ioerr(err, elen, "*** ERROR: cannot read domain_connections");
    goto out;
  }
  H5Sclose(sp); sp = -1;
  H5Dclose(dset); dset = -1;
  /* boundaries */
  if (rd(f, "boundaries", "num", H5T_NATIVE_UINT32, &num, err, elen)) goto out;
  if (num > CS_MAX_BOUNDARIES) { ioerr(err, elen, "*** ERROR: too many boundaries"); goto out; }
  g->num_boundaries = num;
  for (uint32_t i = 0; i < num; i++) {
    char type[CS_NAME_LEN];
    snprintf(name, sizeof name, "name_%u", i);
    if (rd_str(f, "boundaries", name, g->boundary_names[i], CS_NAME_LEN, err, elen)) goto out;
    snprintf(name, sizeof name, "type_%u", i);
    if (rd_str(f, "boundaries", name, type, CS_NAME_LEN, err, elen)) goto out;
    /* Geometry_boundary_type_from_str, geometry.c:1438-1446 */
    if (!strcmp(type, "density")) g->boundary_types[i] = CS_BOUNDARY_DENSITY;
    else if (!strcmp(type, "discrete")) g->boundary_types[i] = CS_BOUNDARY_DISCRETE;
    else { ioerr(err, elen, "*** ERROR: \"%s\" is not a valid boundary type.", type); goto out; }
    char dpath[96];
    snprintf(dpath, sizeof dpath, "/boundaries/%s", g->boundary_names[i]);
    dset = H5Dopen2(f, dpath, H5P_DEFAULT);
    if (dset < 0) { ioerr(err, elen, "*** ERROR: no dataset %s", dpath); goto out; }
    sp = H5Dget_space(dset);
    hsize_t bd[2] = {0, 0};
    int nd = H5Sget_simple_extent_ndims(sp);
    H5Sget_simple_extent_dims(sp, bd, NULL);
    if (nd != 2 || bd[1] != 6) { ioerr(err, elen, "*** ERROR: %s must be [n][6]", dpath); goto out; }
    uint16_t* b = (uint16_t*)malloc(sizeof(uint16_t) * 6 * (bd[0] ? bd[0] : 1));
    if (bd[0] && H5Dread(dset, H5T_NATIVE_UINT16, H5S_ALL, H5S_ALL, H5P_DEFAULT, b) < 0) {
      free(b);
      ioerr(err, elen, "*** ERROR: cannot read %s", dpath);
      goto out;
    }
    g->boundaries[i] = b;
    g->boundary_sizes[i] = bd[0];
    H5Sclose(sp); sp = -1;
    H5Dclose(dset); dset = -1;
  }
  rc = 0;
out:
  if (sp >= 0) H5Sclose(sp);
  if (dset >= 0) H5Dclose(dset);
  H5Fclose(f);
  if (rc) cs_io_free_geometry(g);
  return rc;
}

static int rd_channel(hid_t f, const char* grp, cs_channel_model* c, char* err, size_t elen)
{
  struct { const char* n; double* p; } fields[] = {
      {"Kd_open", &c->Kd_open}, {"k_min_open", &c->k_min_open}, {"k_max_open", &c->k_max_open},
      {"n_open", &c->n_open}, {"Kd_close", &c->Kd_close}, {"k_min_close", &c->k_min_close},
      {"k_max_close", &c->k_max_close}, {"n_close", &c->n_close}, {"i_unitary", &c->i_unitary}};
  for (size_t i = 0; i < sizeof fields / sizeof fields[0]; i++)
    if (rd(f, grp, fields[i].n, H5T_NATIVE_DOUBLE, fields[i].p, err, elen)) return -1;
  if (rd_str(f, grp, "boundary", c->boundary, CS_NAME_LEN, err, elen)) return -1;
  if (rd_str(f, grp, "species", c->species, CS_NAME_LEN, err, elen)) return -1;
  return 0;
}

int cs_io_read_species(const char* path, cs_species_model* m, char* err, size_t elen)
{
  quiet_hdf5();
  memset(m, 0, sizeof *m);
  hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) return ioerr(err, elen, "*** ERROR: cannot open species file %s", path);
  int rc = -1;
  char name[64], grp[96];
  if (rd(f, "/", "num_domains", H5T_NATIVE_UINT32, &m->num_domains, err, elen)) goto out;
  if (rd(f, "/", "num_species", H5T_NATIVE_UINT32, &m->num_species, err, elen)) goto out;
  if (m->num_domains == 0 || m->num_domains > CS_MAX_DOMAINS || m->num_species == 0 ||
      m->num_species > CS_MAX_SPECIES) {
    ioerr(err, elen, "*** ERROR: unsupported domain/species counts %u/%u", m->num_domains, m->num_species);
    goto out;
  }
  for (uint32_t s = 0; s < m->num_species; s++) {
    snprintf(name, sizeof name, "species_name_%u", s);
    if (rd_str(f, "/", name, m->species_names[s], CS_NAME_LEN, err, elen)) goto out;
  }
  for (uint32_t d = 0; d < m->num_domains; d++) {
    snprintf(name, sizeof name, "domain_name_%u", d);
    if (rd_str(f, "/", name, m->domain_names[d], CS_NAME_LEN, err, elen)) goto out;
    const char* dn = m->domain_names[d];
    if (rd(f, dn, "num_diffusive", H5T_NATIVE_UINT32, &m->num_diffusive[d], err, elen)) goto out;
    if (m->num_diffusive[d] > m->num_species) { ioerr(err, elen, "*** ERROR: bad num_diffusive in %s", dn); goto out; }
    double init[CS_MAX_SPECIES];
    hsize_t cnt = 0;
    if (rd_count(f, dn, "init", &cnt) || cnt != m->num_species) { ioerr(err, elen, "*** ERROR: %s/init needs %u values", dn, m->num_species); goto out; }
    if (rd(f, dn, "init", H5T_NATIVE_DOUBLE, init, err, elen)) goto out;
    for (uint32_t s = 0; s < m->num_species; s++) m->init[d][s] = init[s];
    if (m->num_diffusive[d] > 0) {
      if (rd_count(f, dn, "diffusive", &cnt) || cnt != m->num_diffusive[d] ||
          rd_count(f, dn, "sigma", &cnt) || cnt != m->num_diffusive[d]) {
        ioerr(err, elen, "*** ERROR: %s/diffusive and sigma need %u values", dn, m->num_diffusive[d]);
        goto out;
      }
      if (rd(f, dn, "diffusive", H5T_NATIVE_UINT8, m->diffusive[d], err, elen)) goto out;
      if (rd(f, dn, "sigma", H5T_NATIVE_DOUBLE, m->sigma[d], err, elen)) goto out;
    }
    if (rd(f, dn, "num_buffers", H5T_NATIVE_UINT32, &m->num_buffers[d], err, elen)) goto out;
    if (m->num_buffers[d] > m->num_species) { ioerr(err, elen, "*** ERROR: bad num_buffers in %s", dn); goto out; }
    for (uint32_t b = 0; b < m->num_buffers[d]; b++) {
      snprintf(grp, sizeof grp, "/%s/buffer_%u/", dn, b);
      if (rd(f, grp, "k_off", H5T_NATIVE_DOUBLE, &m->k_off[d][b], err, elen)) goto out;
      if (rd(f, grp, "k_on", H5T_NATIVE_DOUBLE, &m->k_on[d][b], err, elen)) goto out;
      if (rd(f, grp, "tot", H5T_NATIVE_DOUBLE, &m->tot[d][b], err, elen)) goto out;
      if (rd(f, grp, "species", H5T_NATIVE_UINT8, m->buffer_species[d][b], err, elen)) goto out;
    }
  }
  if (rd(f, "/", "num_fixed_domain_species", H5T_NATIVE_UINT32, &m->num_fixed, err, elen)) goto out;
  if (m->num_fixed > CS_MAX_DOMAINS * CS_MAX_SPECIES) { ioerr(err, elen, "*** ERROR: too many fixed domain species"); goto out; }
  if (m->num_fixed && rd(f, "/", "fixed_domain_species", H5T_NATIVE_UINT8, m->fixed, err, elen)) goto out;
  if (rd(f, "/", "use_ryr", H5T_NATIVE_UINT32, &m->use_ryr, err, elen)) goto out;
  if (rd(f, "/", "use_p_ryr", H5T_NATIVE_UINT32, &m->use_p_ryr, err, elen)) goto out;
  if (rd(f, "/", "use_serca", H5T_NATIVE_UINT32, &m->use_serca, err, elen)) goto out;
  if (m->use_ryr && rd_channel(f, "/ryr", &m->ryr, err, elen)) goto out;
  if (m->use_p_ryr && rd_channel(f, "/p_ryr", &m->p_ryr, err, elen)) goto out;
  if (m->use_serca) {
    if (rd(f, "/serca", "v_cyt_to_A_sr", H5T_NATIVE_DOUBLE, &m->serca.v_cyt_to_A_sr, err, elen)) goto out;
    if (rd(f, "/serca", "density", H5T_NATIVE_DOUBLE, &m->serca.density, err, elen)) goto out;
    if (rd(f, "/serca", "scale", H5T_NATIVE_DOUBLE, &m->serca.scale, err, elen)) goto out;
    if (rd_str(f, "/serca", "boundary", m->serca.boundary, CS_NAME_LEN, err, elen)) goto out;
    /* boundaryfluxes.c:518 reads the SERCA species from /ryr */
    if (rd_str(f, "/ryr", "species", m->serca.species, CS_NAME_LEN, err, elen)) goto out;
  }
  if (rd(f, "/", "convolution_constants", H5T_NATIVE_DOUBLE, m->conv, err, elen)) goto out;
  rc = 0;
out:
  H5Fclose(f);
  return rc;
}

/* ------------------------------------------------------------ writing */
static int wr_scalar(hid_t loc, const char* name, hid_t ftype, hid_t mtype, const void* v)
{
  hid_t sp = H5Screate(H5S_SCALAR);
  hid_t a = H5Acreate2(loc, name, ftype, sp, H5P_DEFAULT, H5P_DEFAULT);
  herr_t rc = a < 0 ? -1 : H5Awrite(a, mtype, v);
  if (a >= 0) H5Aclose(a);
  H5Sclose(sp);
  return rc < 0 ? -1 : 0;
}

static int wr_array(hid_t loc, const char* name, hid_t ftype, hid_t mtype, const void* v, hsize_t n)
{
  hid_t sp = H5Screate_simple(1, &n, NULL);
  hid_t a = H5Acreate2(loc, name, ftype, sp, H5P_DEFAULT, H5P_DEFAULT);
  herr_t rc = a < 0 ? -1 : (n ? H5Awrite(a, mtype, v) : 0);
  if (a >= 0) H5Aclose(a);
  H5Sclose(sp);
  return rc < 0 ? -1 : 0;
}

/* fixed-length string attribute; pad NULLPAD like the shipped inputs or
 * NULLTERM like the reference's own writer (utils.c:101-138) */
static int wr_str(hid_t loc, const char* name, const char* s, int nullterm)
{
  size_t len = strlen(s);
  hid_t t = H5Tcopy(H5T_C_S1);
  H5Tset_size(t, len ? len : 1);
  H5Tset_strpad(t, nullterm ? H5T_STR_NULLTERM : H5T_STR_NULLPAD);
  hid_t sp = H5Screate(H5S_SCALAR);
  hid_t a = H5Acreate2(loc, name, t, sp, H5P_DEFAULT, H5P_DEFAULT);
  herr_t rc = a < 0 ? -1 : H5Awrite(a, t, s);
  if (a >= 0) H5Aclose(a);
  H5Sclose(sp);
  H5Tclose(t);
  return rc < 0 ? -1 : 0;
}

static hid_t mkgroup(hid_t loc, const char* name)
{
  return H5Gcreate2(loc, name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
}

int cs_io_write_geometry(const char* path, const cs_geometry* g, char* err, size_t elen)
{
  quiet_hdf5();
  hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (f < 0) return ioerr(err, elen, "cannot create %s", path);
  int bad = 0;
  char name[64];
  bad |= wr_array(f, "global_size", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, g->global_size, 3);
  bad |= wr_scalar(f, "h", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, &g->h);
  hid_t dg = mkgroup(f, "domains");
  uint8_t nd = (uint8_t)g->num_domains;
  bad |= wr_scalar(dg, "num", H5T_STD_U8LE, H5T_NATIVE_UINT8, &nd);
  for (uint32_t i = 0; i < g->num_domains; i++) {
    snprintf(name, sizeof name, "name_%u", i);
    bad |= wr_str(dg, name, g->domain_names[i], 0);
  }
  bad |= wr_array(dg, "indices", H5T_STD_U8LE, H5T_NATIVE_UINT8, g->domain_indices, g->num_domains);
  hsize_t dims[3] = {g->N[0], g->N[1], g->N[2]};
  hid_t sp = H5Screate_simple(3, dims, NULL);
  hid_t pl = H5Pcreate(H5P_DATASET_CREATE);
  hsize_t chunk[3] = {dims[0] < 64 ? dims[0] : 64, dims[1] < 64 ? dims[1] : 64, dims[2] < 64 ? dims[2] : 64};
  H5Pset_chunk(pl, 3, chunk);
  H5Pset_deflate(pl, 4);
  hid_t ds = H5Dcreate2(dg, "voxels", H5T_STD_U8LE, sp, H5P_DEFAULT, pl, H5P_DEFAULT);
  bad |= ds < 0 || H5Dwrite(ds, H5T_NATIVE_UINT8, H5S_ALL, H5S_ALL, H5P_DEFAULT, g->voxels) < 0;
  if (ds >= 0) H5Dclose(ds);
  H5Pclose(pl);
  H5Sclose(sp);
  hsize_t cd[2] = {g->num_domains, g->num_domains};
  uint8_t conn[CS_MAX_DOMAINS * CS_MAX_DOMAINS];
  for (uint32_t i = 0; i < g->num_domains * g->num_domains; i++) conn[i] = g->domain_connections[i];
  sp = H5Screate_simple(2, cd, NULL);
  ds = H5Dcreate2(dg, "domain_connections", H5T_STD_U8LE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  bad |= ds < 0 || H5Dwrite(ds, H5T_NATIVE_UINT8, H5S_ALL, H5S_ALL, H5P_DEFAULT, conn) < 0;
  if (ds >= 0) H5Dclose(ds);
  H5Sclose(sp);
  H5Gclose(dg);
  hid_t bg = mkgroup(f, "boundaries");
  uint8_t nb = (uint8_t)g->num_boundaries;
  bad |= wr_scalar(bg, "num", H5T_STD_U8LE, H5T_NATIVE_UINT8, &nb);
  for (uint32_t i = 0; i < g->num_boundaries; i++) {
    snprintf(name, sizeof name, "name_%u", i);
    bad |= wr_str(bg, name, g->boundary_names[i], 0);
    snprintf(name, sizeof name, "type_%u", i);
    bad |= wr_str(bg, name, g->boundary_types[i] == CS_BOUNDARY_DISCRETE ? "discrete" : "density", 0);
    hsize_t bd[2] = {g->boundary_sizes[i], 6};
    sp = H5Screate_simple(2, bd, NULL);
    ds = H5Dcreate2(bg, g->boundary_names[i], H5T_STD_U16LE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    bad |= ds < 0 || (bd[0] && H5Dwrite(ds, H5T_NATIVE_UINT16, H5S_ALL, H5S_ALL, H5P_DEFAULT, g->boundaries[i]) < 0);
    if (ds >= 0) H5Dclose(ds);
    H5Sclose(sp);
  }
  H5Gclose(bg);
  H5Fclose(f);
  return bad ? ioerr(err, elen, "failed writing %s", path) : 0;
}

static int wr_channel(hid_t f, const char* gname, const cs_channel_model* c)
{
  hid_t g = mkgroup(f, gname);
  int bad = 0;
  struct { const char* n; double v; } fields[] = {
      {"Kd_open", c->Kd_open}, {"k_min_open", c->k_min_open}, {"k_max_open", c->k_max_open},
      {"n_open", c->n_open}, {"Kd_close", c->Kd_close}, {"k_min_close", c->k_min_close},
      {"k_max_close", c->k_max_close}, {"n_close", c->n_close}, {"i_unitary", c->i_unitary}};
  for (size_t i = 0; i < sizeof fields / sizeof fields[0]; i++)
    bad |= wr_scalar(g, fields[i].n, H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, &fields[i].v);
  bad |= wr_str(g, "boundary", c->boundary, 0);
  bad |= wr_str(g, "species", c->species, 0);
  H5Gclose(g);
  return bad;
}

int cs_io_write_species(const char* path, const cs_species_model* m, char* err, size_t elen)
{
  quiet_hdf5();
  hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (f < 0) return ioerr(err, elen, "cannot create %s", path);
  int bad = 0;
  char name[64];
  uint8_t u8;
  u8 = (uint8_t)m->num_domains; bad |= wr_scalar(f, "num_domains", H5T_STD_U8LE, H5T_NATIVE_UINT8, &u8);
  u8 = (uint8_t)m->num_species; bad |= wr_scalar(f, "num_species", H5T_STD_U8LE, H5T_NATIVE_UINT8, &u8);
  for (uint32_t d = 0; d < m->num_domains; d++) {
    snprintf(name, sizeof name, "domain_name_%u", d);
    bad |= wr_str(f, name, m->domain_names[d], 0);
  }
  for (uint32_t s = 0; s < m->num_species; s++) {
    snprintf(name, sizeof name, "species_name_%u", s);
    bad |= wr_str(f, name, m->species_names[s], 0);
  }
  for (uint32_t d = 0; d < m->num_domains; d++) {
    hid_t g = mkgroup(f, m->domain_names[d]);
    u8 = (uint8_t)m->num_diffusive[d]; bad |= wr_scalar(g, "num_diffusive", H5T_STD_U8LE, H5T_NATIVE_UINT8, &u8);
    bad |= wr_array(g, "diffusive", H5T_STD_U8LE, H5T_NATIVE_UINT8, m->diffusive[d], m->num_diffusive[d]);
    bad |= wr_array(g, "sigma", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, m->sigma[d], m->num_diffusive[d]);
    bad |= wr_array(g, "init", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, m->init[d], m->num_species);
    u8 = (uint8_t)m->num_buffers[d]; bad |= wr_scalar(g, "num_buffers", H5T_STD_U8LE, H5T_NATIVE_UINT8, &u8);
    for (uint32_t b = 0; b < m->num_buffers[d]; b++) {
      snprintf(name, sizeof name, "buffer_%u", b);
      hid_t bg = mkgroup(g, name);
      bad |= wr_scalar(bg, "k_off", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, &m->k_off[d][b]);
      bad |= wr_scalar(bg, "k_on", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, &m->k_on[d][b]);
      bad |= wr_scalar(bg, "tot", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, &m->tot[d][b]);
      bad |= wr_array(bg, "species", H5T_STD_U8LE, H5T_NATIVE_UINT8, m->buffer_species[d][b], 2);
      H5Gclose(bg);
    }
    H5Gclose(g);
  }
  u8 = (uint8_t)m->num_fixed; bad |= wr_scalar(f, "num_fixed_domain_species", H5T_STD_U8LE, H5T_NATIVE_UINT8, &u8);
  bad |= wr_array(f, "fixed_domain_species", H5T_STD_U8LE, H5T_NATIVE_UINT8, m->fixed, 2 * (hsize_t)m->num_fixed);
  u8 = (uint8_t)m->use_ryr; bad |= wr_scalar(f, "use_ryr", H5T_STD_U8LE, H5T_NATIVE_UINT8, &u8);
  u8 = (uint8_t)m->use_p_ryr; bad |= wr_scalar(f, "use_p_ryr", H5T_STD_U8LE, H5T_NATIVE_UINT8, &u8);
  u8 = (uint8_t)m->use_serca; bad |= wr_scalar(f, "use_serca", H5T_STD_U8LE, H5T_NATIVE_UINT8, &u8);
  if (m->use_ryr) bad |= wr_channel(f, "ryr", &m->ryr);
  if (m->use_p_ryr) bad |= wr_channel(f, "p_ryr", &m->p_ryr);
  if (m->use_serca) {
    hid_t g = mkgroup(f, "serca");
    bad |= wr_scalar(g, "v_cyt_to_A_sr", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, &m->serca.v_cyt_to_A_sr);
    bad |= wr_scalar(g, "density", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, &m->serca.density);
    bad |= wr_scalar(g, "scale", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, &m->serca.scale);
    bad |= wr_str(g, "boundary", m->serca.boundary, 0);
    bad |= wr_str(g, "species", m->serca.species, 0);
    H5Gclose(g);
  }
  bad |= wr_array(f, "convolution_constants", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, m->conv, 3);
  H5Fclose(f);
  return bad ? ioerr(err, elen, "failed writing %s", path) : 0;
}

/* ------------------------------------------------------------- output */
struct cs_h5out {
  hid_t file, group;
  char groupname[64];
};

/* write_h5_attr (utils.c:101-138): numeric attributes get a 1-D space of
 * `dims` elements, strings a scalar NULLTERM type of strlen bytes */
static int out_attr(hid_t loc, const char* name, hid_t ftype, hid_t mtype, const void* v, hsize_t n)
{
  return wr_array(loc, name, ftype, mtype, v, n);
}

cs_h5out* cs_out_open(const char* filename, uint32_t output_ind, double t, uint64_t time_ind, double h,
                      const uint64_t N[3], const char* gfile, const char* sfile, char* err, size_t elen)
{
  quiet_hdf5();
  cs_h5out* o = (cs_h5out*)calloc(1, sizeof *o);
  if (output_ind == 0) {
    o->file = H5Fcreate(filename, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (o->file < 0) { ioerr(err, elen, "cannot create %s", filename); free(o); return NULL; }
    int bad = out_attr(o->file, "h", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, &h, 1);
    bad |= out_attr(o->file, "N", H5T_STD_U64LE, H5T_NATIVE_UINT64, N, 3);
    bad |= wr_str(o->file, "geometry_file", gfile, 1);
    bad |= wr_str(o->file, "species_file", sfile, 1);
    if (bad) { ioerr(err, elen, "cannot write root attributes of %s", filename); H5Fclose(o->file); free(o); return NULL; }
  } else {
    o->file = H5Fopen(filename, H5F_ACC_RDWR, H5P_DEFAULT);
    if (o->file < 0) { ioerr(err, elen, "cannot open %s", filename); free(o); return NULL; }
  }
  snprintf(o->groupname, sizeof o->groupname, "/data_%07u", output_ind);
  o->group = mkgroup(o->file, o->groupname);
  if (o->group < 0) { ioerr(err, elen, "cannot create %s", o->groupname); H5Fclose(o->file); free(o); return NULL; }
  int bad = out_attr(o->group, "time", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, &t, 1);
  bad |= out_attr(o->group, "time_index", H5T_STD_U64LE, H5T_NATIVE_UINT64, &time_ind, 1);
  if (bad) { ioerr(err, elen, "cannot write group attributes"); cs_out_close(o); return NULL; }
  return o;
}

int cs_out_scalars(cs_h5out* o, uint32_t nd, uint32_t ns, const char (*dn)[CS_NAME_LEN],
                   const char (*sn)[CS_NAME_LEN], const double* mn, const double* mx, const double* avg)
{
  int bad = 0;
  char path[256];
  for (uint32_t d = 0; d < nd; d++) {
    snprintf(path, sizeof path, "%s/%s", o->groupname, dn[d]);
    hid_t dg = mkgroup(o->file, path);
    if (dg < 0) return -1;
    for (uint32_t s = 0; s < ns; s++) {
      snprintf(path, sizeof path, "%s/%s/%s", o->groupname, dn[d], sn[s]);
      hid_t sg = mkgroup(o->file, path);
      if (sg < 0) { H5Gclose(dg); return -1; }
      bad |= out_attr(sg, "max", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, &mx[d * ns + s], 1);
      bad |= out_attr(sg, "min", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, &mn[d * ns + s], 1);
      bad |= out_attr(sg, "average", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, &avg[d * ns + s], 1);
      H5Gclose(sg);
    }
    H5Gclose(dg);
  }
  return bad ? -1 : 0;
}

int cs_out_discrete(cs_h5out* o, const char* name, const int32_t* states, uint64_t n)
{
  char path[256];
  snprintf(path, sizeof path, "%s/discrete_%s", o->groupname, name);
  hsize_t dims = n;
  hid_t sp = H5Screate_simple(1, &dims, NULL);
  hid_t ds = H5Dcreate2(o->file, path, H5T_IEEE_F64LE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  int bad = ds < 0 || (n && H5Dwrite(ds, H5T_NATIVE_INT32, H5S_ALL, H5S_ALL, H5P_DEFAULT, states) < 0);
  if (ds >= 0) H5Dclose(ds);
  H5Sclose(sp);
  return bad ? -1 : 0;
}

int cs_out_dataset(cs_h5out* o, const char* name, const double* data, int rank, const uint64_t* dims,
                   int chunked)
{
  if (rank < 1 || rank > 3) return -1;
  char path[256];
  snprintf(path, sizeof path, "%s/%s", o->groupname, name);
  hsize_t d[3];
  for (int i = 0; i < rank; i++) d[i] = dims[i];
  hid_t sp = H5Screate_simple(rank, d, NULL);
  hid_t pl = H5Pcreate(H5P_DATASET_CREATE);
  if (chunked) H5Pset_chunk(pl, rank, d);
  hid_t ds = H5Dcreate2(o->file, path, H5T_IEEE_F64LE, sp, H5P_DEFAULT, pl, H5P_DEFAULT);
  int bad = ds < 0 || H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0;
  if (ds >= 0) H5Dclose(ds);
  H5Pclose(pl);
  H5Sclose(sp);
  return bad ? -1 : 0;
}

void cs_out_close(cs_h5out* o)
{
  if (!o) return;
  if (o->group >= 0) H5Gclose(o->group);
  if (o->file >= 0) H5Fclose(o->file);
  free(o);
}

/* ------------------------------------------------------- output readers */
int cs_io_read_dataset(const char* path, const char* dataset, double* out, uint64_t max_elems,
                       uint64_t dims[3], int* rank)
{
  quiet_hdf5();
  hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) return -1;
  int rc = -1;
  hid_t ds = H5Dopen2(f, dataset, H5P_DEFAULT);
  if (ds >= 0) {
    hid_t sp = H5Dget_space(ds);
    int r = H5Sget_simple_extent_ndims(sp);
    hsize_t d[3] = {0, 0, 0};
    if (r >= 1 && r <= 3) {
      H5Sget_simple_extent_dims(sp, d, NULL);
      uint64_t n = 1;
      for (int i = 0; i < r; i++) { n *= d[i]; dims[i] = d[i]; }
      *rank = r;
      if (n <= max_elems && (n == 0 || H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) >= 0)) rc = 0;
    }
    H5Sclose(sp);
    H5Dclose(ds);
  }
  H5Fclose(f);
  return rc;
}

int cs_io_read_attr(const char* path, const char* object, const char* attr, double* out, uint64_t max_elems,
                    uint64_t* count)
{
  quiet_hdf5();
  hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) return -1;
  int rc = -1;
  hid_t a = H5Aopen_by_name(f, object, attr, H5P_DEFAULT, H5P_DEFAULT);
  if (a >= 0) {
    hid_t sp = H5Aget_space(a);
    hssize_t n = H5Sget_simple_extent_npoints(sp);
    if (n >= 0 && (uint64_t)n <= max_elems && H5Aread(a, H5T_NATIVE_DOUBLE, out) >= 0) {
      *count = (uint64_t)n;
      rc = 0;
    }
    H5Sclose(sp);
    H5Aclose(a);
  }
  H5Fclose(f);
  return rc;
}

int cs_io_read_attr_str(const char* path, const char* object, const char* attr, char* out, size_t out_len)
{
  quiet_hdf5();
  hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) return -1;
  int rc = rd_str(f, object, attr, out, out_len, NULL, 0);
  H5Fclose(f);
  return rc;
}
