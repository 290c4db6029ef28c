/*
 * cs_io.h -- the reference's HDF5 file formats (libcsio.so, C + HDF5).
 *
 *  geometry file  /global_size, /h, /domains/{num, name_i, indices},
 *                 /domains/voxels (uint8 [N0][N1][N2]),
 *                 /domains/domain_connections (uint8 [nd][nd]),
 *                 /boundaries/{num, name_i, type_i} + one uint16 [n][6]
 *                 dataset per boundary            (geometry.c:92-141,
 *                                                  :487-574, :576-620)
 *  species file   attributes listed in species.c:103-258, :346-362, :780
 *                 and boundaryfluxes.c:376-540
 *  output file    <casename>.h5: root attributes h, N, geometry_file,
 *                 species_file; per output a group /data_%07d with
 *                 time/time_index, /<domain>/<species> groups carrying
 *                 max/min/average, discrete_<boundary> datasets, sheets,
 *                 <species>_all_data and linescans   (species.c:2253-2551,
 *                                                     :2785-3092, :2040-2102)
 */
#ifndef CS_IO_H
#define CS_IO_H

#include <stddef.h>
#include <stdint.h>
#include "cs_model.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Read a geometry file.  Allocates g->voxels and g->boundaries[i]; free
 * them with cs_io_free_geometry. */
int  cs_io_read_geometry(const char* path, cs_geometry* g, char* err, size_t err_len);
void cs_io_free_geometry(cs_geometry* g);

/* Read a species file (the SERCA species comes from /ryr/species, as the
 * reference reads it, boundaryfluxes.c:518). */
int  cs_io_read_species(const char* path, cs_species_model* m, char* err, size_t err_len);

/* Write input files in the reference formats (used to generate dyads). */
int  cs_io_write_geometry(const char* path, const cs_geometry* g, char* err, size_t err_len);
int  cs_io_write_species(const char* path, const cs_species_model* m, char* err, size_t err_len);

/* Output file.  cs_out_open creates the file at output_ind 0 (with the
 * root attributes) and reopens it later, then creates /data_%07d. */
typedef struct cs_h5out cs_h5out;

cs_h5out* cs_out_open(const char* filename, uint32_t output_ind, double t, uint64_t time_ind,
                      double h, const uint64_t N[3], const char* geometry_file,
                      const char* species_file, char* err, size_t err_len);
/* /data_n/<domain>/<species> groups with max/min/average, arrays [nd][ns] */
int  cs_out_scalars(cs_h5out* o, uint32_t nd, uint32_t ns, const char (*domain_names)[CS_NAME_LEN],
                    const char (*species_names)[CS_NAME_LEN], const double* mn, const double* mx,
                    const double* avg);
/* /data_n/discrete_<name>: float64 file type written from int32 states */
int  cs_out_discrete(cs_h5out* o, const char* name, const int32_t* states, uint64_t n);
/* 1-, 2- or 3-D float64 dataset under the group; chunked=1 uses one chunk
 * of the whole extent (the reference's all_data layout on one rank) */
int  cs_out_dataset(cs_h5out* o, const char* name, const double* data, int rank,
                    const uint64_t* dims, int chunked);
void cs_out_close(cs_h5out* o);

/* Readers for output files: a float64 dataset (any rank <= 3; dims
 * receives its extent, at most max_elems values are read) and a float64
 * or integer attribute converted to double.  Return 0 or -1. */
int  cs_io_read_dataset(const char* path, const char* dataset, double* out, uint64_t max_elems,
                        uint64_t dims[3], int* rank);
int  cs_io_read_attr(const char* path, const char* object, const char* attr, double* out,
                     uint64_t max_elems, uint64_t* count);
int  cs_io_read_attr_str(const char* path, const char* object, const char* attr, char* out,
                         size_t out_len);

#ifdef __cplusplus
}
#endif
#endif
