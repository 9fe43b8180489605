/*
 * goosefem_amd_h5.h -- C ABI of libgoosefem_amd_h5.so: HDF5 field output of the examples, with the
 * semantics of the H5Easy calls the reference's examples make (HighFive H5Easy::File / dump / load /
 * exist / getShape; docs/examples/dynamics/Elastic-VelocityVerlet/main.cpp:161-176 writes
 * "/global/{t,Ekin,Epot}" and "/mesh/{conn,coor,disp}").  Host arrays, row-major.  Built over the
 * HDF5 C library (1.10) as its own library, so that libgoosefem_amd.so needs no HDF5.
 * Every function returns GF_OK or GF_EINVAL; gf_h5_last_error() gives the message.
 */
#ifndef GOOSEFEM_AMD_H5_H
#define GOOSEFEM_AMD_H5_H

#include <stdint.h>

#include "goosefem_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gf_h5_s* gf_h5;

/* file modes: H5Easy::File::Overwrite, ::ReadOnly, ::ReadWrite (created if absent) */
#define GF_H5_OVERWRITE 0
#define GF_H5_READONLY 1
#define GF_H5_READWRITE 2

/* element types of dumped / loaded arrays: double, int64, uint64 (size_t arrays such as conn) */
#define GF_H5_F64 0
#define GF_H5_I64 1
#define GF_H5_U64 2

const char* gf_h5_last_error(void);
/* H5Easy::File(path, mode) */
int gf_h5_open(gf_h5* out, const char* path, int mode);
int gf_h5_close(gf_h5 f);
/* H5Easy::exist(file, path) */
int gf_h5_exist(gf_h5 f, const char* path, int* exist);
/* H5Easy::dump(file, path, data[, H5Easy::DumpMode::Overwrite]): a dataset of rank (0..8) and shape,
 * intermediate groups created; an existing dataset is an error unless overwrite != 0, and then must
 * have the same shape */
int gf_h5_dump(gf_h5 f, const char* path, const void* data_h, int dtype, int rank, const int64_t* shape,
               int overwrite);
/* H5Easy::dump(file, path, value, {index}): entry index of an extendible 1-D dataset (created chunked
 * on first use, grown to index + 1 as needed) */
int gf_h5_dump_at(gf_h5 f, const char* path, const void* value_h, int dtype, int64_t index);
/* H5Easy::getShape(file, path): rank and shape[rank] (rank <= 8) */
int gf_h5_shape(gf_h5 f, const char* path, int* rank, int64_t* shape);
/* H5Easy::load<T>(file, path): the whole dataset into data_h (converted to dtype) */
int gf_h5_load(gf_h5 f, const char* path, void* data_h, int dtype);

#ifdef __cplusplus
}
#endif

#endif
