/* dim_h5.h - HDF5 matrix and vector files of the dim tool chain (C ABI).
 *
 * Host-side companion of dim_gpu.h, in its own library (libdim_h5.so, linked
 * against the HDF5 C library) so that the GPU library carries no HDF5
 * dependency. It reads and writes the two matrix group layouts that the
 * reference writes, byte-compatible with it:
 *
 *   CSR5 group (dim::io::h5::store / load_csr5, libs/dim/io/h5.cpp:45-74):
 *     attribute column_count (u32); datasets vals (f64), col_idx (u32), both
 *     in the reference's tile-transposed order (omega = 4, sigma = 16,
 *     include/dim/mat/storage_formats/csr5.h:463-493), row_ptr (u32, CSR),
 *     tile_ptr (u32, bit 31 marks a tile holding an empty row), tile_desc
 *     (H5T array of 4 u32 per tile: y_offset bits 0-6, scansum_offset bits
 *     7-9, bit_flag bits 10-25, csr5.h:39-63), tile_desc_offset_ptr and
 *     tile_desc_offset (u32).
 *   MATLAB CSR group (write_matlab_compatible / read_matlab_compatible,
 *     include/dim/io/h5.h:68-97): data (f64), ir (u32 column indices), jc
 *     (u32 row offsets), attribute MATLAB_sparse (u32 column count).
 *
 * Errors are returned as dim_status codes with a message in
 * dim_h5_last_error() (thread-local); the reference throws instead
 * (include/dim/io/h5/err.h:8-16). The two dim_csr5_load_h5* calls also build
 * the GPU matrix, so this library links libdim_gpu.so. */
#ifndef DIM_H5_H
#define DIM_H5_H

#include "dim_gpu.h"

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* A CSR matrix in host memory. Arrays returned by the readers are owned by
 * the library and released with dim_h5_free_csr. */
typedef struct dim_h5_csr {
    uint32_t rows, cols;
    uint64_t nnz;
    uint32_t* row_ptr; /* rows + 1 */
    uint32_t* col_idx; /* nnz, sorted within each row */
    double* vals;      /* nnz */
} dim_h5_csr;

/* Storage properties of one dataset (dim::io::h5::dataset_props_t,
 * include/dim/io/h5.h:30-36): chunk = 0 writes a contiguous dataset (chunking
 * also applies only when chunk < dataset size, h5.cpp:34-43); deflate < 0
 * writes uncompressed, 0..9 is the deflate level. */
typedef struct dim_h5_dataset_props {
    uint64_t chunk;
    int32_t deflate;
} dim_h5_dataset_props;

/* matrix_storage_props_t (MATLAB CSR group) */
typedef struct dim_h5_csr_props {
    dim_h5_dataset_props values, col_idx, row_idx;
} dim_h5_csr_props;

/* csr5_storage_props_t (CSR5 group) */
typedef struct dim_h5_csr5_props {
    dim_h5_dataset_props vals, col_idx, row_ptr, tile_ptr, tile_desc, tile_desc_offset, tile_desc_offset_ptr;
} dim_h5_csr5_props;

typedef enum dim_h5_group_kind {
    DIM_H5_UNKNOWN = 0,
    DIM_H5_CSR5 = 1,       /* reference CSR5 group */
    DIM_H5_MATLAB_CSR = 2  /* MATLAB-compatible CSR group */
} dim_h5_group_kind;

const char* dim_h5_last_error(void);

/* dim::io::h5::is_hdf5 (h5.cpp:76-86): 1 if the file starts with the HDF5
 * signature, 0 if not, -1 if it cannot be read. */
int dim_h5_is_hdf5(const char* path);

/* The layout of group `group` in `path` (by its datasets). */
dim_status dim_h5_group_kind_of(const char* path, const char* group, dim_h5_group_kind* kind);

/* Reads a matrix group of either layout into CSR (load_csr5 followed by the
 * inverse of the tile transpose, or read_matlab_compatible). */
dim_status dim_h5_read_csr(const char* path, const char* group, dim_h5_csr* out);
void dim_h5_free_csr(dim_h5_csr* m);

/* One rank's share of a CSR5 group: dim::io::h5::load_csr5_partial
 * (libs/dim/io/h5.cpp:137-173) of part part_idx of part_count, read by
 * hyperslabs. The tiles are split as the reference splits them
 * (ceil(T / part_count) per part, h5.cpp:95-103); the part covers the rows
 * first_row_idx() .. last_row_idx() (csr5.h:303-320, 850-855) and the stored
 * elements [val_offset, val_offset + block.nnz) (the last part also takes the
 * tail elements after the last tile; skip_tail marks the others, csr5.h:457-458,
 * 756-782). `block` holds those elements as CSR with block.rows =
 * last_row - first_row + 1 rows of the global column count, row i being the
 * part's elements of row first_row + i in CSR order: its product
 * (dim_csr5_from_csr, dim_csr5_spmv) is the part's share of y, the first and
 * last rows being partial sums when the part starts or ends inside them, as
 * with spmv<partial>. Parts that would start past the last tile (the
 * reference wraps around there) are DIM_INVALID_ARGUMENT. Release with
 * dim_h5_free_csr5_part. */
typedef struct dim_h5_csr5_part {
    uint32_t rows, cols;   /* the whole matrix */
    uint32_t first_row;    /* first_row_idx() */
    uint32_t last_row;     /* last_row_idx() */
    uint64_t tile_begin;   /* the part's first reference tile (omega = 4, sigma = 16) */
    uint64_t tile_count;   /* its full tiles */
    uint64_t val_offset;   /* index of its first element in the group's vals */
    int32_t skip_tail;     /* 1 unless this is the last part */
    dim_h5_csr block;
    int32_t first_tile_uncapped; /* first_tile_uncapped() (csr5.h:845-847): the part's first element does
                                    not start a row (the row began in an earlier part) */
} dim_h5_csr5_part;
dim_status dim_h5_load_csr5_partial(const char* path, const char* group, uint32_t part_idx, uint32_t part_count,
                                    dim_h5_csr5_part* out);
void dim_h5_free_csr5_part(dim_h5_csr5_part* p);

/* h5::load_csr5 (libs/dim/io/h5.cpp:57-74, 174-177) straight to a GPU matrix:
 * dim_h5_read_csr of either group layout, then dim_csr5_from_csr with `opts`
 * (may be NULL: sigma 16), on the current device. The single call that the
 * reference's benchmarks make (benchmarks/csr5/dim/main.cpp:51). */
dim_status dim_csr5_load_h5(const char* path, const char* group, const dim_csr5_opts* opts, dim_csr5** out);

/* load_csr5_partial straight to a GPU matrix holding one rank's share: the
 * part's block (see dim_h5_load_csr5_partial) built with dim_csr5_from_csr.
 * Its products (dim_csr5_spmv with y of *n_rows entries) give y rows
 * *first_row .. *first_row + *n_rows - 1, as spmv<partial> of the loaded part
 * does (include/dim/mpi/csr5.h:51-54). first_row / n_rows may be NULL. */
dim_status dim_csr5_load_h5_partial(const char* path, const char* group, uint32_t part_idx, uint32_t part_count,
                                    const dim_csr5_opts* opts, dim_csr5** out, uint32_t* first_row,
                                    uint32_t* n_rows);

/* Writes `m` as a MATLAB CSR group (write_matlab_compatible) or as a
 * reference CSR5 group (csr5::from_csr + h5::store). The file is created (or
 * truncated) unless `append` is set, in which case an existing file is opened
 * and the group added (store_matrix.cpp:96-100). props may be null (no
 * chunking, no compression). */
dim_status dim_h5_write_matlab_csr(const char* path, const char* group, int append, const dim_h5_csr* m,
                                   const dim_h5_csr_props* props);
dim_status dim_h5_write_csr5(const char* path, const char* group, int append, const dim_h5_csr* m,
                             const dim_h5_csr5_props* props);

/* FP64 vectors as 1-D datasets (the SpMV benchmark's result output,
 * benchmarks/csr5/dim/main.cpp:34-42, and compare_results'
 * inputs, apps/dim_cli/subcommands/compare_results.cpp:12-19). `name` may
 * hold a group path ("/Y", "grp/Y"). write: the file is created if missing;
 * an existing dataset of that name is replaced only when `overwrite` is set
 * (otherwise DIM_INVALID_ARGUMENT, as the benchmark refuses). read: *data is
 * released with dim_h5_free. */
dim_status dim_h5_write_vector(const char* path, const char* name, const double* v, uint64_t n, int overwrite);
dim_status dim_h5_read_vector(const char* path, const char* name, double** data, uint64_t* n);
void dim_h5_free(void* p);

#ifdef __cplusplus
}
#endif

#endif /* DIM_H5_H */
