/*
 * libdim_gpu: MI355X (gfx950) CSR5 sparse matrix-vector multiply and
 * conjugate-gradient solver, C ABI.
 *
 * Drop-in boundary for the CSR5 path of dim (rurabori/diploma_research).
 * Every entry point names the reference interface it replaces
 * (file:line in the reference tree). Conventions:
 *   - plain pointers and sizes; no exceptions cross this boundary: every call
 *     returns a dim_status and dim_last_error() describes the last failure on
 *     the calling thread;
 *   - matrix construction takes HOST arrays (the reference's csr is a host
 *     container); products and solves take DEVICE pointers and a hipStream_t
 *     passed as void* (NULL = the default stream) and are stream-ordered;
 *   - a dim_csr5 handle owns its device arrays and a small scratch area (the
 *     counterpart of the reference's calibrator, csr5/calibrator.h:14-31), so
 *     calls on ONE handle must not overlap (same rule as the reference's
 *     shared calibrator); different handles are independent.
 *   - y = A*x OVERWRITES y, including 0 in empty rows (the reference adds into
 *     a zero-filled y and skips empty rows, csr5.h:526-538, 750-754).
 */
#ifndef DIM_GPU_H
#define DIM_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DIM_GPU_VERSION 1

typedef enum dim_status {
    DIM_SUCCESS = 0,
    DIM_INVALID_ARGUMENT = 1, /* reference: std::invalid_argument (csr5.h:820-821) */
    DIM_HIP_ERROR = 2,        /* a HIP runtime call failed */
    DIM_OUT_OF_MEMORY = 3,
    DIM_NOT_CONVERGED = 4,    /* CG reached max_iters above tolerance */
    DIM_BREAKDOWN = 5,        /* CG: p'Ap <= 0 or not finite (matrix not SPD) */
    DIM_IO_ERROR = 6,         /* reference: runtime_error("HDF5 failed.") (include/dim/io/h5/err.h:8-16) */
    DIM_UNSUPPORTED = 7
} dim_status;

typedef enum dim_dtype { DIM_F64 = 0, DIM_F32 = 1 } dim_dtype;

typedef struct dim_csr5 dim_csr5;

/* Construction options. */
typedef struct dim_csr5_opts {
    uint32_t sigma; /* tile height: 4, 5, 6, 7, 8, 10, 12, 14, 16, 20, 24, 28 or 32; 0 = pick by the
                       built-in autotuner (times the candidates near the mean row length) */
    uint32_t flags; /* DIM_CSR5_* below */
} dim_csr5_opts;
#define DIM_CSR5_KEEP_F32 1u /* also keep FP32 values (for DIM_F32 products and mixed-precision CG) */
#define DIM_CSR5_XCD_SLICES 2u /* also keep the column-sliced form for dim_csr5_spmv (scattered columns,
                                  large x: each XCD gathers from one eighth of x). With sigma = 0 the
                                  autotuner adds it on its own when it is faster, and the flag forces
                                  it. No reference counterpart. */
#define DIM_CSR5_PLAIN_COLUMNS 4u /* do not keep column codes: products always stream the 32-bit column
                                     indices. By default a matrix with no empty rows whose elements
                                     lie on few diagonals (at most 16 for sigma <= 6, 8 for sigma 7
                                     and 8) also keeps a 3- or 4-bit code per element, packed with
                                     the row-start flags into one word per lane and tile, and the
                                     products read those. No reference counterpart. */

typedef struct dim_csr5_info {
    uint32_t rows, cols;
    uint64_t nnz;
    uint32_t sigma, omega;
    uint32_t tile_count;           /* ceil(nnz / (omega*sigma)); the last tile is padded */
    uint32_t tail_partition_start; /* always = rows: no separate tail (reference: csr5.h:806) */
    uint32_t n_empty_rows;
    uint32_t n_dirty_tiles;
    uint32_t has_f32;
    uint64_t device_bytes;
    uint32_t xcd_slices;   /* column slices used by dim_csr5_spmv (0: none) */
    uint32_t xcd_sigma;    /* tile height of the sliced form */
    uint32_t column_codes; /* 1: products read column codes instead of 32-bit indices */
    uint32_t reserved;
} dim_csr5_info;

const char* dim_last_error(void);
const char* dim_status_string(dim_status s);
int dim_version(void);

/* csr5::from_csr (csr5.h:791-803): builds the CSR5 matrix from host CSR
 * arrays (row_ptr: rows+1, col_idx/vals: nnz; uint32 indices as in csr.h:19-78)
 * and uploads it to the current device. */
dim_status dim_csr5_from_csr(uint32_t rows, uint32_t cols, uint64_t nnz, const uint32_t* row_ptr,
                             const uint32_t* col_idx, const double* vals, const dim_csr5_opts* opts,
                             dim_csr5** out);

/* Same, from DEVICE CSR arrays: the conversion runs on the GPU. */
dim_status dim_csr5_from_csr_device(uint32_t rows, uint32_t cols, uint64_t nnz, const uint32_t* row_ptr,
                                    const uint32_t* col_idx, const double* vals, const dim_csr5_opts* opts,
                                    void* stream, dim_csr5** out);

dim_status dim_csr5_destroy(dim_csr5* A);
dim_status dim_csr5_get_info(const dim_csr5* A, dim_csr5_info* info);

/* Copies the format arrays to host buffers (any may be NULL): tile_ptr
 * (tile_count+1), bit_flag (tile_count*64, widened to u32), tile_flags
 * (tile_count+1), tile_desc_offset_ptr (tile_count+1), tile_desc_offset
 * (*n_tdo entries; pass n_tdo to receive the count), empty_rows. */
dim_status dim_csr5_copy_format(const dim_csr5* A, uint32_t* tile_ptr, uint32_t* bit_flag, uint8_t* tile_flags,
                                uint32_t* tile_desc_offset_ptr, uint32_t* tile_desc_offset, uint64_t* n_tdo,
                                uint32_t* empty_rows);

/* csr5::spmv (csr5.h:818-829): y = A*x. x: cols entries, y: rows entries, of
 * `dtype` (DIM_F32 needs DIM_CSR5_KEEP_F32). x_len != cols is
 * DIM_INVALID_ARGUMENT, as the reference throws. */
dim_status dim_csr5_spmv(const dim_csr5* A, dim_dtype dtype, const void* x, uint64_t x_len, void* y,
                         uint64_t y_len, void* stream);

/* csr5::spmv (csr5.h:818-829) for callers that hold x and y in HOST memory,
 * as the reference's callers do (benchmarks/csr5/dim/main.cpp:58-68): copies
 * x to the device, runs dim_csr5_spmv's kernels and copies y back; returns
 * when y is written. The device copies live in the handle (cols + rows
 * entries of 8 bytes, allocated on the first call); the same argument rules
 * as dim_csr5_spmv. */
dim_status dim_csr5_spmv_host(const dim_csr5* A, dim_dtype dtype, const void* x, uint64_t x_len, void* y,
                              uint64_t y_len, void* stream);

/* Measurement variant of dim_csr5_spmv: records HIP events around the tile
 * kernel and the carry/fix-up kernel on `stream`, waits for them and returns
 * both durations in milliseconds (ms[0] tiles, ms[1] fix-up; ms[1] is 0 when
 * the product needs no fix-up kernel and none was launched). With column
 * slices, ms[0] is the sliced tile kernel and ms[1] its fix-up plus the slice
 * sum. No reference
 * counterpart (the reference times whole calls with a stopwatch,
 * include/dim/bench/stopwatch.h:11-22). */
dim_status dim_csr5_spmv_timed(const dim_csr5* A, dim_dtype dtype, const void* x, uint64_t x_len, void* y,
                               uint64_t y_len, void* stream, float* ms);

/* csr5::spmv<spmv_strategy::partial> (csr5.h:585-594, 756-782) on the tile
 * range [tile_begin, tile_end) of the whole matrix: y_block[i] receives the
 * contribution of those tiles to row first_row + i, where first_row is the row
 * of tile_begin's first element (reference: first_row_idx(), csr5.h:303) and
 * the block spans last_row_idx() - first_row_idx() + 1 rows (csr5.h:304-320).
 * Rows shared with the neighbouring blocks hold partial sums, as in the
 * reference, whose edge sync adds them up (libs/dim/mpi/csr5/edge_sync.cpp). */
dim_status dim_csr5_block_rows(const dim_csr5* A, uint32_t tile_begin, uint32_t tile_end, uint32_t* first_row,
                               uint32_t* n_rows);
dim_status dim_csr5_spmv_block(const dim_csr5* A, dim_dtype dtype, uint32_t tile_begin, uint32_t tile_end,
                               const void* x, uint64_t x_len, void* y_block, uint64_t y_len, void* stream);

/* Multi-RHS product Y = A*X with k in 1..16 right-hand sides decoded from
 * one pass over the tiles. X is cols x k and Y rows x k, row-major (entry
 * (i, j) at i*k + j). For k >= 6 the blocks are staged column-major in a
 * scratch buffer that the matrix keeps (allocated on first use, grown when a
 * larger block comes). No reference counterpart (the reference has only
 * single-vector spmv). */
dim_status dim_csr5_spmm(const dim_csr5* A, dim_dtype dtype, uint32_t k, const void* X, void* Y, void* stream);

/* The same product on column-major blocks: entry (i, j) of X at X[i + j*ldx]
 * (ldx >= cols) and of Y at Y[i + j*ldy] (ldy >= rows). The faster layout:
 * each right-hand side is gathered from its own contiguous column. No
 * reference counterpart. */
dim_status dim_csr5_spmm_colmajor(const dim_csr5* A, dim_dtype dtype, uint32_t k, const void* X, uint64_t ldx,
                                  void* Y, uint64_t ldy, void* stream);

/* y = A^T x (x: rows entries, y: cols entries). No reference counterpart;
 * used by the adjoint. */
dim_status dim_csr5_spmv_transpose(const dim_csr5* A, dim_dtype dtype, const void* x, void* y, void* stream);

/* out[e] = scale * u[row(e)] * v[col(e)] for every stored element, in the
 * CSR order of the arrays the matrix was built from (FP64). With scale = -1,
 * u = lambda and v = x this is dL/dA on the sparsity pattern. */
dim_status dim_csr5_sampled_outer(const dim_csr5* A, double scale, const double* u, const double* v, double* out,
                                  void* stream);

/* ------------------------------------------------------------------ CG ---- */

typedef enum dim_cg_precision {
    DIM_CG_F64 = 0,       /* plain FP64 CG */
    DIM_CG_F32_REFINED = 1 /* FP32 iteration, FP64 residual/solution/norm (needs DIM_CSR5_KEEP_F32) */
} dim_cg_precision;

typedef struct dim_cg_opts {
    double rtol;           /* stop when ||r|| / ||b|| <= rtol */
    uint32_t max_iters;    /* iteration cap (for F32_REFINED: inner iterations in total) */
    uint32_t precision;    /* dim_cg_precision */
    uint32_t check_every;  /* iterations between host convergence checks (0 = 16) */
    double refine_delta;   /* F32_REFINED: FP64 residual replacement when ||r|| drops by this factor (0 = 0.1) */
    double* history;       /* optional HOST array of max_iters+1: ||r_i||/||b|| (recursive residual) */
    uint32_t flags;        /* DIM_CG_* below; batched only */
} dim_cg_opts;
/* iterate with A^T (the adjoint solve A^T l = g) */
#define DIM_CG_TRANSPOSE 1u
/* run the whole iteration loop in one persistent cooperative kernel with
 * grid barriers instead of replaying a captured launch sequence (not with
 * DIM_CG_TRANSPOSE; DIM_UNSUPPORTED when the kernel cannot be resident) */
#define DIM_CG_PERSISTENT 2u

typedef struct dim_cg_result {
    uint32_t iterations;
    int32_t converged;
    double relres;         /* the stopping quantity at exit */
    double relres_true;    /* ||b - A x||_2 / ||b||_2 recomputed in FP64 at exit */
    uint32_t refinements;  /* F32_REFINED: FP64 residual replacements */
    double seconds;        /* wall time of the solve (device work, stream-synchronised) */
} dim_cg_result;

/* conjugate_gradient (thesis.tex:2133-2160; step order of
 * benchmarks/conjugate_gradient/dim_sp/main.cpp:102-127): solves A x = b for
 * SPD A. b and x are DEVICE FP64 arrays of rows entries; x holds x0 on
 * entry. Synchronises `stream` before returning. The handle keeps the work
 * buffers (when they take at most 256 MiB) and the captured iterations of its
 * last dim_cg / dim_cg_multi / dim_cg_host solve for the next solve of the
 * same shape, and frees them with the handle. */
dim_status dim_cg(const dim_csr5* A, const double* b, double* x, const dim_cg_opts* opts, dim_cg_result* res,
                  void* stream);

/* dim_cg with b and x in HOST memory (benchmarks/conjugate_gradient/dim_sp/
 * main.cpp:84-127 keeps its vectors on the host): x holds x0 on entry and
 * the solution on return (also when the solve stops unconverged). Staged
 * through the handle's device copies, as dim_csr5_spmv_host. */
dim_status dim_cg_host(const dim_csr5* A, const double* b, double* x, const dim_cg_opts* opts, dim_cg_result* res,
                       void* stream);

/* The solve of dim_cg (one right-hand side, DIM_CG_F64 or DIM_CG_F32_REFINED)
 * for callers that cannot allocate device memory: loops that re-solve many
 * times, harnesses that forbid hipMalloc. No counterpart in the reference.
 * It calls no hipMalloc, hipFree or other allocating runtime routine (no
 * pinned memory, graphs or events): its state, reduction partials and work
 * vectors live in one static device array of DIM_CG_NOALLOC_BYTES (48.1 MiB)
 * that every GPU the library runs on holds until the process ends, and the
 * product's own carries are the handle's (allocated with the matrix).
 * Limits: rows <= DIM_CG_NOALLOC_MAX_ROWS (else DIM_INVALID_ARGUMENT),
 * opts->history NULL, the matrix on the current device. Calls that overlap
 * on one GPU (other threads or streams) race on the array: serialise them.
 * res->seconds is the host wall time of the call. Results equal dim_cg's
 * bit for bit (same kernels in the same order). */
#define DIM_CG_NOALLOC_MAX_ROWS 2097152u
#define DIM_CG_NOALLOC_BYTES (1024u + 4u * 256u * 16u * 8u + 24u * DIM_CG_NOALLOC_MAX_ROWS + 8u * 256u)
dim_status dim_cg_noalloc(const dim_csr5* A, const double* b, double* x, const dim_cg_opts* opts,
                          dim_cg_result* res, void* stream);

/* k independent right-hand sides (B, X: rows x k row-major, DEVICE, FP64),
 * iterated together with one SpMM per iteration; res has k entries. */
dim_status dim_cg_multi(const dim_csr5* A, uint32_t k, const double* B, double* X, const dim_cg_opts* opts,
                        dim_cg_result* res, void* stream);

/* Many independent SPD systems in one launch sequence: `batch` is a
 * block-diagonal matrix built with dim_csr5_from_csr_batched (host CSR arrays
 * per system, each square; rows[i] rows and nnz[i] entries), B/X hold the
 * systems' vectors back to back (DEVICE, FP64). Every system keeps its own
 * step lengths and stop test; a batch of iterations is captured once in a
 * hipGraph and replayed until every system has stopped. res has n_systems
 * entries. With opts->flags & DIM_CG_TRANSPOSE the products use A^T (the
 * adjoint solve). No reference counterpart (the reference solves one system
 * per process). Padding elements of each system's last tile are stored as
 * 0 * x[first row of the system], so x must be finite. The batch handle keeps
 * its block tables, vectors (when they take at most 256 MiB) and captured
 * iterations from its first dim_cg_batched solve, and frees them with the
 * handle. */
dim_status dim_csr5_from_csr_batched(uint32_t n_systems, const uint32_t* rows, const uint64_t* nnz,
                                     const uint32_t* const* row_ptr, const uint32_t* const* col_idx,
                                     const double* const* vals, const dim_csr5_opts* opts, dim_csr5** out);
dim_status dim_cg_batched(const dim_csr5* batch, const double* B, double* X, const dim_cg_opts* opts,
                          dim_cg_result* res, void* stream);
uint32_t dim_csr5_n_systems(const dim_csr5* A);

/* Replaces the stored values (same sparsity pattern) from a DEVICE FP64 array
 * in the CSR order the matrix was built from (for batched matrices: the
 * systems' arrays back to back). Keeps the FP32 copy in step when present and
 * drops the column-sliced form (DIM_CSR5_XCD_SLICES). No reference
 * counterpart; used when differentiating with respect to A. */
dim_status dim_csr5_set_values(dim_csr5* A, const double* vals, void* stream);

/* Measurement helper with no reference counterpart: reads `bytes` (a multiple
 * of 16, 16-byte aligned) of DEVICE memory at `data` once with non-temporal
 * 16-byte loads and writes their sum to the DEVICE double *sum, on `stream`.
 * Timing it gives the HBM read rate the GPU reaches on a plain stream, the
 * measured ceiling reported beside the 8 TB/s peak. */
dim_status dim_stream_read(const void* data, uint64_t bytes, double* sum, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* DIM_GPU_H */
