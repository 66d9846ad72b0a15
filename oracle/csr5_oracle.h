/*
 * The oracle: a CPU restatement of dim's CSR5 conversion, CSR5 SpMV, dense
 * vector kernels and the conjugate-gradient loop.
 *
 * TEST INFRASTRUCTURE ONLY. It is the checker that the HIP path is compared
 * against (tests/, __graft_entry__.smoke()) and the CPU leg of bench.py. The
 * product library never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates.
 */
#ifndef DIM_ORACLE_CSR5_ORACLE_H
#define DIM_ORACLE_CSR5_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_OK 0
#define ORC_INVALID_ARGUMENT 1 /* reference: std::invalid_argument (csr5.h:820-821) */
#define ORC_NO_MEMORY 2
#define ORC_BREAKDOWN 3

#define ORC_DIRTY_BIT 0x80000000u /* tile_ptr MSB marks a tile with empty rows (csr5.h:156-169) */

/*
 * CSR5 matrix in the reference layout, generalised over sigma (<= 64) and
 * omega (<= 64). The tile descriptor is kept unpacked, one entry per tile
 * column; orc_csr5_pack_desc() produces the reference's packed bit-field words.
 */
typedef struct orc_csr5 {
    uint32_t sigma, omega;
    uint32_t rows, cols;
    uint64_t nnz;
    uint64_t tile_count;            /* floor(nnz / (sigma*omega)), csr5.h:438-439 */
    uint32_t* tile_ptr;             /* tile_count + 1 */
    uint32_t* y_offset;             /* tile_count * omega */
    uint32_t* scansum_offset;       /* tile_count * omega */
    uint64_t* bit_flag;             /* tile_count * omega */
    uint32_t* tile_desc_offset_ptr; /* tile_count + 1 */
    uint32_t* tile_desc_offset;     /* n_tile_desc_offset */
    uint64_t n_tile_desc_offset;
    double* vals;      /* nnz, full tiles transposed (csr5.h:463-493) */
    uint32_t* col_idx; /* nnz, same permutation */
    uint32_t* row_ptr; /* rows + 1, plain CSR */
} orc_csr5;

/* csr5::from_csr (csr5.h:791-803) -> csr5_info_t::from_csr (csr5.h:434-446). */
int orc_csr5_from_csr(uint32_t sigma, uint32_t omega, uint32_t rows, uint32_t cols, uint64_t nnz,
                      const uint32_t* row_ptr, const uint32_t* col_idx, const double* vals, orc_csr5** out);
void orc_csr5_free(orc_csr5* m);

/* Packs tile t's descriptor into the reference's per-column words (sigma=16,
 * omega=4: y_offset bits 0-6, scansum_offset bits 7-9, bit_flag bits 10-25;
 * csr5.h:33-63). Writes omega words. */
int orc_csr5_pack_desc(const orc_csr5* m, uint64_t tile, uint32_t* words);

/* y = A*x with the reference's tile loop, calibrator and tail
 * (csr5.h:608-829), emulating `nthreads` OpenMP threads. y is zeroed and the
 * calibrator is fresh on every call (the reference requires both). */
int orc_csr5_spmv(const orc_csr5* m, uint32_t nthreads, const double* x, uint64_t x_len, double* y, uint64_t y_len);

/* Row-block share of a CSR5 SpMV: the contributions of the non-zeros
 * [nz_begin, nz_end) (in CSR order) to rows first_row .. first_row+n_out-1,
 * written to y_block[row - first_row]. A rank's tile range [t0, t1) covers the
 * non-zeros [t0*sigma*omega, t1*sigma*omega) (nnz for the last rank), which is
 * the partial strategy (csr5.h:585-594, 756-782; h5.cpp:268-304) restated by
 * brute force over the un-transposed CSR arrays. */
int orc_csr_block_spmv(uint32_t rows, const uint32_t* row_ptr, const uint32_t* col_idx, const double* vals,
                       uint64_t nz_begin, uint64_t nz_end, const double* x, uint32_t first_row,
                       uint32_t n_out, double* y_block);

/* Brute-force CSR SpMV: y[i] = sum_j a_ij x_j in row order. */
void orc_csr_spmv(uint32_t rows, const uint32_t* row_ptr, const uint32_t* col_idx, const double* vals,
                  const double* x, double* y);

/* dim::vec kernels (vec.h:29-84) with a fixed, thread-count independent
 * reduction order (blocks of 4096, blocks summed in order). */
double orc_dot(uint64_t n, const double* a, const double* b);
void orc_axpy(uint64_t n, double* y, const double* x, double alpha); /* y += alpha*x (vec.h:49-52) */
void orc_aypx(uint64_t n, double* y, const double* x, double alpha); /* y = x + alpha*y (vec.h:57-60) */

typedef struct orc_cg_result {
    uint32_t iterations;
    int32_t converged;
    double relres;      /* recursive residual sqrt(r.r)/||b|| at exit */
    double relres_true; /* ||b - A x|| / ||b|| recomputed at exit */
} orc_cg_result;

/* Textbook CG (thesis.tex:2133-2160) in the step order of
 * benchmarks/conjugate_gradient/dim_sp/main.cpp:102-127, with the stop rule
 * sqrt(r.r)/||b|| <= rtol. x holds x0 on entry. `history` (maxit+1 entries,
 * may be NULL) receives sqrt(r.r)/||b|| after every iteration (index 0: start). */
int orc_cg(const orc_csr5* m, uint32_t nthreads, const double* b, double* x, double rtol, uint32_t maxit,
           orc_cg_result* res, double* history);

/* Test hooks for the 4-lane helpers restated above, checked against the
 * reference's own known answers (tests/test_simd.cpp:17-30). */
void orc_lane_inclusive_scan(uint32_t omega, const double* a, double* s);       /* simd.h:34-45 */
void orc_fast_segmented_sum(uint32_t omega, const double* vec, const uint32_t* off, double* out); /* csr5.h:200-215 */
void orc_shuffle_relative(uint32_t omega, const double* vec, const uint32_t* off, double* out); /* simd.h:70-81 */
double orc_lane_sum(uint32_t omega, const double* a);                           /* simd.h:53-61 */

#ifdef __cplusplus
}
#endif

#endif
