/*
 * The oracle: CPU restatement of dim's CSR5 path (conversion, SpMV, CG).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the CPU leg of
 * the benchmark. Nothing in the product links or calls this file.
 *
 * Reference files restated (paths relative to the reference tree):
 *   include/dim/mat/storage_formats/csr5.h        format, conversion, SpMV
 *   include/dim/mat/storage_formats/csr5/calibrator.h
 *   include/dim/simd.h                            4-lane scan / sum / shuffle
 *   include/dim/vec.h                             dot / axpy / aypx
 *   benchmarks/conjugate_gradient/dim_sp/main.cpp CG step order
 *   docs/paper/thesis.tex:2133-2160               CG with a residual stop test
 *
 * Deliberate differences from the reference (its defects, not its semantics):
 *   - y is zeroed and the calibrator is fresh on every SpMV call: the reference
 *     requires both of its caller (csr5.h:526-538, 750-754) and its own CG
 *     drivers violate that.
 *   - a matrix with zero full tiles skips the full-tile and calibrator passes
 *     instead of dividing by a zero chunk size (csr5.h:608-617).
 * Arithmetic follows the reference's operation order, including its fused
 * multiply-adds (csr5.h:569,723) and the 4-lane scan association
 * (simd.h:34-45), so that at omega=4 the results are bit-identical to the
 * compiled reference header.
 */
#include "csr5_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define MAX_OMEGA 64

static uint32_t strip(uint32_t v) { return v & ~ORC_DIRTY_BIT; }
static int is_dirty(uint32_t v) { return (v & ORC_DIRTY_BIT) != 0; }

/* detail::upper_bound_idx (csr5.h:194-198): index of the last element <= v. */
static uint64_t upper_bound_idx(const uint32_t* a, uint64_t n, uint64_t v) {
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        uint64_t mid = lo + (hi - lo) / 2;
        if ((uint64_t)a[mid] <= v)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo - 1;
}

static int popcount64(uint64_t v) { return __builtin_popcountll(v); }

void orc_csr5_free(orc_csr5* m) {
    if (!m)
        return;
    free(m->tile_ptr);
    free(m->y_offset);
    free(m->scansum_offset);
    free(m->bit_flag);
    free(m->tile_desc_offset_ptr);
    free(m->tile_desc_offset);
    free(m->vals);
    free(m->col_idx);
    free(m->row_ptr);
    free(m);
}

/* csr5_info_t::generate_tile_pointer (csr5.h:411-431). */
static void generate_tile_pointer(orc_csr5* m) {
    const uint64_t tile = (uint64_t)m->sigma * m->omega;
    for (uint64_t t = 0; t <= m->tile_count; ++t) {
        uint64_t boundary = t * tile < m->nnz ? t * tile : m->nnz;
        m->tile_ptr[t] = (uint32_t)upper_bound_idx(m->row_ptr, (uint64_t)m->rows + 1, (uint32_t)boundary);
    }
    for (uint64_t t = 0; t < m->tile_count; ++t) {
        uint32_t start = strip(m->tile_ptr[t]), stop = strip(m->tile_ptr[t + 1]);
        if (start == stop)
            continue;
        /* detail::is_dirty(row span) (csr5.h:183-186): an empty row in [start, stop]. */
        for (uint32_t r = start; r < stop; ++r) {
            if (m->row_ptr[r] == m->row_ptr[r + 1]) {
                m->tile_ptr[t] |= ORC_DIRTY_BIT;
                break;
            }
        }
    }
}

/* set_bit_flag (csr5.h:323-340). */
static void set_bit_flag(orc_csr5* m) {
    const uint64_t sigma = m->sigma, omega = m->omega, tile = sigma * omega;
    for (uint64_t t = 0; t < m->tile_count; ++t) {
        uint32_t start = strip(m->tile_ptr[t]), stop = strip(m->tile_ptr[t + 1]);
        for (uint64_t rid = start; rid <= stop; ++rid) {
            uint64_t idx = m->row_ptr[rid];
            if (t != idx / tile)
                continue;
            uint64_t trow = idx % sigma, tcol = (idx / sigma) % omega;
            m->bit_flag[t * omega + tcol] |= (uint64_t)1 << trow;
        }
    }
}

/* tile_descriptor_t::set_scansum_and_y_offsets (csr5.h:74-97); returns the
 * number of segments in the tile (segn_scan[omega]). */
static uint32_t set_scansum_and_y_offsets(orc_csr5* m, uint64_t t) {
    const uint32_t omega = m->omega;
    uint32_t segn[MAX_OMEGA + 1];
    int present[MAX_OMEGA + 1];
    for (uint32_t c = 0; c < omega; ++c) {
        uint64_t bf = m->bit_flag[t * omega + c] | (c == 0 ? 1u : 0u); /* num_bits_set(!col_idx) */
        segn[c] = (uint32_t)popcount64(bf);
        present[c] = segn[c] != 0;
    }
    present[omega] = 1; /* the sentinel bit `1 << Omega` (csr5.h:71) */
    uint32_t acc = 0;
    for (uint32_t c = 0; c < omega; ++c) {
        uint32_t v = segn[c];
        m->y_offset[t * omega + c] = acc;
        acc += v;
    }
    for (uint32_t c = 0; c < omega; ++c) {
        uint32_t s = 0;
        if (present[c]) {
            /* countr_zero(present >> (c + 1)) */
            uint32_t k = c + 1;
            while (!present[k])
                ++k;
            s = k - c - 1;
        }
        m->scansum_offset[t * omega + c] = s;
    }
    return acc;
}

/* generate_tile_descriptor (csr5.h:399-409) with
 * set_tile_descriptor_y_and_segsum_offsets (342-361) and generate_tile_offset
 * (363-397). */
static int generate_tile_descriptor(orc_csr5* m) {
    const uint64_t sigma = m->sigma, omega = m->omega, tile = sigma * omega;
    const uint64_t T = m->tile_count;
    set_bit_flag(m);
    uint64_t total = 0;
    for (uint64_t t = 0; t < T; ++t) {
        uint32_t num_set = set_scansum_and_y_offsets(m, t);
        uint32_t start = strip(m->tile_ptr[t]), stop = strip(m->tile_ptr[t + 1]);
        if (start == stop)
            continue;
        if (is_dirty(m->tile_ptr[t])) {
            m->tile_desc_offset_ptr[t] = num_set + 1;
            total += num_set + 1;
        }
    }
    if (!total)
        return ORC_OK;
    /* std::exclusive_scan over all T+1 entries: entry T becomes the total. */
    uint32_t acc = 0;
    for (uint64_t t = 0; t <= T; ++t) {
        uint32_t v = m->tile_desc_offset_ptr[t];
        m->tile_desc_offset_ptr[t] = acc;
        acc += v;
    }
    m->n_tile_desc_offset = m->tile_desc_offset_ptr[T];
    m->tile_desc_offset = (uint32_t*)calloc(m->n_tile_desc_offset ? m->n_tile_desc_offset : 1, sizeof(uint32_t));
    if (!m->tile_desc_offset)
        return ORC_NO_MEMORY;
    for (uint64_t t = 0; t < T; ++t) {
        if (!is_dirty(m->tile_ptr[t]))
            continue;
        uint32_t start = strip(m->tile_ptr[t]), stop = strip(m->tile_ptr[t + 1]);
        const uint32_t* rows = m->row_ptr + start;
        const uint64_t nrows = (uint64_t)stop - start + 1;
        const uint32_t offset_ptr = m->tile_desc_offset_ptr[t];
        for (uint64_t c = 0; c < omega; ++c) {
            uint64_t bf = m->bit_flag[t * omega + c] | (c == 0 ? 1u : 0u);
            uint32_t y = m->y_offset[t * omega + c];
            for (uint64_t r = 0; r < sigma; ++r) {
                if (!((bf >> r) & 1))
                    continue;
                uint64_t global_idx = t * tile + c * sigma + r;
                m->tile_desc_offset[offset_ptr + y] = (uint32_t)upper_bound_idx(rows, nrows, global_idx);
                ++y;
            }
        }
    }
    return ORC_OK;
}

/* csr5::transpose (csr5.h:463-493): in full, non-fast-track tiles the element
 * at CSR offset c*sigma + r moves to r*omega + c. */
static int transpose(orc_csr5* m) {
    const uint64_t sigma = m->sigma, omega = m->omega, tile = sigma * omega;
    double* tv = (double*)malloc(tile * sizeof(double));
    uint32_t* tc = (uint32_t*)malloc(tile * sizeof(uint32_t));
    if (!tv || !tc) {
        free(tv);
        free(tc);
        return ORC_NO_MEMORY;
    }
    for (uint64_t t = 0; t < m->tile_count; ++t) {
        if (strip(m->tile_ptr[t]) == strip(m->tile_ptr[t + 1]))
            continue;
        double* v = m->vals + t * tile;
        uint32_t* ci = m->col_idx + t * tile;
        for (uint64_t c = 0; c < omega; ++c)
            for (uint64_t r = 0; r < sigma; ++r) {
                tv[r * omega + c] = v[c * sigma + r];
                tc[r * omega + c] = ci[c * sigma + r];
            }
        memcpy(v, tv, tile * sizeof(double));
        memcpy(ci, tc, tile * sizeof(uint32_t));
    }
    free(tv);
    free(tc);
    return ORC_OK;
}

int orc_csr5_from_csr(uint32_t sigma, uint32_t omega, uint32_t rows, uint32_t cols, uint64_t nnz,
                      const uint32_t* row_ptr, const uint32_t* col_idx, const double* vals, orc_csr5** out) {
    *out = NULL;
    if (sigma == 0 || sigma > 64 || omega == 0 || omega > MAX_OMEGA || !row_ptr)
        return ORC_INVALID_ARGUMENT;
    if (row_ptr[rows] != nnz || nnz >= ORC_DIRTY_BIT)
        return ORC_INVALID_ARGUMENT;
    orc_csr5* m = (orc_csr5*)calloc(1, sizeof(orc_csr5));
    if (!m)
        return ORC_NO_MEMORY;
    m->sigma = sigma;
    m->omega = omega;
    m->rows = rows;
    m->cols = cols;
    m->nnz = nnz;
    m->tile_count = nnz / ((uint64_t)sigma * omega);
    const uint64_t T = m->tile_count;
    m->tile_ptr = (uint32_t*)calloc(T + 1, sizeof(uint32_t));
    m->y_offset = (uint32_t*)calloc(T * omega + 1, sizeof(uint32_t));
    m->scansum_offset = (uint32_t*)calloc(T * omega + 1, sizeof(uint32_t));
    m->bit_flag = (uint64_t*)calloc(T * omega + 1, sizeof(uint64_t));
    m->tile_desc_offset_ptr = (uint32_t*)calloc(T + 1, sizeof(uint32_t));
    m->vals = (double*)malloc((nnz ? nnz : 1) * sizeof(double));
    m->col_idx = (uint32_t*)malloc((nnz ? nnz : 1) * sizeof(uint32_t));
    m->row_ptr = (uint32_t*)malloc(((uint64_t)rows + 1) * sizeof(uint32_t));
    if (!m->tile_ptr || !m->y_offset || !m->scansum_offset || !m->bit_flag || !m->tile_desc_offset_ptr || !m->vals ||
        !m->col_idx || !m->row_ptr) {
        orc_csr5_free(m);
        return ORC_NO_MEMORY;
    }
    if (nnz) {
        memcpy(m->vals, vals, nnz * sizeof(double));
        memcpy(m->col_idx, col_idx, nnz * sizeof(uint32_t));
    }
    memcpy(m->row_ptr, row_ptr, ((uint64_t)rows + 1) * sizeof(uint32_t));
    generate_tile_pointer(m);
    int st = generate_tile_descriptor(m);
    if (st == ORC_OK)
        st = transpose(m);
    if (st != ORC_OK) {
        orc_csr5_free(m);
        return st;
    }
    *out = m;
    return ORC_OK;
}

int orc_csr5_pack_desc(const orc_csr5* m, uint64_t tile, uint32_t* words) {
    if (m->sigma != 16 || m->omega != 4 || tile >= m->tile_count)
        return ORC_INVALID_ARGUMENT;
    for (uint32_t c = 0; c < 4; ++c) {
        uint64_t i = tile * 4 + c;
        words[c] = (m->y_offset[i] & 0x7Fu) | ((m->scansum_offset[i] & 0x7u) << 7) |
                   ((uint32_t)(m->bit_flag[i] & 0xFFFFu) << 10);
    }
    return ORC_OK;
}

/* ---------------------------------------------------------------- SpMV --- */

/* simd::inclusive_scan (simd.h:34-45) at 4 lanes, kept in its association;
 * a sequential scan for other lane counts. */
static void lane_inclusive_scan(uint32_t omega, const double* a, double* s) {
    if (omega == 4) {
        double u0 = a[0] + 0.0, u1 = a[1] + a[0], u2 = a[2] + a[1], u3 = a[3] + a[2];
        u0 = u0 + 0.0;
        u1 = u1 + 0.0;
        u2 = u2 + a[0];
        u3 = u3 + a[1];
        s[0] = u0 + 0.0;
        s[1] = u1 + 0.0;
        s[2] = u2 + 0.0;
        s[3] = u3 + a[0];
        return;
    }
    double acc = 0.0;
    for (uint32_t i = 0; i < omega; ++i) {
        acc = i ? acc + a[i] : a[i];
        s[i] = acc;
    }
}

/* detail::fast_segmented_sum (csr5.h:200-215). */
static void fast_segmented_sum(uint32_t omega, const double* vec, const uint32_t* off, double* out) {
    double t[MAX_OMEGA], s[MAX_OMEGA];
    for (uint32_t i = 0; i < omega; ++i)
        t[i] = i + 1 < omega ? vec[i + 1] : 0.0;
    lane_inclusive_scan(omega, t, s);
    for (uint32_t i = 0; i < omega; ++i)
        out[i] = (s[i + off[i]] - s[i]) + t[i];
}

/* simd::sum (simd.h:53-61) at 4 lanes; pairwise otherwise. */
static double lane_sum(uint32_t omega, const double* a) {
    if (omega == 4)
        return (a[0] + a[2]) + (a[1] + a[3]);
    double tmp[MAX_OMEGA];
    memcpy(tmp, a, omega * sizeof(double));
    for (uint32_t w = omega; w > 1;) {
        uint32_t h = (w + 1) / 2;
        for (uint32_t i = 0; i + h < w; ++i)
            tmp[i] = tmp[i] + tmp[i + h];
        w = h;
    }
    return tmp[0];
}

typedef struct chunk_ctx {
    const orc_csr5* m;
    const double* x;
    double* y;
    double* calibrator;
    uint64_t chunk;
} chunk_ctx;

/* spmv_thread_data::store_to_row_start (csr5.h:526-538). */
static void store_to_row_start(const chunk_ctx* ctx, uint64_t tid, uint32_t start_row_start, uint32_t row_start,
                               double val, int direct) {
    if (row_start == start_row_start && !direct) {
        ctx->calibrator[tid] += val;
        return;
    }
    if (direct)
        ctx->y[row_start] = val;
    else
        ctx->y[row_start] += val;
}

/* One thread's share of spmv_full_partitions (csr5.h:619-748). */
static void spmv_chunk(const chunk_ctx* ctx, uint64_t tid) {
    const orc_csr5* m = ctx->m;
    const uint32_t sigma = m->sigma, omega = m->omega;
    const uint64_t tile = (uint64_t)sigma * omega;
    const uint64_t T = m->tile_count;
    const uint64_t first = tid * ctx->chunk;
    const uint64_t last = first + ctx->chunk < T ? first + ctx->chunk : T;
    const uint32_t start_row_start = strip(m->tile_ptr[first]);
    const double* x = ctx->x;

    for (uint64_t t = first; t < last; ++t) {
        const double* vals = m->vals + t * tile;
        const uint32_t* cols = m->col_idx + t * tile;
        const uint32_t raw = m->tile_ptr[t];
        const uint32_t row_stop = strip(m->tile_ptr[t + 1]);
        const uint64_t* bf_in = m->bit_flag + t * omega;
        const uint32_t* yo_in = m->y_offset + t * omega;
        const uint32_t* ss_in = m->scansum_offset + t * omega;
        const int fast_direct = (int)(bf_in[0] & 1);

        if (raw == row_stop) {
            /* partition_fast_track (csr5.h:562-574) */
            double acc[MAX_OMEGA];
            for (uint32_t l = 0; l < omega; ++l)
                acc[l] = 0.0;
            for (uint32_t i = 0; i < sigma; ++i)
                for (uint32_t l = 0; l < omega; ++l)
                    acc[l] = fma(vals[i * omega + l], x[cols[i * omega + l]], acc[l]);
            store_to_row_start(ctx, tid, start_row_start, raw, lane_sum(omega, acc), fast_direct);
            continue;
        }

        const int empty_rows = is_dirty(raw);
        const uint32_t row_start = strip(raw);
        double* y_local = ctx->y + row_start;
        const uint32_t offset_pointer = empty_rows ? m->tile_desc_offset_ptr[t] - m->tile_desc_offset_ptr[0] : 0;
        const uint32_t* offs = m->tile_desc_offset;

        uint64_t bf[MAX_OMEGA];
        uint32_t y_off[MAX_OMEGA];
        int64_t start[MAX_OMEGA], stop[MAX_OMEGA], any_row_started[MAX_OMEGA];
        double sum[MAX_OMEGA], first_sum[MAX_OMEGA];

        const int first_all_direct = t == first && fast_direct;
        for (uint32_t l = 0; l < omega; ++l) {
            bf[l] = bf_in[l] | (l == 0 ? 1u : 0u);
            y_off[l] = yo_in[l] + (l == 0 ? 1u : 0u);
            int64_t lb = (int64_t)(bf[l] & 1);
            start[l] = 1 - lb;
            stop[l] = 0;
            any_row_started[l] = l == 0 ? 0 : lb;
            sum[l] = vals[l] * x[cols[l]];
            first_sum[l] = 0.0;
        }

        for (uint32_t i = 1; i < sigma; ++i) {
            int any = 0;
            for (uint32_t l = 0; l < omega; ++l)
                any |= (int)((bf[l] >> i) & 1);
            if (any) {
                for (uint32_t l = 0; l < omega; ++l) {
                    const int64_t lb = (int64_t)((bf[l] >> i) & 1);
                    if (any_row_started[l] && lb) {
                        const uint32_t yi = empty_rows ? offs[offset_pointer + y_off[l]] : y_off[l];
                        y_local[yi] = sum[l];
                        ++y_off[l];
                    }
                    if (!any_row_started[l] && lb)
                        first_sum[l] = sum[l];
                    if (lb)
                        sum[l] = 0.0;
                    any_row_started[l] |= lb;
                    stop[l] += lb;
                }
            }
            for (uint32_t l = 0; l < omega; ++l)
                sum[l] = fma(vals[i * omega + l], x[cols[i * omega + l]], sum[l]);
        }

        /* first_sum = row_active ? first_sum : sum (csr5.h:730-732) */
        for (uint32_t l = 0; l < omega; ++l)
            if (!any_row_started[l])
                first_sum[l] = sum[l];

        /* detail::compute_last_sum (csr5.h:217-235) */
        double fs[MAX_OMEGA], nps[MAX_OMEGA], last_sum[MAX_OMEGA];
        for (uint32_t l = 0; l < omega; ++l)
            fs[l] = start[l] == 1 ? first_sum[l] : 0.0;
        fast_segmented_sum(omega, fs, ss_in, nps);
        for (uint32_t l = 0; l < omega; ++l) {
            const double b = start[l] > stop[l] ? 0.0 : nps[l];
            last_sum[l] = sum[l] + b;
        }
        for (uint32_t l = 0; l < omega; ++l) {
            if (any_row_started[l]) {
                const uint32_t yi = empty_rows ? offs[offset_pointer + y_off[l]] : y_off[l];
                y_local[yi] = last_sum[l];
            }
        }
        store_to_row_start(ctx, tid, start_row_start, row_start, any_row_started[0] ? first_sum[0] : last_sum[0],
                           first_all_direct);
    }
}

int orc_csr5_spmv(const orc_csr5* m, uint32_t nthreads, const double* x, uint64_t x_len, double* y,
                  uint64_t y_len) {
    if (x_len != m->cols || y_len != m->rows || nthreads == 0)
        return ORC_INVALID_ARGUMENT;
    memset(y, 0, (size_t)y_len * sizeof(double));
    const uint64_t T = m->tile_count;

    if (T) {
        /* get_parallel_data (csr5.h:608-617) */
        const uint64_t chunk = (T + nthreads - 1) / nthreads;
        const uint64_t active = (T + chunk - 1) / chunk;
        double* calibrator = (double*)calloc(nthreads, sizeof(double));
        if (!calibrator)
            return ORC_NO_MEMORY;
        chunk_ctx ctx = {m, x, y, calibrator, chunk};
        /* every emulated thread's chunk is processed in tile order by one
         * worker, so the result does not depend on the worker count */
#pragma omp parallel for schedule(dynamic, 1)
        for (int64_t tid = 0; tid < (int64_t)active; ++tid)
            spmv_chunk(&ctx, (uint64_t)tid);
        /* spmv_calibrator (csr5.h:750-754) */
        const uint64_t ncal = nthreads < active ? nthreads : active;
        for (uint64_t i = 0; i < ncal; ++i)
            y[strip(m->tile_ptr[i * chunk])] += calibrator[i];
        free(calibrator);
    }

    /* spmv_tail_partition (csr5.h:756-782) */
    const uint64_t tile = (uint64_t)m->sigma * m->omega;
    const uint64_t first_tail_element = T * tile;
    const uint32_t tail_start = strip(m->tile_ptr[T]);
    const uint32_t tail_stop = m->rows;
#pragma omp parallel for schedule(static)
    for (int64_t row = tail_start; row < (int64_t)tail_stop; ++row) {
        uint64_t idx_start = row == tail_start ? first_tail_element : m->row_ptr[row];
        uint64_t idx_stop = m->row_ptr[row + 1];
        double s = 0.0;
        for (uint64_t idx = idx_start; idx < idx_stop; ++idx) {
            double p = m->vals[idx] * x[m->col_idx[idx]];
            s = s + p;
        }
        if (row == tail_start && m->row_ptr[row] != first_tail_element)
            y[row] += s;
        else
            y[row] = s;
    }
    return ORC_OK;
}

void orc_csr_spmv(uint32_t rows, const uint32_t* row_ptr, const uint32_t* col_idx, const double* vals,
                  const double* x, double* y) {
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < (int64_t)rows; ++r) {
        double s = 0.0;
        for (uint64_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k) {
            double p = vals[k] * x[col_idx[k]];
            s = s + p;
        }
        y[r] = s;
    }
}

int orc_csr_block_spmv(uint32_t rows, const uint32_t* row_ptr, const uint32_t* col_idx, const double* vals,
                       uint64_t nz_begin, uint64_t nz_end, const double* x, uint32_t first_row, uint32_t n_out,
                       double* y_block) {
    if (nz_begin > nz_end || nz_end > row_ptr[rows])
        return ORC_INVALID_ARGUMENT;
    memset(y_block, 0, (size_t)n_out * sizeof(double));
    for (uint64_t r = first_row; r < rows && r < (uint64_t)first_row + n_out; ++r) {
        uint64_t b = row_ptr[r] > nz_begin ? row_ptr[r] : nz_begin;
        uint64_t e = row_ptr[r + 1] < nz_end ? row_ptr[r + 1] : nz_end;
        double s = 0.0;
        for (uint64_t k = b; k < e; ++k) {
            double p = vals[k] * x[col_idx[k]];
            s = s + p;
        }
        y_block[r - first_row] = s;
    }
    return ORC_OK;
}

/* ------------------------------------------------------- vectors and CG --- */

#define DOT_BLOCK 4096

double orc_dot(uint64_t n, const double* a, const double* b) {
    const uint64_t nb = (n + DOT_BLOCK - 1) / DOT_BLOCK;
    double* part = (double*)malloc((nb ? nb : 1) * sizeof(double));
    if (!part)
        return NAN;
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < (int64_t)nb; ++k) {
        uint64_t lo = (uint64_t)k * DOT_BLOCK, hi = lo + DOT_BLOCK < n ? lo + DOT_BLOCK : n;
        double s = 0.0;
        for (uint64_t i = lo; i < hi; ++i) {
            double p = a[i] * b[i];
            s = s + p;
        }
        part[k] = s;
    }
    double s = 0.0;
    for (uint64_t k = 0; k < nb; ++k)
        s += part[k];
    free(part);
    return s;
}

void orc_axpy(uint64_t n, double* y, const double* x, double alpha) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        double p = alpha * x[i];
        y[i] = y[i] + p;
    }
}

void orc_aypx(uint64_t n, double* y, const double* x, double alpha) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        double p = alpha * y[i];
        y[i] = x[i] + p;
    }
}

int orc_cg(const orc_csr5* m, uint32_t nthreads, const double* b, double* x, double rtol, uint32_t maxit,
           orc_cg_result* res, double* history) {
    if (m->rows != m->cols)
        return ORC_INVALID_ARGUMENT;
    const uint64_t n = m->rows;
    double* r = (double*)malloc((n ? n : 1) * sizeof(double));
    double* s = (double*)malloc((n ? n : 1) * sizeof(double));
    double* As = (double*)malloc((n ? n : 1) * sizeof(double));
    if (!r || !s || !As) {
        free(r);
        free(s);
        free(As);
        return ORC_NO_MEMORY;
    }
    int st = ORC_OK;
    /* r = b - A x0 ; s = r ; r_r = r.r (thesis.tex:2137-2139) */
    orc_csr5_spmv(m, nthreads, x, n, As, n);
    for (uint64_t i = 0; i < n; ++i)
        r[i] = b[i] - As[i];
    memcpy(s, r, n * sizeof(double));
    double r_r = orc_dot(n, r, r);
    const double bnorm = sqrt(orc_dot(n, b, b));
    const double scale = bnorm > 0.0 ? bnorm : 1.0;
    double relres = sqrt(r_r) / scale;
    if (history)
        history[0] = relres;
    uint32_t it = 0;
    int converged = relres <= rtol;
    while (!converged && it < maxit) {
        orc_csr5_spmv(m, nthreads, s, n, As, n); /* As = A*s */
        const double sAs = orc_dot(n, s, As);
        if (!(sAs > 0.0)) {
            st = ORC_BREAKDOWN;
            break;
        }
        const double alpha = r_r / sAs;
        orc_axpy(n, x, s, alpha);   /* x += s*alpha */
        orc_axpy(n, r, As, -alpha); /* r -= As*alpha */
        const double r_r_new = orc_dot(n, r, r);
        ++it;
        relres = sqrt(r_r_new) / scale;
        if (history)
            history[it] = relres;
        if (relres <= rtol) {
            converged = 1;
            break;
        }
        const double beta = r_r_new / r_r;
        r_r = r_r_new;
        orc_aypx(n, s, r, beta); /* s = r + beta*s */
    }
    /* true residual */
    orc_csr5_spmv(m, nthreads, x, n, As, n);
    for (uint64_t i = 0; i < n; ++i)
        r[i] = b[i] - As[i];
    if (res) {
        res->iterations = it;
        res->converged = converged;
        res->relres = relres;
        res->relres_true = sqrt(orc_dot(n, r, r)) / scale;
    }
    free(r);
    free(s);
    free(As);
    return st;
}

/* ------------------------------------------------------------ test hooks --- */

void orc_lane_inclusive_scan(uint32_t omega, const double* a, double* s) { lane_inclusive_scan(omega, a, s); }
void orc_fast_segmented_sum(uint32_t omega, const double* vec, const uint32_t* off, double* out) {
    fast_segmented_sum(omega, vec, off, out);
}
void orc_shuffle_relative(uint32_t omega, const double* vec, const uint32_t* off, double* out) {
    for (uint32_t i = 0; i < omega; ++i)
        out[i] = vec[i + off[i]];
}
double orc_lane_sum(uint32_t omega, const double* a) { return lane_sum(omega, a); }
