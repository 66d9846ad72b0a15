// C ABI of libdim_gpu.so: matrix construction, products and the autotuner.
// The CG entry points live in cg.hip.
#include "handle.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <string>

namespace dimgpu {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string& msg) { g_last_error = msg; }

dim_status fail(dim_status st, const std::string& msg) {
    set_error(msg);
    return st;
}

dim_status hip_fail(hipError_t e, const char* what) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? DIM_OUT_OF_MEMORY : DIM_HIP_ERROR;
}

void free_handle(dim_csr5* h) {
    if (!h)
        return;
    int prev = 0;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(h->device);
    free_cg_cache(h->cg_cache);
    free_batch_cache(h->batch_cache);
    for (void* p : h->allocations)
        (void)hipFree(p);
    (void)hipSetDevice(prev);
    free_handle(h->xs);
    delete h;
}

dim_status upload_format(const HostFormat& f, bool keep_f32, dim_csr5** out, bool plain_columns) {
    std::unique_ptr<dim_csr5, void (*)(dim_csr5*)> h(new dim_csr5, free_handle);
    DIM_HIP_TRY(hipGetDevice(&h->device));
    DevFormat& d = h->f;
    d.rows = f.rows;
    d.cols = f.cols;
    d.sigma = f.sigma;
    d.tiles = f.tiles;
    d.nnz = f.nnz;
    d.padded = f.padded;
    d.any_dirty = f.any_dirty;
    d.n_empty = static_cast<uint32_t>(f.empty_rows.size());
    DIM_HIP_TRY(upload(h.get(), d.tile_ptr, f.tile_ptr));
    if (f.sigma <= 16) {
        std::vector<uint16_t> narrow(f.bit_flag.begin(), f.bit_flag.end());
        uint16_t* p = nullptr;
        DIM_HIP_TRY(upload(h.get(), p, narrow));
        d.bit_flag = p;
    } else {
        uint32_t* p = nullptr;
        DIM_HIP_TRY(upload(h.get(), p, f.bit_flag));
        d.bit_flag = p;
    }
    {
        // flag words are read 4 bytes at a time: pad to a multiple of 4 plus 4
        std::vector<uint8_t> flags(f.tile_flags);
        flags.resize((flags.size() + 3) / 4 * 4 + 4, 0);
        DIM_HIP_TRY(upload(h.get(), d.tile_flags, flags));
    }
    h->binfo.assign(f.tile_flags);
    d.binfo = &h->binfo;
    DIM_HIP_TRY(upload(h.get(), d.tile_desc_offset_ptr, f.tile_desc_offset_ptr));
    DIM_HIP_TRY(upload(h.get(), d.tile_desc_offset, f.tile_desc_offset));
    DIM_HIP_TRY(upload(h.get(), d.col_idx, f.col_idx));
    DIM_HIP_TRY(upload(h.get(), d.vals64, f.vals));
    if (keep_f32) {
        std::vector<float> v32(f.vals.begin(), f.vals.end());
        DIM_HIP_TRY(upload(h.get(), d.vals32, v32));
    }
    DIM_HIP_TRY(upload(h.get(), d.empty_rows, f.empty_rows));
    if (!f.col_dict.empty() && !plain_columns) {
        std::vector<int32_t> dict(f.col_dict);
        dict.resize(kMaxDiagonals, 0);
        DIM_HIP_TRY(upload(h.get(), d.col_dict, dict));
        DIM_HIP_TRY(upload(h.get(), d.col_code, f.col_code));
    }
    for (uint32_t t = 0; t < f.tiles; ++t)
        h->n_dirty += (f.tile_ptr[t] & kDirty) ? 1 : 0;

    Scratch& s = h->s;
    const size_t T = std::max<size_t>(f.tiles, 1);
    double* hp = nullptr;
    double* tp = nullptr;
    DIM_HIP_TRY(dev_alloc(h.get(), hp, T * kMaxK));
    DIM_HIP_TRY(dev_alloc(h.get(), tp, T * kMaxK));
    s.head = hp;
    s.tail = tp;
    DIM_HIP_TRY(dev_alloc(h.get(), s.chunk_flags, T));
    DIM_HIP_TRY(dev_alloc(h.get(), s.dot_chunk, T * kMaxK));
    DIM_HIP_TRY(dev_alloc(h.get(), s.dot_blk, ((T + 255) / 256) * kMaxK));
    DIM_HIP_TRY(upload(h.get(), s.queue, std::vector<uint32_t>(kQueueWords, 0u)));
    h->sys_row = {0, f.rows};
    h->sys_tile = {0, f.tiles};
    h->sys_nnz_csr = {0, f.nnz};
    h->nnz_logical = f.nnz;
    *out = h.release();
    return DIM_SUCCESS;
}

// Candidate tile heights: the supported sigma nearest to 1..4 times the mean
// row length (a lane then tends to start on a row and keep one row per step
// of the wavefront, so the x gathers of a step are close together), plus 4,
// 8 and 16 for irregular rows.
std::vector<uint32_t> sigma_candidates(uint32_t rows, uint64_t nnz) {
    static const uint32_t supported[] = {
#define DIM_SIGMA_ITEM(S) S,
      DIM_FOR_EACH_SIGMA(DIM_SIGMA_ITEM)
#undef DIM_SIGMA_ITEM
    };
    std::vector<uint32_t> c = {4, 8, 16};
    const double len = rows ? static_cast<double>(nnz) / rows : 8.0;
    for (int m = 1; m <= 4; ++m) {
        const double want = m * len;
        uint32_t best = supported[0];
        for (uint32_t s : supported)
            if (std::abs(static_cast<double>(s) - want) < std::abs(static_cast<double>(best) - want))
                best = s;
        c.push_back(best);
    }
    std::sort(c.begin(), c.end());
    c.erase(std::unique(c.begin(), c.end()), c.end());
    return c;
}

DevFormat system_view(const dim_csr5* A, uint32_t s) {
    DevFormat view = A->f;
    const uint32_t t0 = A->sys_tile[s], t1 = A->sys_tile[s + 1];
    const size_t tile = size_t{kOmega} * A->f.sigma;
    view.tiles = t1 - t0;
    view.tile_ptr = A->f.tile_ptr + t0;
    view.bit_flag = static_cast<char*>(A->f.bit_flag) + size_t{t0} * kOmega * (A->f.sigma <= 16 ? 2 : 4);
    view.tile_flags = A->f.tile_flags + t0;
    view.binfo_base = A->f.binfo_base + t0;
    view.tile_desc_offset_ptr = A->f.tile_desc_offset_ptr + t0;
    view.col_idx = A->f.col_idx + t0 * tile;
    view.vals64 = A->f.vals64 + t0 * tile;
    view.vals32 = A->f.vals32 ? A->f.vals32 + t0 * tile : nullptr;
    view.col_code = A->f.col_code ? A->f.col_code + size_t{t0} * kOmega : nullptr;
    // a batch with column codes keeps one table of diagonals per system
    view.col_dict = A->f.col_dict && A->n_systems > 1 ? A->f.col_dict + size_t{s} * kMaxDiagonals : A->f.col_dict;
    view.chunk_tile = nullptr;
    view.chunk_sys = nullptr;
    view.n_chunk_tab = 0;
    view.nnz = A->sys_nnz_csr[s + 1] - A->sys_nnz_csr[s];
    view.padded = size_t{view.tiles} * tile;
    return view;
}

namespace {

dim_status check_handle(const dim_csr5* A) {
    if (!A)
        return fail(DIM_INVALID_ARGUMENT, "matrix handle is null");
    return DIM_SUCCESS;
}

// row-major multi-RHS blocks with at least this many sides go through
// column-major staging
constexpr uint32_t kRowMajorStagingK = 6;

// ---- Column slices ---------------------------------------------------------
// A product that gathers x entries from all over a large x (random columns)
// is bound by the rate of L2 misses, not by HBM: every XCD's L2 (4 MB) holds a
// small share of x. Splitting the columns into kXcdSlices slices of equal
// element count and stacking the slices as row blocks (row s*rows + i of the
// stacked matrix holds the elements of row i whose columns lie in slice s, in
// CSR order) lets the tile kernel, whose XCD remap gives each XCD one
// contiguous eighth of the tiles, gather each XCD's elements from about one
// eighth of x, which then stays in that XCD's L2. k_slice_sum adds the
// per-slice partial products of each row in slice order. Measured on the
// MI355X: 62M gathers from 32 MB take 1.01 ms at random and 0.26 ms sliced.
constexpr uint32_t kXcdSlices = 8;

// The stacked matrix keeps only the (slice, row) pairs that hold elements, so
// that it has no empty rows (and no dirty tiles): stacked row k is the k-th
// set bit of `present`, bit s*rows + i. rank[w] counts the set bits in the
// words before word w.
struct StackedCsr {
    uint32_t rows = 0;
    std::vector<uint32_t> row_ptr, col_idx;
    std::vector<double> vals;
    std::vector<uint32_t> present, rank;
};

// x far beyond one L2, rows long enough that one partial per slice and row
// costs less than the gathers it localises, and row indices that still fit
bool slices_eligible(uint32_t rows, uint32_t cols, uint64_t nnz) {
    return uint64_t{cols} * sizeof(double) >= (16ull << 20) && nnz >= uint64_t{kXcdSlices} * rows &&
           uint64_t{rows} * kXcdSlices < (1ull << 31);
}

StackedCsr stack_slices(uint32_t rows, uint32_t cols, uint64_t nnz, const uint32_t* row_ptr, const uint32_t* col_idx,
                        const double* vals, uint32_t S) {
    // column cuts at equal element counts: slice of column c = number of cuts <= c
    std::vector<uint64_t> below(size_t{cols} + 1, 0); // below[c]: elements with column < c
    for (uint64_t e = 0; e < nnz; ++e)
        ++below[size_t{col_idx[e]} + 1];
    for (size_t c = 1; c <= cols; ++c)
        below[c] += below[c - 1];
    std::vector<uint8_t> slice_of(cols, 0);
    {
        uint32_t q = 0;
        for (uint32_t c = 0; c < cols; ++c) {
            while (q + 1 < S && below[c] * S >= nnz * (q + 1) && nnz)
                ++q;
            slice_of[c] = static_cast<uint8_t>(q);
        }
    }
    const size_t pairs = size_t{S} * rows;
    std::vector<uint32_t> count(pairs, 0); // elements of (slice, row), slice-major
    for (uint32_t i = 0; i < rows; ++i)
        for (uint32_t e = row_ptr[i]; e < row_ptr[i + 1]; ++e)
            ++count[size_t{slice_of[col_idx[e]]} * rows + i];
    StackedCsr out;
    out.present.assign((pairs + 31) / 32, 0);
    out.rank.assign(out.present.size(), 0);
    out.row_ptr.push_back(0);
    std::vector<uint32_t> at(pairs, 0); // next element slot of each pair
    for (size_t r = 0; r < pairs; ++r) {
        if (!(r & 31))
            out.rank[r >> 5] = out.rows;
        if (count[r]) {
            out.present[r >> 5] |= 1u << (r & 31);
            at[r] = out.row_ptr.back();
            out.row_ptr.push_back(out.row_ptr.back() + count[r]);
            ++out.rows;
        }
    }
    out.col_idx.resize(nnz);
    out.vals.resize(nnz);
    for (uint32_t i = 0; i < rows; ++i)
        for (uint32_t e = row_ptr[i]; e < row_ptr[i + 1]; ++e) {
            const uint32_t k = at[size_t{slice_of[col_idx[e]]} * rows + i]++;
            out.col_idx[k] = col_idx[e];
            out.vals[k] = vals[e];
        }
    return out;
}

// Builds the stacked matrix's format with tile height sigma, with its
// partial-product buffer and the bit map (and word ranks) of the pairs.
dim_status build_slices(const StackedCsr& sc, uint32_t cols, uint32_t sigma, bool keep_f32, dim_csr5** out) {
    HostFormat f;
    const std::string err = build_format(sc.rows, cols, sc.vals.size(), sc.row_ptr.data(), sc.col_idx.data(),
                                         sc.vals.data(), sigma, f);
    if (!err.empty())
        return fail(DIM_INVALID_ARGUMENT, err);
    dim_csr5* h = nullptr;
    if (dim_status st = upload_format(f, keep_f32, &h))
        return st;
    std::unique_ptr<dim_csr5, void (*)(dim_csr5*)> xs(h, free_handle);
    DIM_HIP_TRY(upload(xs.get(), xs->present, sc.present));
    DIM_HIP_TRY(upload(xs.get(), xs->present_rank, sc.rank));
    double* ys = nullptr;
    DIM_HIP_TRY(dev_alloc(xs.get(), ys, sc.rows));
    xs->ystack = ys;
    *out = xs.release();
    return DIM_SUCCESS;
}

// The whole-matrix single-vector product, through the column slices when the
// matrix has them.
hipError_t plain_product(dim_csr5* h, bool f32, const void* x, void* y, hipStream_t st) {
    if (h->xs) {
        dim_csr5* xs = h->xs;
        hipError_t e = launch_spmv(xs->f, xs->s, f32, 1, x, xs->ystack, 0, xs->f.tiles, 0, 0, xs->f.n_empty, false,
                                   nullptr, st);
        if (e != hipSuccess)
            return e;
        return launch_slice_sum(f32, xs->ystack, xs->present, xs->present_rank, h->f.rows, h->xs_slices, y, st);
    }
    return launch_spmv(h->f, h->s, f32, 1, x, y, 0, h->f.tiles, 0, 0, h->f.n_empty, false, nullptr, st);
}

// Times candidate tile heights on this matrix and keeps the fastest:
// sigma trades per-tile fixed work (descriptor decode, wavefront scan, carry
// slots) against registers per lane and load-balance granularity.
dim_status autotune_sigma(uint32_t rows, uint32_t cols, uint64_t nnz, const uint32_t* row_ptr,
                          const uint32_t* col_idx, const double* vals, bool keep_f32, bool force_slices,
                          bool plain_columns, dim_csr5** out) {
    const std::vector<uint32_t> candidates = sigma_candidates(rows, nnz);
    dim_csr5* best = nullptr;
    struct Tried {
        dim_csr5* h;
        float ms;
    };
    std::vector<Tried> tried;
    double* x = nullptr;
    double* y = nullptr;
    DIM_HIP_TRY(hipMalloc(&x, std::max<size_t>(cols, 1) * sizeof(double)));
    if (hipMalloc(&y, std::max<size_t>(rows, 1) * sizeof(double)) != hipSuccess) {
        (void)hipFree(x);
        return fail(DIM_OUT_OF_MEMORY, "autotune: cannot allocate y");
    }
    {
        std::vector<double> hx(std::max<size_t>(cols, 1));
        std::mt19937_64 rng(12345);
        std::uniform_real_distribution<double> U(-1.0, 1.0);
        for (auto& v : hx)
            v = U(rng);
        (void)hipMemcpy(x, hx.data(), hx.size() * sizeof(double), hipMemcpyHostToDevice);
    }
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    // best of three rounds of ten products: the rounds absorb clock and queue
    // noise that a single round shows
    auto time_products = [&](dim_csr5* h, float* ms) -> dim_status {
        for (int i = 0; i < 3; ++i)
            (void)plain_product(h, false, x, y, 0);
        const int reps = 10;
        hipError_t e = hipSuccess;
        for (int round = 0; round < 3 && e == hipSuccess; ++round) {
            (void)hipEventRecord(e0, 0);
            for (int i = 0; i < reps; ++i)
                (void)plain_product(h, false, x, y, 0);
            (void)hipEventRecord(e1, 0);
            e = hipEventSynchronize(e1);
            float r = 0;
            (void)hipEventElapsedTime(&r, e0, e1);
            if (round == 0 || r < *ms)
                *ms = r;
        }
        return e == hipSuccess ? DIM_SUCCESS : hip_fail(e, "autotune timing");
    };
    dim_status st = DIM_SUCCESS;
    for (uint32_t sigma : candidates) {
        HostFormat f;
        std::string err = build_format(rows, cols, nnz, row_ptr, col_idx, vals, sigma, f);
        if (!err.empty()) {
            st = fail(DIM_INVALID_ARGUMENT, err);
            break;
        }
        dim_csr5* h = nullptr;
        st = upload_format(f, keep_f32, &h, plain_columns);
        if (st != DIM_SUCCESS)
            break;
        float ms = 0;
        st = time_products(h, &ms);
        if (st != DIM_SUCCESS) {
            free_handle(h);
            break;
        }
        tried.push_back({h, ms});
    }
    // Column slices: the stacked matrix at the tile heights near its (short)
    // mean row length, timed as whole products through the first candidate.
    dim_csr5* xs_best = nullptr;
    float xs_ms = 0;
    if (st == DIM_SUCCESS && !tried.empty() && nnz && (force_slices || slices_eligible(rows, cols, nnz))) {
        const StackedCsr sc = stack_slices(rows, cols, nnz, row_ptr, col_idx, vals, kXcdSlices);
        dim_csr5* probe = tried[0].h;
        for (uint32_t sigma : {4u, 8u, 16u}) {
            dim_csr5* xs = nullptr;
            st = build_slices(sc, cols, sigma, keep_f32, &xs);
            if (st != DIM_SUCCESS)
                break;
            probe->xs = xs;
            probe->xs_slices = kXcdSlices;
            float ms = 0;
            st = time_products(probe, &ms);
            probe->xs = nullptr;
            probe->xs_slices = 0;
            if (st != DIM_SUCCESS) {
                free_handle(xs);
                break;
            }
            if (!xs_best || ms < xs_ms) {
                free_handle(xs_best);
                xs_best = xs;
                xs_ms = ms;
            } else {
                free_handle(xs);
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(x);
    (void)hipFree(y);
    if (st == DIM_SUCCESS && !tried.empty()) {
        // Candidates within 3% of the fastest are ties (timing noise): among
        // them take the sigma nearest the mean row length, so that repeated
        // conversions of one matrix pick the same format.
        float fastest = tried[0].ms;
        for (const auto& c : tried)
            fastest = std::min(fastest, c.ms);
        const double mean_len = rows ? static_cast<double>(nnz) / rows : 0.0;
        for (const auto& c : tried) {
            if (c.ms > 1.03f * fastest)
                continue;
            if (!best || std::abs(c.h->f.sigma - mean_len) < std::abs(best->f.sigma - mean_len))
                best = c.h;
        }
    }
    if (st == DIM_SUCCESS && best && xs_best) {
        float fastest = tried[0].ms;
        for (const auto& c : tried)
            fastest = std::min(fastest, c.ms);
        if (force_slices || xs_ms < 0.97f * fastest) {
            best->xs = xs_best;
            best->xs_slices = kXcdSlices;
            xs_best = nullptr;
        }
    }
    free_handle(xs_best);
    for (const auto& c : tried)
        if (c.h != best)
            free_handle(c.h);
    if (st != DIM_SUCCESS) {
        free_handle(best);
        return st;
    }
    *out = best;
    return DIM_SUCCESS;
}

} // namespace

dim_status stage_host_vectors(dim_csr5* A, char** x, char** y) {
    if (!A->host_stage) {
        double* buf = nullptr;
        DIM_HIP_TRY(dev_alloc(A, buf, size_t{A->f.cols} + A->f.rows));
        A->host_stage = buf;
    }
    *x = static_cast<char*>(A->host_stage);
    *y = *x + size_t{A->f.cols} * sizeof(double);
    return DIM_SUCCESS;
}

} // namespace dimgpu

using namespace dimgpu;

extern "C" {

const char* dim_last_error(void) { return g_last_error.c_str(); }

const char* dim_status_string(dim_status s) {
    switch (s) {
        case DIM_SUCCESS: return "success";
        case DIM_INVALID_ARGUMENT: return "invalid argument";
        case DIM_HIP_ERROR: return "HIP error";
        case DIM_OUT_OF_MEMORY: return "out of memory";
        case DIM_NOT_CONVERGED: return "not converged";
        case DIM_BREAKDOWN: return "breakdown (matrix not SPD)";
        case DIM_IO_ERROR: return "I/O error";
        case DIM_UNSUPPORTED: return "unsupported";
    }
    return "unknown status";
}

int dim_version(void) { return DIM_GPU_VERSION; }

dim_status dim_csr5_from_csr(uint32_t rows, uint32_t cols, uint64_t nnz, const uint32_t* row_ptr,
                             const uint32_t* col_idx, const double* vals, const dim_csr5_opts* opts,
                             dim_csr5** out) {
    if (!out)
        return fail(DIM_INVALID_ARGUMENT, "out is null");
    *out = nullptr;
    if (nnz && !vals)
        return fail(DIM_INVALID_ARGUMENT, "vals is null");
    std::string err = validate_csr(rows, cols, nnz, row_ptr, col_idx);
    if (!err.empty())
        return fail(DIM_INVALID_ARGUMENT, err);
    const uint32_t sigma = opts ? opts->sigma : 16;
    const bool keep_f32 = opts && (opts->flags & DIM_CSR5_KEEP_F32);
    const bool slices = opts && (opts->flags & DIM_CSR5_XCD_SLICES);
    const bool plain_columns = opts && (opts->flags & DIM_CSR5_PLAIN_COLUMNS);
    if (slices && uint64_t{rows} * kXcdSlices >= (1ull << 31))
        return fail(DIM_INVALID_ARGUMENT, "DIM_CSR5_XCD_SLICES needs rows < 2^28");
    if (sigma == 0)
        return autotune_sigma(rows, cols, nnz, row_ptr, col_idx, vals, keep_f32, slices, plain_columns, out);
    HostFormat f;
    err = build_format(rows, cols, nnz, row_ptr, col_idx, vals, sigma, f);
    if (!err.empty())
        return fail(DIM_INVALID_ARGUMENT, err);
    dim_csr5* h = nullptr;
    if (dim_status st = upload_format(f, keep_f32, &h, plain_columns))
        return st;
    if (slices && nnz) {
        const StackedCsr sc = stack_slices(rows, cols, nnz, row_ptr, col_idx, vals, kXcdSlices);
        if (dim_status st = build_slices(sc, cols, sigma, keep_f32, &h->xs)) {
            free_handle(h);
            return st;
        }
        h->xs_slices = kXcdSlices;
    }
    *out = h;
    return DIM_SUCCESS;
}

dim_status dim_csr5_destroy(dim_csr5* A) {
    free_handle(A);
    return DIM_SUCCESS;
}

dim_status dim_csr5_get_info(const dim_csr5* A, dim_csr5_info* info) {
    if (dim_status st = check_handle(A))
        return st;
    if (!info)
        return fail(DIM_INVALID_ARGUMENT, "info is null");
    info->rows = A->f.rows;
    info->cols = A->f.cols;
    info->nnz = A->nnz_logical;
    info->sigma = A->f.sigma;
    info->omega = kOmega;
    info->tile_count = A->f.tiles;
    info->tail_partition_start = A->f.rows;
    info->n_empty_rows = A->f.n_empty;
    info->n_dirty_tiles = A->n_dirty;
    info->has_f32 = A->f.vals32 != nullptr;
    info->device_bytes = A->device_bytes + (A->xs ? A->xs->device_bytes : 0);
    info->xcd_slices = A->xs ? A->xs_slices : 0;
    info->xcd_sigma = A->xs ? A->xs->f.sigma : 0;
    info->column_codes = A->f.col_code != nullptr;
    return DIM_SUCCESS;
}

dim_status dim_csr5_copy_format(const dim_csr5* A, uint32_t* tile_ptr, uint32_t* bit_flag, uint8_t* tile_flags,
                                uint32_t* tile_desc_offset_ptr, uint32_t* tile_desc_offset, uint64_t* n_tdo,
                                uint32_t* empty_rows) {
    if (dim_status st = check_handle(A))
        return st;
    const DevFormat& d = A->f;
    const size_t T = d.tiles;
    if (tile_ptr)
        DIM_HIP_TRY(hipMemcpy(tile_ptr, d.tile_ptr, (T + 1) * 4, hipMemcpyDeviceToHost));
    if (bit_flag && T) {
        if (d.sigma <= 16) {
            std::vector<uint16_t> tmp(T * kOmega);
            DIM_HIP_TRY(hipMemcpy(tmp.data(), d.bit_flag, tmp.size() * 2, hipMemcpyDeviceToHost));
            std::copy(tmp.begin(), tmp.end(), bit_flag);
        } else {
            DIM_HIP_TRY(hipMemcpy(bit_flag, d.bit_flag, T * kOmega * 4, hipMemcpyDeviceToHost));
        }
    }
    if (tile_flags)
        DIM_HIP_TRY(hipMemcpy(tile_flags, d.tile_flags, T + 1, hipMemcpyDeviceToHost));
    uint32_t ntdo = 0;
    DIM_HIP_TRY(hipMemcpy(&ntdo, d.tile_desc_offset_ptr + T, 4, hipMemcpyDeviceToHost));
    if (tile_desc_offset_ptr)
        DIM_HIP_TRY(hipMemcpy(tile_desc_offset_ptr, d.tile_desc_offset_ptr, (T + 1) * 4, hipMemcpyDeviceToHost));
    if (tile_desc_offset && ntdo)
        DIM_HIP_TRY(hipMemcpy(tile_desc_offset, d.tile_desc_offset, size_t{ntdo} * 4, hipMemcpyDeviceToHost));
    if (n_tdo)
        *n_tdo = ntdo;
    if (empty_rows && d.n_empty)
        DIM_HIP_TRY(hipMemcpy(empty_rows, d.empty_rows, size_t{d.n_empty} * 4, hipMemcpyDeviceToHost));
    return DIM_SUCCESS;
}

static dim_status check_dtype(const dim_csr5* A, dim_dtype dtype) {
    if (dtype != DIM_F64 && dtype != DIM_F32)
        return fail(DIM_INVALID_ARGUMENT, "dtype must be DIM_F64 or DIM_F32");
    if (dtype == DIM_F32 && !A->f.vals32)
        return fail(DIM_INVALID_ARGUMENT, "matrix was built without DIM_CSR5_KEEP_F32");
    return DIM_SUCCESS;
}

dim_status dim_csr5_spmv(const dim_csr5* A, dim_dtype dtype, const void* x, uint64_t x_len, void* y,
                         uint64_t y_len, void* stream) {
    if (dim_status st = check_handle(A))
        return st;
    if (dim_status st = check_dtype(A, dtype))
        return st;
    if (x_len != A->f.cols)
        return fail(DIM_INVALID_ARGUMENT, "can't multiply, vector size doesn't match column count of the matrix");
    if (y_len != A->f.rows)
        return fail(DIM_INVALID_ARGUMENT, "output vector size doesn't match row count of the matrix");
    if ((A->f.cols && !x) || (A->f.rows && !y))
        return fail(DIM_INVALID_ARGUMENT, "null vector");
    auto* self = const_cast<dim_csr5*>(A);
    DIM_HIP_TRY(plain_product(self, dtype == DIM_F32, x, y, static_cast<hipStream_t>(stream)));
    return DIM_SUCCESS;
}

dim_status dim_csr5_spmv_host(const dim_csr5* A, dim_dtype dtype, const void* x, uint64_t x_len, void* y,
                              uint64_t y_len, void* stream) {
    if (dim_status st = check_handle(A))
        return st;
    if (dim_status st = check_dtype(A, dtype))
        return st;
    if (x_len != A->f.cols)
        return fail(DIM_INVALID_ARGUMENT, "can't multiply, vector size doesn't match column count of the matrix");
    if (y_len != A->f.rows)
        return fail(DIM_INVALID_ARGUMENT, "output vector size doesn't match row count of the matrix");
    if ((A->f.cols && !x) || (A->f.rows && !y))
        return fail(DIM_INVALID_ARGUMENT, "null vector");
    auto* self = const_cast<dim_csr5*>(A);
    const auto st = static_cast<hipStream_t>(stream);
    const size_t esz = dtype == DIM_F32 ? sizeof(float) : sizeof(double);
    char *dx = nullptr, *dy = nullptr;
    if (dim_status s = stage_host_vectors(self, &dx, &dy))
        return s;
    DIM_HIP_TRY(hipMemcpyAsync(dx, x, size_t{A->f.cols} * esz, hipMemcpyHostToDevice, st));
    DIM_HIP_TRY(plain_product(self, dtype == DIM_F32, dx, dy, st));
    DIM_HIP_TRY(hipMemcpyAsync(y, dy, size_t{A->f.rows} * esz, hipMemcpyDeviceToHost, st));
    DIM_HIP_TRY(hipStreamSynchronize(st));
    return DIM_SUCCESS;
}

dim_status dim_csr5_spmv_timed(const dim_csr5* A, dim_dtype dtype, const void* x, uint64_t x_len, void* y,
                               uint64_t y_len, void* stream, float* ms) {
    if (dim_status st = check_handle(A))
        return st;
    if (dim_status st = check_dtype(A, dtype))
        return st;
    if (x_len != A->f.cols || y_len != A->f.rows || !ms)
        return fail(DIM_INVALID_ARGUMENT, "vector size mismatch or null ms");
    hipEvent_t ev[3] = {};
    for (auto& e : ev)
        DIM_HIP_TRY(hipEventCreate(&e));
    auto* self = const_cast<dim_csr5*>(A);
    bool fixup = true;
    hipError_t e;
    if (self->xs) {
        // the stacked matrix's tile kernel; its fix-up (if any) and the slice sum
        const auto st = static_cast<hipStream_t>(stream);
        e = launch_spmv_events(self->xs->f, self->xs->s, dtype == DIM_F32, x, self->xs->ystack, ev[0], ev[1], ev[2],
                               st, &fixup);
        if (e == hipSuccess)
            e = launch_slice_sum(dtype == DIM_F32, self->xs->ystack, self->xs->present, self->xs->present_rank,
                                 self->f.rows, self->xs_slices, y, st);
        if (e == hipSuccess)
            e = hipEventRecord(ev[2], st);
        fixup = true;
    } else {
        e = launch_spmv_events(self->f, self->s, dtype == DIM_F32, x, y, ev[0], ev[1], ev[2],
                               static_cast<hipStream_t>(stream), &fixup);
    }
    if (e == hipSuccess)
        e = hipEventSynchronize(ev[2]);
    if (e == hipSuccess)
        e = hipEventElapsedTime(&ms[0], ev[0], ev[1]);
    if (e == hipSuccess)
        e = hipEventElapsedTime(&ms[1], ev[1], ev[2]);
    if (!fixup)
        ms[1] = 0.0f;
    for (auto& ee : ev)
        (void)hipEventDestroy(ee);
    if (e != hipSuccess)
        return hip_fail(e, "dim_csr5_spmv_timed");
    return DIM_SUCCESS;
}

dim_status dim_csr5_block_rows(const dim_csr5* A, uint32_t tile_begin, uint32_t tile_end, uint32_t* first_row,
                               uint32_t* n_rows) {
    if (dim_status st = check_handle(A))
        return st;
    if (tile_begin >= tile_end || tile_end > A->f.tiles)
        return fail(DIM_INVALID_ARGUMENT, "tile range must satisfy tile_begin < tile_end <= tile_count");
    uint32_t tp[2];
    DIM_HIP_TRY(hipMemcpy(&tp[0], A->f.tile_ptr + tile_begin, 4, hipMemcpyDeviceToHost));
    DIM_HIP_TRY(hipMemcpy(&tp[1], A->f.tile_ptr + tile_end, 4, hipMemcpyDeviceToHost));
    const uint32_t first = tp[0] & ~kDirty;
    // the block's last row: the row holding its last element
    uint32_t last;
    if (tile_end == A->f.tiles) {
        last = A->f.rows ? A->f.rows - 1 : 0;
    } else {
        uint8_t fl = 0;
        DIM_HIP_TRY(hipMemcpy(&fl, A->f.tile_flags + tile_end, 1, hipMemcpyDeviceToHost));
        const uint32_t next = tp[1] & ~kDirty;
        last = (fl & kStartsRow) ? next - 1 : next;
    }
    if (first_row)
        *first_row = first;
    if (n_rows)
        *n_rows = last >= first ? last - first + 1 : 0;
    return DIM_SUCCESS;
}

dim_status dim_csr5_spmv_block(const dim_csr5* A, dim_dtype dtype, uint32_t tile_begin, uint32_t tile_end,
                               const void* x, uint64_t x_len, void* y_block, uint64_t y_len, void* stream) {
    if (dim_status st = check_handle(A))
        return st;
    if (dim_status st = check_dtype(A, dtype))
        return st;
    if (x_len != A->f.cols)
        return fail(DIM_INVALID_ARGUMENT, "can't multiply, vector size doesn't match column count of the matrix");
    if (A->n_systems > 1 && A->f.col_code)
        return fail(DIM_UNSUPPORTED, "row-block products of a batch with per-system column codes");
    uint32_t first = 0, n = 0;
    if (dim_status st = dim_csr5_block_rows(A, tile_begin, tile_end, &first, &n))
        return st;
    if (y_len != n)
        return fail(DIM_INVALID_ARGUMENT, "y_block must have last_row - first_row + 1 entries");
    // empty rows inside [first, first + n)
    std::vector<uint32_t> er(A->f.n_empty);
    if (A->f.n_empty)
        DIM_HIP_TRY(hipMemcpy(er.data(), A->f.empty_rows, er.size() * 4, hipMemcpyDeviceToHost));
    const uint32_t eb = static_cast<uint32_t>(std::lower_bound(er.begin(), er.end(), first) - er.begin());
    const uint32_t ee = static_cast<uint32_t>(std::lower_bound(er.begin(), er.end(), first + n) - er.begin());
    auto* self = const_cast<dim_csr5*>(A);
    DIM_HIP_TRY(launch_spmv(self->f, self->s, dtype == DIM_F32, 1, x, y_block, tile_begin, tile_end, first, eb, ee,
                            false, nullptr, static_cast<hipStream_t>(stream)));
    return DIM_SUCCESS;
}

dim_status dim_csr5_spmm(const dim_csr5* A, dim_dtype dtype, uint32_t k, const void* X, void* Y, void* stream) {
    if (dim_status st = check_handle(A))
        return st;
    if (dim_status st = check_dtype(A, dtype))
        return st;
    if (k < 1 || k > static_cast<uint32_t>(kMaxK))
        return fail(DIM_INVALID_ARGUMENT, "k must be in 1..16");
    if ((A->f.cols && !X) || (A->f.rows && !Y))
        return fail(DIM_INVALID_ARGUMENT, "null matrix argument");
    auto* self = const_cast<dim_csr5*>(A);
    const auto st = static_cast<hipStream_t>(stream);
    const bool f32 = dtype == DIM_F32;
    if (k >= kRowMajorStagingK && A->f.rows && A->f.cols) {
        // From 6 sides on, the column-major product plus two transposes beats
        // gathering k-wide rows (one cache line per lane and side).
        const size_t esz = f32 ? 4 : 8;
        const size_t xb = size_t{A->f.cols} * k * esz, need = xb + size_t{A->f.rows} * k * esz;
        if (self->mm_bytes < need) {
            if (self->mm_scratch) {
                auto it = std::find(self->allocations.begin(), self->allocations.end(), self->mm_scratch);
                if (it != self->allocations.end())
                    self->allocations.erase(it);
                (void)hipFree(self->mm_scratch);
                self->device_bytes -= self->mm_bytes;
                self->mm_scratch = nullptr;
                self->mm_bytes = 0;
            }
            char* buf = nullptr;
            DIM_HIP_TRY(dev_alloc(self, buf, need));
            self->mm_scratch = buf;
            self->mm_bytes = need;
        }
        char* xt = static_cast<char*>(self->mm_scratch);
        char* yt = xt + xb;
        DIM_HIP_TRY(launch_block_transpose(f32, true, A->f.cols, static_cast<int>(k), X, xt, st));
        DIM_HIP_TRY(launch_spmv(self->f, self->s, f32, static_cast<int>(k), xt, yt, 0, A->f.tiles, 0, 0, A->f.n_empty,
                                false, nullptr, st, nullptr, A->f.cols, A->f.rows));
        DIM_HIP_TRY(launch_block_transpose(f32, false, A->f.rows, static_cast<int>(k), yt, Y, st));
        return DIM_SUCCESS;
    }
    DIM_HIP_TRY(launch_spmv(self->f, self->s, f32, static_cast<int>(k), X, Y, 0, A->f.tiles, 0, 0, A->f.n_empty, false,
                            nullptr, st));
    return DIM_SUCCESS;
}

dim_status dim_csr5_spmm_colmajor(const dim_csr5* A, dim_dtype dtype, uint32_t k, const void* X, uint64_t ldx,
                                  void* Y, uint64_t ldy, void* stream) {
    if (dim_status st = check_handle(A))
        return st;
    if (dim_status st = check_dtype(A, dtype))
        return st;
    if (k < 1 || k > static_cast<uint32_t>(kMaxK))
        return fail(DIM_INVALID_ARGUMENT, "k must be in 1..16");
    if ((A->f.cols && !X) || (A->f.rows && !Y))
        return fail(DIM_INVALID_ARGUMENT, "null matrix argument");
    if (ldx < std::max<uint64_t>(A->f.cols, 1) || ldy < std::max<uint64_t>(A->f.rows, 1))
        return fail(DIM_INVALID_ARGUMENT, "leading dimension smaller than the column length");
    auto* self = const_cast<dim_csr5*>(A);
    DIM_HIP_TRY(launch_spmv(self->f, self->s, dtype == DIM_F32, static_cast<int>(k), X, Y, 0, A->f.tiles, 0, 0,
                            A->f.n_empty, false, nullptr, static_cast<hipStream_t>(stream), nullptr, ldx, ldy));
    return DIM_SUCCESS;
}

dim_status dim_csr5_spmv_transpose(const dim_csr5* A, dim_dtype dtype, const void* x, void* y, void* stream) {
    if (dim_status st = check_handle(A))
        return st;
    if (dim_status st = check_dtype(A, dtype))
        return st;
    if ((A->f.rows && !x) || (A->f.cols && !y))
        return fail(DIM_INVALID_ARGUMENT, "null vector");
    DIM_HIP_TRY(launch_spmv_transpose(A->f, dtype == DIM_F32, x, y, static_cast<hipStream_t>(stream)));
    return DIM_SUCCESS;
}

dim_status dim_csr5_sampled_outer(const dim_csr5* A, double scale, const double* u, const double* v, double* out,
                                  void* stream) {
    if (dim_status st = check_handle(A))
        return st;
    if ((A->f.rows && !u) || (A->f.cols && !v) || (A->f.nnz && !out))
        return fail(DIM_INVALID_ARGUMENT, "null vector");
    for (uint32_t s = 0; s < A->n_systems; ++s) {
        const DevFormat view = system_view(A, s);
        DIM_HIP_TRY(launch_sampled_outer(view, scale, u, v, out + A->sys_nnz_csr[s], static_cast<hipStream_t>(stream)));
    }
    return DIM_SUCCESS;
}

uint32_t dim_csr5_n_systems(const dim_csr5* A) { return A ? A->n_systems : 0; }

dim_status dim_csr5_set_values(dim_csr5* A, const double* vals, void* stream) {
    if (dim_status st = check_handle(A))
        return st;
    if (A->nnz_logical && !vals)
        return fail(DIM_INVALID_ARGUMENT, "vals is null");
    // the column-sliced copy would go stale: products fall back to the
    // unsliced form (hipFree waits for work in flight)
    free_handle(A->xs);
    A->xs = nullptr;
    A->xs_slices = 0;
    for (uint32_t s = 0; s < A->n_systems; ++s) {
        const DevFormat view = system_view(A, s);
        DIM_HIP_TRY(launch_set_values(view, vals + A->sys_nnz_csr[s], static_cast<hipStream_t>(stream)));
    }
    return DIM_SUCCESS;
}

} // extern "C"
