// HDF5 matrix and vector files (include/dim_h5.h), host code linked against
// the HDF5 C library. Group layouts follow the reference's writers
// (libs/dim/io/h5.cpp:45-74, include/dim/io/h5.h:68-97) so that files pass
// between the two tool chains unchanged.
#include "../../include/dim_h5.h"

#include <hdf5.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

thread_local std::string g_error;

dim_status fail(dim_status s, const std::string& msg) {
    g_error = msg;
    return s;
}

// HDF5 handles closed on scope exit
struct Id {
    hid_t id = -1;
    herr_t (*close)(hid_t) = nullptr;
    Id() = default;
    Id(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    Id(const Id&) = delete;
    Id& operator=(const Id&) = delete;
    Id(Id&& o) noexcept : id(o.id), close(o.close) { o.id = -1; }
    ~Id() {
        if (id >= 0 && close)
            close(id);
    }
    bool ok() const { return id >= 0; }
};

void quiet_hdf5() {
    static bool done = false;
    if (!done) {
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
        done = true;
    }
}

bool file_exists(const char* path) {
    FILE* f = std::fopen(path, "rb");
    if (!f)
        return false;
    std::fclose(f);
    return true;
}

Id open_file(const char* path, bool write) {
    quiet_hdf5();
    return Id(H5Fopen(path, write ? H5F_ACC_RDWR : H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
}

Id create_or_open(const char* path, bool append) {
    quiet_hdf5();
    if (append && file_exists(path))
        return Id(H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT), H5Fclose);
    return Id(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
}

bool link_exists(hid_t loc, const char* name) {
    // every component of a path must exist for H5Lexists on the full path
    std::string p(name), cur;
    size_t i = 0;
    if (!p.empty() && p[0] == '/') {
        cur = "/";
        i = 1;
    }
    while (i <= p.size()) {
        size_t j = p.find('/', i);
        if (j == std::string::npos)
            j = p.size();
        if (j > i) {
            cur += p.substr(i, j - i);
            if (H5Lexists(loc, cur.c_str(), H5P_DEFAULT) <= 0)
                return false;
            cur += "/";
        }
        i = j + 1;
    }
    return true;
}

// dataset creation properties: chunk only when smaller than the dataset
Id dcpl(const dim_h5_dataset_props* p, hsize_t size) {
    Id pl(H5Pcreate(H5P_DATASET_CREATE), H5Pclose);
    if (p && p->chunk && p->chunk < size) {
        const hsize_t c = p->chunk;
        H5Pset_chunk(pl.id, 1, &c);
        if (p->deflate >= 0)
            H5Pset_deflate(pl.id, static_cast<unsigned>(p->deflate));
    } else if (p && p->deflate >= 0 && size > 0) {
        // deflate needs a chunked layout: one chunk for the whole dataset
        const hsize_t c = size;
        H5Pset_chunk(pl.id, 1, &c);
        H5Pset_deflate(pl.id, static_cast<unsigned>(p->deflate));
    }
    return pl;
}

template<typename T>
struct Types;
template<>
struct Types<double> {
    static hid_t mem() { return H5T_NATIVE_DOUBLE; }
    static hid_t disk() { return H5T_IEEE_F64LE; }
};
template<>
struct Types<uint32_t> {
    static hid_t mem() { return H5T_NATIVE_UINT32; }
    static hid_t disk() { return H5T_STD_U32LE; }
};

template<typename T>
bool write_1d(hid_t loc, const char* name, const T* data, hsize_t n, const dim_h5_dataset_props* props) {
    Id space(H5Screate_simple(1, &n, nullptr), H5Sclose);
    Id pl = dcpl(props, n);
    Id ds(H5Dcreate2(loc, name, Types<T>::disk(), space.id, H5P_DEFAULT, pl.id, H5P_DEFAULT), H5Dclose);
    if (!ds.ok())
        return false;
    return n == 0 || H5Dwrite(ds.id, Types<T>::mem(), H5S_ALL, H5S_ALL, H5P_DEFAULT, data) >= 0;
}

bool write_u32_attr(hid_t loc, const char* name, uint32_t v) {
    Id space(H5Screate(H5S_SCALAR), H5Sclose);
    Id a(H5Acreate2(loc, name, H5T_STD_U32LE, space.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    return a.ok() && H5Awrite(a.id, H5T_NATIVE_UINT32, &v) >= 0;
}

bool read_u32_attr(hid_t loc, const char* name, uint32_t* v) {
    Id a(H5Aopen(loc, name, H5P_DEFAULT), H5Aclose);
    return a.ok() && H5Aread(a.id, H5T_NATIVE_UINT32, v) >= 0;
}

// 1-D dataset length (or -1)
int64_t dataset_len(hid_t loc, const char* name) {
    Id ds(H5Dopen2(loc, name, H5P_DEFAULT), H5Dclose);
    if (!ds.ok())
        return -1;
    Id space(H5Dget_space(ds.id), H5Sclose);
    if (H5Sget_simple_extent_ndims(space.id) != 1)
        return -1;
    hsize_t n = 0;
    H5Sget_simple_extent_dims(space.id, &n, nullptr);
    return static_cast<int64_t>(n);
}

template<typename T>
bool read_1d(hid_t loc, const char* name, std::vector<T>& out) {
    const int64_t n = dataset_len(loc, name);
    if (n < 0)
        return false;
    out.resize(static_cast<size_t>(n));
    if (!n)
        return true;
    Id ds(H5Dopen2(loc, name, H5P_DEFAULT), H5Dclose);
    return ds.ok() && H5Dread(ds.id, Types<T>::mem(), H5S_ALL, H5S_ALL, H5P_DEFAULT, out.data()) >= 0;
}

// Elements [start, start + count) of a 1-D dataset (a hyperslab read). `per`
// memory values make one dataset element (4 for the tile_desc arrays).
template<typename T>
bool read_slab(hid_t loc, const char* name, uint64_t start, uint64_t count, std::vector<T>& out,
               hid_t mem_type = Types<T>::mem(), size_t per = 1) {
    out.assign(count * per, T{});
    if (!count)
        return true;
    Id ds(H5Dopen2(loc, name, H5P_DEFAULT), H5Dclose);
    if (!ds.ok())
        return false;
    Id fs(H5Dget_space(ds.id), H5Sclose);
    hsize_t n = 0;
    if (H5Sget_simple_extent_ndims(fs.id) != 1 || H5Sget_simple_extent_dims(fs.id, &n, nullptr) < 0 ||
        start + count > n)
        return false;
    const hsize_t st = start, ct = count;
    if (H5Sselect_hyperslab(fs.id, H5S_SELECT_SET, &st, nullptr, &ct, nullptr) < 0)
        return false;
    Id ms(H5Screate_simple(1, &ct, nullptr), H5Sclose);
    return H5Dread(ds.id, mem_type, ms.id, fs.id, H5P_DEFAULT, out.data()) >= 0;
}

Id array4_type(hid_t base) {
    const hsize_t four = 4;
    return Id(H5Tarray_create2(base, 1, &four), H5Tclose);
}

// ---- the reference's CSR5 layout (omega = 4, sigma = 16) ------------------
// Built on the host from CSR by the rules of csr5_info_t::from_csr
// (include/dim/mat/storage_formats/csr5.h:323-446) and csr5::transpose
// (:463-493).
constexpr uint32_t kRefOmega = 4, kRefSigma = 16, kRefTile = kRefOmega * kRefSigma;
constexpr uint32_t kDirtyBit = 0x80000000u;

// index of the last entry <= v in the sorted array a[0, n)
uint64_t last_not_above(const uint32_t* a, uint64_t n, uint64_t v) {
    return static_cast<uint64_t>(std::upper_bound(a, a + n, v, [](uint64_t x, uint32_t e) { return x < e; }) - a) - 1;
}

struct RefCsr5 {
    uint64_t tiles = 0;
    std::vector<uint32_t> tile_ptr, tile_desc, tdo_ptr, tdo, col_idx;
    std::vector<double> vals;
};

void build_ref_csr5(const dim_h5_csr& m, RefCsr5& f) {
    const uint64_t nnz = m.nnz, T = nnz / kRefTile;
    const uint64_t nrp = uint64_t{m.rows} + 1;
    f.tiles = T;
    // tile pointers: the row of each tile's first element; bit 31 when an
    // empty row lies inside the tile's row span
    f.tile_ptr.assign(T + 1, 0);
    for (uint64_t t = 0; t <= T; ++t)
        f.tile_ptr[t] = static_cast<uint32_t>(last_not_above(m.row_ptr, nrp, std::min(t * kRefTile, nnz)));
    for (uint64_t t = 0; t < T; ++t) {
        const uint32_t a = f.tile_ptr[t] & ~kDirtyBit, b = f.tile_ptr[t + 1] & ~kDirtyBit;
        for (uint32_t r = a; r < b; ++r)
            if (m.row_ptr[r] == m.row_ptr[r + 1]) {
                f.tile_ptr[t] |= kDirtyBit;
                break;
            }
    }
    // bit flags: the start of every row whose first element lies in the tile
    std::vector<uint32_t> bits(T * kRefOmega, 0), y_off(T * kRefOmega, 0), scan(T * kRefOmega, 0);
    for (uint64_t t = 0; t < T; ++t) {
        const uint32_t a = f.tile_ptr[t] & ~kDirtyBit, b = f.tile_ptr[t + 1] & ~kDirtyBit;
        for (uint64_t r = a; r <= b; ++r) {
            const uint64_t e = m.row_ptr[r];
            if (e / kRefTile != t)
                continue;
            bits[t * kRefOmega + (e / kRefSigma) % kRefOmega] |= 1u << (e % kRefSigma);
        }
    }
    // y offsets (exclusive scan of the columns' start counts, column 0 with
    // its implied start), scansum offsets (columns to the right up to the next
    // column holding a start), and the segment counts of dirty tiles
    f.tdo_ptr.assign(T + 1, 0);
    std::vector<uint32_t> nseg(T, 0);
    for (uint64_t t = 0; t < T; ++t) {
        uint32_t cnt[kRefOmega + 1];
        for (uint32_t c = 0; c < kRefOmega; ++c)
            cnt[c] = static_cast<uint32_t>(__builtin_popcount(bits[t * kRefOmega + c] | (c == 0 ? 1u : 0u)));
        cnt[kRefOmega] = 1; // a start beyond the last column ends every scan
        uint32_t acc = 0;
        for (uint32_t c = 0; c < kRefOmega; ++c) {
            y_off[t * kRefOmega + c] = acc;
            acc += cnt[c];
        }
        for (uint32_t c = 0; c < kRefOmega; ++c) {
            uint32_t s = 0;
            if (cnt[c]) {
                uint32_t k = c + 1;
                while (!cnt[k])
                    ++k;
                s = k - c - 1;
            }
            scan[t * kRefOmega + c] = s;
        }
        nseg[t] = acc;
        const uint32_t a = f.tile_ptr[t] & ~kDirtyBit, b = f.tile_ptr[t + 1] & ~kDirtyBit;
        if (a != b && (f.tile_ptr[t] & kDirtyBit))
            f.tdo_ptr[t] = acc + 1;
    }
    // exclusive scan over T + 1 entries: entry T holds the total
    uint32_t run = 0;
    for (uint64_t t = 0; t <= T; ++t) {
        const uint32_t v = f.tdo_ptr[t];
        f.tdo_ptr[t] = run;
        run += v;
    }
    f.tdo.assign(run, 0);
    // segment -> row (relative to the tile's first row) in dirty tiles
    for (uint64_t t = 0; t < T; ++t) {
        if (!(f.tile_ptr[t] & kDirtyBit))
            continue;
        const uint32_t a = f.tile_ptr[t] & ~kDirtyBit, b = f.tile_ptr[t + 1] & ~kDirtyBit;
        const uint32_t* rows = m.row_ptr + a;
        const uint64_t nr = uint64_t{b} - a + 1;
        for (uint32_t c = 0; c < kRefOmega; ++c) {
            const uint32_t bf = bits[t * kRefOmega + c] | (c == 0 ? 1u : 0u);
            uint32_t y = y_off[t * kRefOmega + c];
            for (uint32_t r = 0; r < kRefSigma; ++r) {
                if (!((bf >> r) & 1u))
                    continue;
                const uint64_t g = t * kRefTile + uint64_t{c} * kRefSigma + r;
                f.tdo[f.tdo_ptr[t] + y] = static_cast<uint32_t>(last_not_above(rows, nr, g));
                ++y;
            }
        }
    }
    // the packed descriptor words
    f.tile_desc.assign(T * kRefOmega, 0);
    for (uint64_t i = 0; i < T * kRefOmega; ++i)
        f.tile_desc[i] = (y_off[i] & 0x7Fu) | ((scan[i] & 0x7u) << 7) | ((bits[i] & 0xFFFFu) << 10);
    // values and columns: every full tile that spans more than one row is
    // transposed (CSR offset c*sigma + r goes to r*omega + c); the tail keeps
    // CSR order
    f.vals.assign(m.vals, m.vals + nnz);
    f.col_idx.assign(m.col_idx, m.col_idx + nnz);
    for (uint64_t t = 0; t < T; ++t) {
        if ((f.tile_ptr[t] & ~kDirtyBit) == (f.tile_ptr[t + 1] & ~kDirtyBit))
            continue;
        const uint64_t base = t * kRefTile;
        for (uint32_t c = 0; c < kRefOmega; ++c)
            for (uint32_t r = 0; r < kRefSigma; ++r) {
                f.vals[base + uint64_t{r} * kRefOmega + c] = m.vals[base + uint64_t{c} * kRefSigma + r];
                f.col_idx[base + uint64_t{r} * kRefOmega + c] = m.col_idx[base + uint64_t{c} * kRefSigma + r];
            }
    }
}

dim_status check_csr(const dim_h5_csr* m) {
    if (!m || !m->row_ptr || (m->nnz && (!m->col_idx || !m->vals)))
        return fail(DIM_INVALID_ARGUMENT, "null matrix arrays");
    if (m->row_ptr[0] != 0 || m->row_ptr[m->rows] != m->nnz)
        return fail(DIM_INVALID_ARGUMENT, "row_ptr must start at 0 and end at nnz");
    for (uint32_t r = 0; r < m->rows; ++r)
        if (m->row_ptr[r + 1] < m->row_ptr[r])
            return fail(DIM_INVALID_ARGUMENT, "row_ptr must be non-decreasing");
    for (uint64_t k = 0; k < m->nnz; ++k)
        if (m->col_idx[k] >= m->cols)
            return fail(DIM_INVALID_ARGUMENT, "column index out of range");
    return DIM_SUCCESS;
}

Id open_or_create_group(hid_t file, const char* group) {
    if (link_exists(file, group))
        return Id(H5Gopen2(file, group, H5P_DEFAULT), H5Gclose);
    Id lcpl(H5Pcreate(H5P_LINK_CREATE), H5Pclose);
    H5Pset_create_intermediate_group(lcpl.id, 1);
    return Id(H5Gcreate2(file, group, lcpl.id, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
}

template<typename T>
T* copy_out(const std::vector<T>& v) {
    T* p = static_cast<T*>(std::malloc(std::max<size_t>(v.size(), 1) * sizeof(T)));
    if (p && !v.empty())
        std::memcpy(p, v.data(), v.size() * sizeof(T));
    return p;
}

} // namespace

extern "C" {

const char* dim_h5_last_error(void) { return g_error.c_str(); }

int dim_h5_is_hdf5(const char* path) {
    static const unsigned char magic[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
    if (!path)
        return -1;
    FILE* f = std::fopen(path, "rb");
    if (!f)
        return -1;
    unsigned char buf[8];
    const size_t got = std::fread(buf, 1, 8, f);
    std::fclose(f);
    return got == 8 && std::memcmp(buf, magic, 8) == 0 ? 1 : 0;
}

dim_status dim_h5_group_kind_of(const char* path, const char* group, dim_h5_group_kind* kind) {
    if (!path || !group || !kind)
        return fail(DIM_INVALID_ARGUMENT, "null argument");
    *kind = DIM_H5_UNKNOWN;
    Id file = open_file(path, false);
    if (!file.ok())
        return fail(DIM_IO_ERROR, std::string("cannot open HDF5 file ") + path);
    if (!link_exists(file.id, group))
        return fail(DIM_IO_ERROR, std::string("no group ") + group + " in " + path);
    Id g(H5Gopen2(file.id, group, H5P_DEFAULT), H5Gclose);
    if (!g.ok())
        return fail(DIM_IO_ERROR, std::string(group) + " is not a group");
    if (H5Lexists(g.id, "tile_desc", H5P_DEFAULT) > 0 && H5Lexists(g.id, "row_ptr", H5P_DEFAULT) > 0)
        *kind = DIM_H5_CSR5;
    else if (H5Lexists(g.id, "jc", H5P_DEFAULT) > 0 && H5Lexists(g.id, "ir", H5P_DEFAULT) > 0)
        *kind = DIM_H5_MATLAB_CSR;
    return DIM_SUCCESS;
}

dim_status dim_h5_read_csr(const char* path, const char* group, dim_h5_csr* out) {
    if (!out)
        return fail(DIM_INVALID_ARGUMENT, "out is null");
    *out = dim_h5_csr{};
    dim_h5_group_kind kind;
    if (dim_status st = dim_h5_group_kind_of(path, group, &kind))
        return st;
    if (kind == DIM_H5_UNKNOWN)
        return fail(DIM_IO_ERROR, std::string("group ") + group + " holds neither a CSR5 nor a MATLAB CSR matrix");
    Id file = open_file(path, false);
    Id g(H5Gopen2(file.id, group, H5P_DEFAULT), H5Gclose);
    std::vector<uint32_t> row_ptr, col_idx;
    std::vector<double> vals;
    uint32_t cols = 0;
    if (kind == DIM_H5_MATLAB_CSR) {
        if (!read_1d(g.id, "jc", row_ptr) || !read_1d(g.id, "ir", col_idx) || !read_1d(g.id, "data", vals) ||
            !read_u32_attr(g.id, "MATLAB_sparse", &cols))
            return fail(DIM_IO_ERROR, "malformed MATLAB CSR group");
    } else {
        std::vector<uint32_t> tile_ptr;
        if (!read_1d(g.id, "row_ptr", row_ptr) || !read_1d(g.id, "col_idx", col_idx) ||
            !read_1d(g.id, "vals", vals) || !read_1d(g.id, "tile_ptr", tile_ptr) ||
            !read_u32_attr(g.id, "column_count", &cols))
            return fail(DIM_IO_ERROR, "malformed CSR5 group");
        const int64_t T = dataset_len(g.id, "tile_desc");
        if (T < 0 || static_cast<uint64_t>(T) + 1 != tile_ptr.size() ||
            static_cast<uint64_t>(T) * kRefTile > vals.size() || col_idx.size() != vals.size())
            return fail(DIM_IO_ERROR, "CSR5 group: tile arrays do not match the value array");
        // undo the per-tile transpose (fast-track tiles were stored as is)
        std::vector<double> tv(kRefTile);
        std::vector<uint32_t> tc(kRefTile);
        for (int64_t t = 0; t < T; ++t) {
            if ((tile_ptr[t] & ~kDirtyBit) == (tile_ptr[t + 1] & ~kDirtyBit))
                continue;
            const size_t base = static_cast<size_t>(t) * kRefTile;
            for (uint32_t c = 0; c < kRefOmega; ++c)
                for (uint32_t r = 0; r < kRefSigma; ++r) {
                    tv[c * kRefSigma + r] = vals[base + r * kRefOmega + c];
                    tc[c * kRefSigma + r] = col_idx[base + r * kRefOmega + c];
                }
            std::copy(tv.begin(), tv.end(), vals.begin() + base);
            std::copy(tc.begin(), tc.end(), col_idx.begin() + base);
        }
    }
    if (row_ptr.empty() || row_ptr.back() != vals.size() || col_idx.size() != vals.size())
        return fail(DIM_IO_ERROR, "row_ptr does not match the number of stored elements");
    dim_h5_csr m{};
    m.rows = static_cast<uint32_t>(row_ptr.size() - 1);
    m.cols = cols;
    m.nnz = vals.size();
    m.row_ptr = copy_out(row_ptr);
    m.col_idx = copy_out(col_idx);
    m.vals = copy_out(vals);
    if (!m.row_ptr || !m.col_idx || !m.vals) {
        dim_h5_free_csr(&m);
        return fail(DIM_OUT_OF_MEMORY, "host allocation");
    }
    *out = m;
    return DIM_SUCCESS;
}

void dim_h5_free_csr(dim_h5_csr* m) {
    if (!m)
        return;
    std::free(m->row_ptr);
    std::free(m->col_idx);
    std::free(m->vals);
    *m = dim_h5_csr{};
}

// load_csr5_partial (libs/dim/io/h5.cpp:137-173): part part_idx of part_count
// of the group's tiles, read by hyperslabs. The partition and the row range
// follow the reference: ceil(T / part_count) tiles per part
// (calculate_partition, h5.cpp:95-103), the first row from the part's first
// tile_ptr, the last row from its last tile's last descriptor column
// (csr5_info_t::last_row_idx, csr5.h:304-320) or the matrix's last row for the
// last part, which also takes the tail elements after the last tile. The part
// is returned as CSR over rows [first_row, last_row] holding exactly the
// elements [val_offset, val_end) in CSR order (its full tiles un-transposed),
// so that its product is the part's share of y.
dim_status dim_h5_load_csr5_partial(const char* path, const char* group, uint32_t part_idx, uint32_t part_count,
                                    dim_h5_csr5_part* out) {
    if (!out)
        return fail(DIM_INVALID_ARGUMENT, "out is null");
    *out = dim_h5_csr5_part{};
    if (part_count == 0 || part_idx >= part_count)
        return fail(DIM_INVALID_ARGUMENT, "part_idx must be below part_count");
    dim_h5_group_kind kind;
    if (dim_status st = dim_h5_group_kind_of(path, group, &kind))
        return st;
    if (kind != DIM_H5_CSR5)
        return fail(DIM_INVALID_ARGUMENT, std::string("group ") + group + " holds no CSR5 matrix");
    Id file = open_file(path, false);
    Id g(H5Gopen2(file.id, group, H5P_DEFAULT), H5Gclose);
    uint32_t cols = 0;
    const int64_t T = dataset_len(g.id, "tile_desc"), nv = dataset_len(g.id, "vals"),
                  nrp = dataset_len(g.id, "row_ptr");
    if (T < 0 || nv < 0 || nrp < 1 || dataset_len(g.id, "col_idx") != nv || !read_u32_attr(g.id, "column_count", &cols) ||
        static_cast<uint64_t>(T) * kRefTile > static_cast<uint64_t>(nv))
        return fail(DIM_IO_ERROR, "malformed CSR5 group");
    const uint32_t rows = static_cast<uint32_t>(nrp - 1);
    const uint64_t part = (static_cast<uint64_t>(T) + part_count - 1) / part_count;
    const uint64_t t0 = part * part_idx;
    const bool last_part = part_idx + 1 == part_count;
    // The reference's count = min(part, T - t0) wraps around past the last
    // tile, and a non-last part without tiles has no last row: both refused.
    if (t0 > static_cast<uint64_t>(T) || (!last_part && t0 == static_cast<uint64_t>(T)))
        return fail(DIM_INVALID_ARGUMENT, "part " + std::to_string(part_idx) + " of " + std::to_string(part_count) +
                                            " starts beyond the last tile: use fewer parts");
    const uint64_t nt = std::min<uint64_t>(part, static_cast<uint64_t>(T) - t0);
    std::vector<uint32_t> tile_ptr;
    if (!read_slab(g.id, "tile_ptr", t0, nt + 1, tile_ptr))
        return fail(DIM_IO_ERROR, "cannot read the part's tile_ptr");
    const uint32_t first_row = tile_ptr[0] & ~kDirtyBit;
    uint32_t last_row = rows ? rows - 1 : 0;
    if (!last_part) {
        // the segment of the last tile's last element: its last column's
        // y_offset plus that column's row starts, less one
        std::vector<uint32_t> desc, tdo_ptr, tdo;
        Id a4 = array4_type(H5T_NATIVE_UINT32);
        if (!read_slab(g.id, "tile_desc", t0 + nt - 1, 1, desc, a4.id, 4))
            return fail(DIM_IO_ERROR, "cannot read the part's last tile descriptor");
        const uint32_t col = desc[kRefOmega - 1];
        const uint32_t y_off = col & 0x7fu, bits = (col >> 10) & 0xffffu;
        const uint32_t rel = y_off + static_cast<uint32_t>(__builtin_popcount(bits)) - 1u;
        const uint32_t tp = tile_ptr[nt - 1];
        if (!(tp & kDirtyBit)) {
            last_row = (tp & ~kDirtyBit) + rel;
        } else {
            if (!read_slab(g.id, "tile_desc_offset_ptr", t0 + nt - 1, 1, tdo_ptr) ||
                !read_slab(g.id, "tile_desc_offset", uint64_t{tdo_ptr[0]} + rel, 1, tdo))
                return fail(DIM_IO_ERROR, "cannot read the part's last tile_desc_offset");
            last_row = (tp & ~kDirtyBit) + tdo[0];
        }
    }
    if (rows == 0 || last_row < first_row || last_row >= rows)
        return fail(DIM_IO_ERROR, "CSR5 group: tile descriptors do not match row_ptr");
    const uint64_t v0 = t0 * kRefTile;
    const uint64_t v1 = last_part ? static_cast<uint64_t>(nv) : v0 + nt * kRefTile;
    const uint32_t n_rows = last_row - first_row + 1;
    std::vector<uint32_t> rp, ci;
    std::vector<double> vals;
    // row_ptr entries first_row .. last_row (.. rows for the last part)
    if (!read_slab(g.id, "row_ptr", first_row, last_part ? uint64_t{rows} + 1 - first_row : n_rows, rp) ||
        !read_slab(g.id, "vals", v0, v1 - v0, vals) || !read_slab(g.id, "col_idx", v0, v1 - v0, ci))
        return fail(DIM_IO_ERROR, "cannot read the part's row_ptr, vals or col_idx");
    // un-transpose the part's full tiles (fast-track tiles, inside one row,
    // were stored as they are, csr5.h:463-493)
    std::vector<double> tv(kRefTile);
    std::vector<uint32_t> tc(kRefTile);
    for (uint64_t t = 0; t < nt; ++t) {
        if ((tile_ptr[t] & ~kDirtyBit) == (tile_ptr[t + 1] & ~kDirtyBit))
            continue;
        const size_t base = static_cast<size_t>(t) * kRefTile;
        for (uint32_t c = 0; c < kRefOmega; ++c)
            for (uint32_t r = 0; r < kRefSigma; ++r) {
                tv[c * kRefSigma + r] = vals[base + r * kRefOmega + c];
                tc[c * kRefSigma + r] = ci[base + r * kRefOmega + c];
            }
        std::copy(tv.begin(), tv.end(), vals.begin() + base);
        std::copy(tc.begin(), tc.end(), ci.begin() + base);
    }
    for (uint32_t c : ci)
        if (c >= cols)
            return fail(DIM_IO_ERROR, "CSR5 group: column index out of range");
    // block row i holds the part's elements of row first_row + i
    const uint64_t ne = v1 - v0;
    std::vector<uint32_t> brp(size_t{n_rows} + 1);
    for (uint32_t i = 0; i < n_rows; ++i) {
        const uint64_t e = std::min<uint64_t>(std::max<uint64_t>(rp[i], v0), v1);
        brp[i] = static_cast<uint32_t>(e - v0);
    }
    brp[n_rows] = static_cast<uint32_t>(ne);
    for (uint32_t i = 0; i < n_rows; ++i)
        if (brp[i + 1] < brp[i])
            return fail(DIM_IO_ERROR, "CSR5 group: row_ptr is not non-decreasing");
    dim_h5_csr5_part p{};
    p.rows = rows;
    p.cols = cols;
    p.first_row = first_row;
    p.last_row = last_row;
    p.tile_begin = t0;
    p.tile_count = nt;
    p.val_offset = v0;
    p.skip_tail = last_part ? 0 : 1;
    // the part's first element continues a row that began in an earlier part
    p.first_tile_uncapped = !rp.empty() && rp[0] != v0 ? 1 : 0;
    p.block.rows = n_rows;
    p.block.cols = cols;
    p.block.nnz = ne;
    p.block.row_ptr = copy_out(brp);
    p.block.col_idx = copy_out(ci);
    p.block.vals = copy_out(vals);
    if (!p.block.row_ptr || !p.block.col_idx || !p.block.vals) {
        dim_h5_free_csr(&p.block);
        return fail(DIM_OUT_OF_MEMORY, "host allocation");
    }
    *out = p;
    return DIM_SUCCESS;
}

void dim_h5_free_csr5_part(dim_h5_csr5_part* p) {
    if (p)
        dim_h5_free_csr(&p->block);
}

dim_status dim_csr5_load_h5(const char* path, const char* group, const dim_csr5_opts* opts, dim_csr5** out) {
    if (!out)
        return fail(DIM_INVALID_ARGUMENT, "out is null");
    *out = nullptr;
    dim_h5_csr m{};
    if (dim_status st = dim_h5_read_csr(path, group, &m))
        return st;
    const dim_status st = dim_csr5_from_csr(m.rows, m.cols, m.nnz, m.row_ptr, m.col_idx, m.vals, opts, out);
    dim_h5_free_csr(&m);
    return st ? fail(st, dim_last_error()) : DIM_SUCCESS;
}

dim_status dim_csr5_load_h5_partial(const char* path, const char* group, uint32_t part_idx, uint32_t part_count,
                                    const dim_csr5_opts* opts, dim_csr5** out, uint32_t* first_row,
                                    uint32_t* n_rows) {
    if (!out)
        return fail(DIM_INVALID_ARGUMENT, "out is null");
    *out = nullptr;
    dim_h5_csr5_part part{};
    if (dim_status st = dim_h5_load_csr5_partial(path, group, part_idx, part_count, &part))
        return st;
    const dim_h5_csr& b = part.block;
    const dim_status st = dim_csr5_from_csr(b.rows, b.cols, b.nnz, b.row_ptr, b.col_idx, b.vals, opts, out);
    if (st == DIM_SUCCESS) {
        if (first_row)
            *first_row = part.first_row;
        if (n_rows)
            *n_rows = b.rows;
    }
    dim_h5_free_csr5_part(&part);
    return st ? fail(st, dim_last_error()) : DIM_SUCCESS;
}

dim_status dim_h5_write_matlab_csr(const char* path, const char* group, int append, const dim_h5_csr* m,
                                   const dim_h5_csr_props* props) {
    if (!path || !group)
        return fail(DIM_INVALID_ARGUMENT, "null path or group");
    if (dim_status st = check_csr(m))
        return st;
    Id file = create_or_open(path, append != 0);
    if (!file.ok())
        return fail(DIM_IO_ERROR, std::string("cannot create HDF5 file ") + path);
    if (link_exists(file.id, group))
        return fail(DIM_INVALID_ARGUMENT, std::string("group ") + group + " exists already");
    Id g = open_or_create_group(file.id, group);
    if (!g.ok())
        return fail(DIM_IO_ERROR, std::string("cannot create group ") + group);
    const bool ok = write_1d<double>(g.id, "data", m->vals, m->nnz, props ? &props->values : nullptr) &&
                    write_1d<uint32_t>(g.id, "ir", m->col_idx, m->nnz, props ? &props->col_idx : nullptr) &&
                    write_1d<uint32_t>(g.id, "jc", m->row_ptr, uint64_t{m->rows} + 1,
                                       props ? &props->row_idx : nullptr) &&
                    write_u32_attr(g.id, "MATLAB_sparse", m->cols);
    return ok ? DIM_SUCCESS : fail(DIM_IO_ERROR, "HDF5 write failed");
}

dim_status dim_h5_write_csr5(const char* path, const char* group, int append, const dim_h5_csr* m,
                             const dim_h5_csr5_props* props) {
    if (!path || !group)
        return fail(DIM_INVALID_ARGUMENT, "null path or group");
    if (dim_status st = check_csr(m))
        return st;
    if (m->nnz >= kDirtyBit)
        return fail(DIM_INVALID_ARGUMENT, "too many stored elements for 32-bit tile pointers");
    RefCsr5 f;
    build_ref_csr5(*m, f);
    Id file = create_or_open(path, append != 0);
    if (!file.ok())
        return fail(DIM_IO_ERROR, std::string("cannot create HDF5 file ") + path);
    if (link_exists(file.id, group))
        return fail(DIM_INVALID_ARGUMENT, std::string("group ") + group + " exists already");
    Id g = open_or_create_group(file.id, group);
    if (!g.ok())
        return fail(DIM_IO_ERROR, std::string("cannot create group ") + group);
    bool ok = write_u32_attr(g.id, "column_count", m->cols) &&
              write_1d<double>(g.id, "vals", f.vals.data(), f.vals.size(), props ? &props->vals : nullptr) &&
              write_1d<uint32_t>(g.id, "col_idx", f.col_idx.data(), f.col_idx.size(),
                                 props ? &props->col_idx : nullptr) &&
              write_1d<uint32_t>(g.id, "row_ptr", m->row_ptr, uint64_t{m->rows} + 1,
                                 props ? &props->row_ptr : nullptr) &&
              write_1d<uint32_t>(g.id, "tile_ptr", f.tile_ptr.data(), f.tile_ptr.size(),
                                 props ? &props->tile_ptr : nullptr);
    if (ok) {
        // tile_desc: one H5T array of four u32 words per tile
        const hsize_t T = f.tiles;
        Id disk = array4_type(H5T_STD_U32LE), mem = array4_type(H5T_NATIVE_UINT32);
        Id space(H5Screate_simple(1, &T, nullptr), H5Sclose);
        Id pl = dcpl(props ? &props->tile_desc : nullptr, T);
        Id ds(H5Dcreate2(g.id, "tile_desc", disk.id, space.id, H5P_DEFAULT, pl.id, H5P_DEFAULT), H5Dclose);
        ok = ds.ok() && (T == 0 || H5Dwrite(ds.id, mem.id, H5S_ALL, H5S_ALL, H5P_DEFAULT, f.tile_desc.data()) >= 0);
    }
    ok = ok &&
         write_1d<uint32_t>(g.id, "tile_desc_offset_ptr", f.tdo_ptr.data(), f.tdo_ptr.size(),
                            props ? &props->tile_desc_offset_ptr : nullptr) &&
         write_1d<uint32_t>(g.id, "tile_desc_offset", f.tdo.data(), f.tdo.size(),
                            props ? &props->tile_desc_offset : nullptr);
    return ok ? DIM_SUCCESS : fail(DIM_IO_ERROR, "HDF5 write failed");
}

dim_status dim_h5_write_vector(const char* path, const char* name, const double* v, uint64_t n, int overwrite) {
    if (!path || !name || (n && !v))
        return fail(DIM_INVALID_ARGUMENT, "null argument");
    Id file = create_or_open(path, true);
    if (!file.ok())
        return fail(DIM_IO_ERROR, std::string("cannot open or create ") + path);
    if (link_exists(file.id, name)) {
        if (!overwrite)
            return fail(DIM_IO_ERROR, std::string("file already holds ") + name +
                                                  "; pass overwrite to replace it");
        if (H5Ldelete(file.id, name, H5P_DEFAULT) < 0)
            return fail(DIM_IO_ERROR, std::string("cannot remove ") + name);
    }
    Id lcpl(H5Pcreate(H5P_LINK_CREATE), H5Pclose);
    H5Pset_create_intermediate_group(lcpl.id, 1);
    const hsize_t dims = n;
    Id space(H5Screate_simple(1, &dims, nullptr), H5Sclose);
    Id ds(H5Dcreate2(file.id, name, H5T_IEEE_F64LE, space.id, lcpl.id, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    if (!ds.ok() || (n && H5Dwrite(ds.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, v) < 0))
        return fail(DIM_IO_ERROR, std::string("cannot write ") + name);
    return DIM_SUCCESS;
}

dim_status dim_h5_read_vector(const char* path, const char* name, double** data, uint64_t* n) {
    if (!path || !name || !data || !n)
        return fail(DIM_INVALID_ARGUMENT, "null argument");
    *data = nullptr;
    *n = 0;
    Id file = open_file(path, false);
    if (!file.ok())
        return fail(DIM_IO_ERROR, std::string("cannot open HDF5 file ") + path);
    if (!link_exists(file.id, name))
        return fail(DIM_IO_ERROR, std::string("no dataset ") + name + " in " + path);
    std::vector<double> v;
    if (!read_1d(file.id, name, v))
        return fail(DIM_IO_ERROR, std::string(name) + " is not a 1-D FP64 dataset");
    *data = copy_out(v);
    if (!*data)
        return fail(DIM_OUT_OF_MEMORY, "host allocation");
    *n = v.size();
    return DIM_SUCCESS;
}

void dim_h5_free(void* p) { std::free(p); }

} // extern "C"
