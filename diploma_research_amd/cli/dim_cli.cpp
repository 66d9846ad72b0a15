// dim - command-line tool of the MI355X CSR5 path.
//
// Subcommands mirror the reference's tools:
//   store-matrix     apps/dim_cli/subcommands/store_matrix.cpp (MatrixMarket or
//                    MATLAB-CSR HDF5 in; MATLAB CSR or CSR5 group out, with the
//                    chunk/deflate properties of config.yaml)
//   compare-results  apps/dim_cli/subcommands/compare_results.cpp (tolerance
//                    configurable; the reference fixes 1% per element)
//   spmv             benchmarks/csr5/dim/main.cpp (csr5_single_node: load a CSR5
//                    group, x = 1 unless given, time the product, write Y)
//   cg               benchmarks/conjugate_gradient/dim_sp/main.cpp (b = 1 unless
//                    given, x0 = 0; stops at --rtol or after --num-iters)
//   info             sizes and tile statistics of a matrix group
// The products and solves run on the GPU through libdim_gpu; files go through
// libdim_h5.
#include "dim/gpu/csr5.hpp"
#include "dim/gpu/h5.hpp"
#include "dim_h5.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <numeric>
#include <optional>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

[[noreturn]] void die(const std::string& msg) {
    std::fprintf(stderr, "dim: error: %s\n", msg.c_str());
    std::exit(1);
}

void info(const std::string& msg) { std::printf("[dim] %s\n", msg.c_str()); }

void h5check(dim_status st) {
    if (st != DIM_SUCCESS)
        die(std::string(dim_status_string(st)) + ": " + dim_h5_last_error());
}

#define HIPCHECK(x)                                                                                                   \
    do {                                                                                                              \
        hipError_t e_ = (x);                                                                                          \
        if (e_ != hipSuccess)                                                                                         \
            die(std::string("HIP: ") + hipGetErrorString(e_) + " in " #x);                                           \
    } while (0)

// ---- arguments -------------------------------------------------------------

struct Args {
    std::vector<std::string> pos;
    std::map<std::string, std::string> opt;
    std::map<std::string, bool> flag;

    // flags take no value; every other --name (or -x alias) takes one
    Args(int argc, char** argv, int first, const std::map<std::string, std::string>& alias,
         const std::vector<std::string>& flags) {
        for (int i = first; i < argc; ++i) {
            std::string a = argv[i];
            auto al = alias.find(a);
            if (al != alias.end())
                a = al->second;
            if (a.rfind("--", 0) == 0) {
                std::string name = a.substr(2), val;
                const size_t eq = name.find('=');
                if (eq != std::string::npos) {
                    val = name.substr(eq + 1);
                    name = name.substr(0, eq);
                    opt[name] = val;
                    continue;
                }
                if (std::find(flags.begin(), flags.end(), name) != flags.end()) {
                    flag[name] = true;
                    continue;
                }
                if (i + 1 >= argc)
                    die("option --" + name + " needs a value");
                opt[name] = argv[++i];
            } else {
                pos.push_back(a);
            }
        }
    }
    std::string get(const std::string& k, const std::string& dflt) const {
        auto it = opt.find(k);
        return it == opt.end() ? dflt : it->second;
    }
    std::optional<std::string> maybe(const std::string& k) const {
        auto it = opt.find(k);
        if (it == opt.end())
            return std::nullopt;
        return it->second;
    }
    bool has(const std::string& k) const { return flag.count(k) != 0; }
};

// ---- MatrixMarket (include/dim/io/matrix_market.h:25-83) --------------------
// Coordinate files only; 1-based indices shifted to 0-based; pattern entries
// read as 1; symmetric and hermitian files mirror every off-diagonal entry
// (coo::iterate, include/dim/mat/storage_formats/coo.h:42-44); CSR rows are
// then sorted by column (csr::from_coo, csr.h:42-77).
dim::gpu::csr read_matrix_market(const std::string& path) {
    std::ifstream in(path);
    if (!in)
        die("cannot open " + path);
    std::string line;
    std::getline(in, line);
    std::istringstream banner(line);
    std::string mm, object, format, field, symmetry;
    banner >> mm >> object >> format >> field >> symmetry;
    auto lower = [](std::string s) {
        for (auto& c : s)
            c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
        return s;
    };
    if (mm != "%%MatrixMarket" || lower(object) != "matrix")
        die(path + " is not a MatrixMarket file");
    if (lower(format) != "coordinate")
        throw std::invalid_argument{"Only coordinate matrix loading is implemented."};
    const bool pattern = lower(field) == "pattern";
    const bool symmetric = lower(symmetry) == "symmetric" || lower(symmetry) == "hermitian";
    while (std::getline(in, line))
        if (!line.empty() && line[0] != '%')
            break;
    uint64_t rows = 0, cols = 0, nnz = 0;
    std::istringstream(line) >> rows >> cols >> nnz;
    std::vector<uint32_t> r(nnz), c(nnz);
    std::vector<double> v(nnz, 1.0);
    for (uint64_t k = 0; k < nnz; ++k) {
        uint64_t ri, ci;
        if (!(in >> ri >> ci))
            die("scan failed reading " + path);
        if (!pattern && !(in >> v[k]))
            die("scan failed reading " + path);
        if (ri < 1 || ci < 1 || ri > rows || ci > cols)
            die("entry out of range in " + path);
        r[k] = static_cast<uint32_t>(ri - 1);
        c[k] = static_cast<uint32_t>(ci - 1);
    }
    dim::gpu::csr m;
    m.dimensions = {static_cast<uint32_t>(rows), static_cast<uint32_t>(cols)};
    std::vector<uint32_t> count(rows + 1, 0);
    auto each = [&](auto&& fn) {
        for (uint64_t k = 0; k < nnz; ++k) {
            fn(v[k], r[k], c[k]);
            if (symmetric && r[k] != c[k])
                fn(v[k], c[k], r[k]);
        }
    };
    each([&](double, uint32_t row, uint32_t) { ++count[row]; });
    m.row_start_offsets.assign(rows + 1, 0);
    std::exclusive_scan(count.begin(), count.end(), m.row_start_offsets.begin(), 0u);
    const uint64_t total = m.row_start_offsets[rows];
    m.col_indices.resize(total);
    m.values.resize(total);
    std::fill(count.begin(), count.end(), 0);
    each([&](double val, uint32_t row, uint32_t col) {
        const uint64_t at = m.row_start_offsets[row] + count[row]++;
        m.col_indices[at] = col;
        m.values[at] = val;
    });
    std::vector<std::pair<uint32_t, double>> tmp;
    for (uint64_t row = 0; row < rows; ++row) {
        const uint32_t a = m.row_start_offsets[row], b = m.row_start_offsets[row + 1];
        tmp.clear();
        for (uint32_t k = a; k < b; ++k)
            tmp.emplace_back(m.col_indices[k], m.values[k]);
        std::stable_sort(tmp.begin(), tmp.end(), [](auto& x, auto& y) { return x.first < y.first; });
        for (uint32_t k = a; k < b; ++k) {
            m.col_indices[k] = tmp[k - a].first;
            m.values[k] = tmp[k - a].second;
        }
    }
    return m;
}

// ---- config.yaml (resources/config.yaml): the two-level map of
// csr / csr5 -> dataset -> {chunk_size, compression_level} that store-matrix
// reads (store_matrix.cpp:23-73). Only that shape is understood.
struct StoreConfig {
    dim_h5_csr_props csr{};
    dim_h5_csr5_props csr5{};
};

StoreConfig read_config(const std::optional<std::string>& path) {
    StoreConfig cfg;
    auto none = [](dim_h5_dataset_props& p) {
        p.chunk = 0;
        p.deflate = -1;
    };
    for (auto* p : {&cfg.csr.values, &cfg.csr.col_idx, &cfg.csr.row_idx})
        none(*p);
    for (auto* p : {&cfg.csr5.vals, &cfg.csr5.col_idx, &cfg.csr5.row_ptr, &cfg.csr5.tile_ptr, &cfg.csr5.tile_desc,
                    &cfg.csr5.tile_desc_offset, &cfg.csr5.tile_desc_offset_ptr})
        none(*p);
    if (!path)
        return cfg;
    std::ifstream in(*path);
    if (!in)
        die("cannot open config " + *path);
    std::map<std::string, dim_h5_dataset_props*> csr{{"values", &cfg.csr.values},
                                                     {"col_idx", &cfg.csr.col_idx},
                                                     {"row_idx", &cfg.csr.row_idx}};
    std::map<std::string, dim_h5_dataset_props*> csr5{{"vals", &cfg.csr5.vals},
                                                      {"col_idx", &cfg.csr5.col_idx},
                                                      {"row_ptr", &cfg.csr5.row_ptr},
                                                      {"tile_ptr", &cfg.csr5.tile_ptr},
                                                      {"tile_desc", &cfg.csr5.tile_desc},
                                                      {"tile_desc_offset", &cfg.csr5.tile_desc_offset},
                                                      {"tile_desc_offset_ptr", &cfg.csr5.tile_desc_offset_ptr}};
    std::string line, section;
    dim_h5_dataset_props* cur = nullptr;
    while (std::getline(in, line)) {
        const size_t hash = line.find('#');
        if (hash != std::string::npos)
            line = line.substr(0, hash);
        if (line.find_first_not_of(" \t\r") == std::string::npos)
            continue;
        const size_t indent = line.find_first_not_of(' ');
        const size_t colon = line.find(':');
        if (colon == std::string::npos)
            die("config: cannot read line: " + line);
        std::string key = line.substr(indent, colon - indent);
        std::string val = line.substr(colon + 1);
        val.erase(0, val.find_first_not_of(" \t"));
        val.erase(val.find_last_not_of(" \t\r") + 1);
        if (indent == 0) {
            section = key;
            cur = nullptr;
        } else if (indent == 2) {
            auto& table = section == "csr" ? csr : csr5;
            auto it = table.find(key);
            cur = (section == "csr" || section == "csr5") && it != table.end() ? it->second : nullptr;
        } else if (cur) {
            const bool null = val.empty() || val == "null" || val == "~";
            if (key == "chunk_size")
                cur->chunk = null ? 0 : std::stoull(val);
            else if (key == "compression_level")
                cur->deflate = null ? -1 : std::stoi(val);
        }
    }
    return cfg;
}

std::string deduce_output(std::string base, const std::string& format) {
    // strip every extension, then add .<format>.h5 (store_matrix.cpp:83-92)
    const size_t slash = base.find_last_of('/');
    size_t dot;
    while ((dot = base.find_last_of('.')) != std::string::npos && (slash == std::string::npos || dot > slash))
        base = base.substr(0, dot);
    return base + "." + format + ".h5";
}

// ---- subcommands -----------------------------------------------------------

int store_matrix(int argc, char** argv) {
    Args a(argc, argv, 2, {{"-o", "--output"}, {"-g", "--group-name"}, {"-f", "--format"}, {"-c", "--config"}},
           {"append"});
    if (a.pos.size() != 1)
        die("usage: dim store-matrix <input> [-o out] [--in-group-name A] [-g A] [--append] [-f csr|csr5] "
            "[-c config.yaml]");
    const std::string input = a.pos[0];
    const std::string format = a.get("format", "csr");
    if (format != "csr" && format != "csr5")
        die("format must be csr or csr5");
    const std::string out = a.get("output", deduce_output(input, format));
    const std::string group = a.get("group-name", "A");
    const StoreConfig cfg = read_config(a.maybe("config"));
    info("loading matrix from '" + input + "'");
    auto t0 = std::chrono::steady_clock::now();
    dim_h5_csr m{};
    dim::gpu::csr mm;
    const bool from_h5 = dim_h5_is_hdf5(input.c_str()) == 1;
    if (from_h5) {
        h5check(dim_h5_read_csr(input.c_str(), a.get("in-group-name", "A").c_str(), &m));
    } else {
        mm = read_matrix_market(input);
        m.rows = mm.dimensions.rows;
        m.cols = mm.dimensions.cols;
        m.nnz = mm.values.size();
        m.row_ptr = mm.row_start_offsets.data();
        m.col_idx = mm.col_indices.data();
        m.vals = mm.values.data();
    }
    auto t1 = std::chrono::steady_clock::now();
    info("load and conversion to CSR took: " + std::to_string(std::chrono::duration<double>(t1 - t0).count()) + "s");
    info("storing matrix as group '" + group + "' to " + out);
    const dim_status st = format == "csr" ? dim_h5_write_matlab_csr(out.c_str(), group.c_str(), a.has("append"), &m,
                                                                    &cfg.csr)
                                          : dim_h5_write_csr5(out.c_str(), group.c_str(), a.has("append"), &m,
                                                              &cfg.csr5);
    if (from_h5)
        dim_h5_free_csr(&m); // the MatrixMarket arrays belong to mm
    h5check(st);
    auto t2 = std::chrono::steady_clock::now();
    info("storing as HDF5 took: " + std::to_string(std::chrono::duration<double>(t2 - t1).count()) + "s");
    return 0;
}

int compare_results(int argc, char** argv) {
    Args a(argc, argv, 2, {}, {"normwise"});
    if (a.pos.empty() || a.pos.size() > 2)
        die("usage: dim compare-results <file> [file2] [--lhs-group /] [--lhs-dataset Y] [--rhs-group /] "
            "[--rhs-dataset Y] [--rtol 0.01] [--normwise]");
    auto path_of = [](const std::string& group, const std::string& ds) {
        std::string g = group.empty() ? "/" : group;
        if (g.back() != '/')
            g += '/';
        return g + ds;
    };
    const std::string f1 = a.pos[0], f2 = a.pos.size() == 2 ? a.pos[1] : a.pos[0];
    const std::string l = path_of(a.get("lhs-group", "/"), a.get("lhs-dataset", "Y"));
    const std::string r = path_of(a.get("rhs-group", "/"), a.get("rhs-dataset", "Y"));
    const double rtol = std::stod(a.get("rtol", "0.01"));
    double *lhs = nullptr, *rhs = nullptr;
    uint64_t nl = 0, nr = 0;
    h5check(dim_h5_read_vector(f1.c_str(), l.c_str(), &lhs, &nl));
    h5check(dim_h5_read_vector(f2.c_str(), r.c_str(), &rhs, &nr));
    const std::string lid = f1 + ":" + l, rid = f2 + ":" + r;
    bool correct = nl == nr;
    if (nl != nr)
        std::printf("[dim] sizes differ: %llu vs %llu\n", static_cast<unsigned long long>(nl),
                    static_cast<unsigned long long>(nr));
    if (correct && a.has("normwise")) {
        double d = 0, s = 0;
        for (uint64_t i = 0; i < nl; ++i) {
            d += (lhs[i] - rhs[i]) * (lhs[i] - rhs[i]);
            s += rhs[i] * rhs[i];
        }
        const double rel = s > 0 ? std::sqrt(d / s) : std::sqrt(d);
        std::printf("[dim] normwise relative difference %.3e (tolerance %.3e)\n", rel, rtol);
        correct = rel <= rtol;
    } else if (correct) {
        // per element |l - r| <= rtol * |r| (compare_results.cpp:10)
        uint64_t bad = 0;
        for (uint64_t i = 0; i < nl; ++i)
            if (!(std::abs(lhs[i] - rhs[i]) <= rtol * std::abs(rhs[i]))) {
                if (bad++ < 20)
                    std::printf("[dim] at %llu: expected %.17g, got %.17g\n", static_cast<unsigned long long>(i),
                                lhs[i], rhs[i]);
            }
        correct = bad == 0;
    }
    dim_h5_free(lhs);
    dim_h5_free(rhs);
    if (!correct) {
        std::fprintf(stderr, "[dim] vectors %s and %s are not equal\n", lid.c_str(), rid.c_str());
        return 1;
    }
    info("vectors " + lid + " and " + rid + " are equal");
    return 0;
}

dim::gpu::csr5 load_matrix(const std::string& path, const std::string& group, uint32_t sigma, bool keep_f32,
                           dim_h5_csr& m) {
    auto t0 = std::chrono::steady_clock::now();
    h5check(dim_h5_read_csr(path.c_str(), group.c_str(), &m));
    auto t1 = std::chrono::steady_clock::now();
    auto A = dim::gpu::csr5::from_csr(m.rows, m.cols, m.nnz, m.row_ptr, m.col_idx, m.vals, {sigma, keep_f32});
    HIPCHECK(hipDeviceSynchronize());
    auto t2 = std::chrono::steady_clock::now();
    info("CSR5 loading took " + std::to_string(std::chrono::duration<double>(t1 - t0).count()) +
         "s, conversion and upload " + std::to_string(std::chrono::duration<double>(t2 - t1).count()) + "s (sigma " +
         std::to_string(A.sigma()) + ")");
    return A;
}

std::vector<double> vector_or(const std::optional<std::string>& file, const std::optional<std::string>& name,
                              uint64_t n, double fill) {
    if (!name)
        return std::vector<double>(n, fill);
    double* d = nullptr;
    uint64_t len = 0;
    h5check(dim_h5_read_vector(file->c_str(), name->c_str(), &d, &len));
    std::vector<double> v(d, d + len);
    dim_h5_free(d);
    if (len != n)
        die(*name + " has " + std::to_string(len) + " entries, expected " + std::to_string(n));
    return v;
}

// "i/n" -> (i, n)
std::pair<uint32_t, uint32_t> parse_part(const std::string& s) {
    const size_t k = s.find('/');
    if (k == std::string::npos)
        die("--part takes i/n (part i of n, counted from 0)");
    return {static_cast<uint32_t>(std::stoul(s.substr(0, k))), static_cast<uint32_t>(std::stoul(s.substr(k + 1)))};
}

int spmv(int argc, char** argv) {
    Args a(argc, argv, 2, {{"-n", "--num-runs"}, {"-g", "--matrix-group"}}, {"overwrite"});
    if (a.pos.size() != 1)
        die("usage: dim spmv <matrix.h5> [-n 100] [-g A] [--output-file F] [--vector-dataset Y] [--overwrite] "
            "[--x-file F --x-dataset X] [--sigma 0] [--dtype f64|f32] [--part i/n]");
    const std::string input = a.pos[0];
    const bool f32 = a.get("dtype", "f64") == "f32";
    const auto sigma = static_cast<uint32_t>(std::stoul(a.get("sigma", "0")));
    dim_h5_csr m{};
    // --part i/n: one rank's share (h5::load_csr5_partial and spmv<partial>,
    // include/dim/mpi/csr5.h:40-54); y then holds its rows first .. last
    std::optional<dim::gpu::csr5> part;
    if (auto p = a.maybe("part")) {
        const auto [idx, count] = parse_part(*p);
        part.emplace(dim::io::h5::load_csr5_partial(input, a.get("matrix-group", "A"), {idx, count}, {sigma, f32}));
        m.rows = part->last_row_idx() - part->first_row_idx() + 1;
        m.cols = part->dimensions().cols;
        m.nnz = part->non_zero_count();
        info("part " + *p + ": rows " + std::to_string(part->first_row_idx()) + ".." +
             std::to_string(part->last_row_idx()) + ", elements from " + std::to_string(part->val_offset()) +
             " (sigma " + std::to_string(part->sigma()) + ")");
    }
    auto A = part ? std::move(*part) : load_matrix(input, a.get("matrix-group", "A"), sigma, f32, m);
    const uint64_t n_runs = std::stoull(a.get("num-runs", "100"));
    auto xh = vector_or(a.maybe("x-file") ? a.maybe("x-file") : std::optional<std::string>(input), a.maybe("x-dataset"),
                        m.cols, 1.0);
    const size_t esz = f32 ? 4 : 8;
    void *x = nullptr, *y = nullptr;
    HIPCHECK(hipMalloc(&x, std::max<size_t>(m.cols, 1) * esz));
    HIPCHECK(hipMalloc(&y, std::max<size_t>(m.rows, 1) * esz));
    if (f32) {
        std::vector<float> xf(xh.begin(), xh.end());
        HIPCHECK(hipMemcpy(x, xf.data(), xf.size() * 4, hipMemcpyHostToDevice));
    } else {
        HIPCHECK(hipMemcpy(x, xh.data(), xh.size() * 8, hipMemcpyHostToDevice));
    }
    auto calibrator = A.create_calibrator();
    const bool partial = A.first_row_idx() != 0 || A.skip_tail() || a.maybe("part");
    auto run = [&] {
        using S = dim::gpu::spmv_strategy;
        if (f32) {
            dim::gpu::spmv_data_t<float> d{{static_cast<const float*>(x), m.cols}, {static_cast<float*>(y), m.rows},
                                           calibrator};
            partial ? A.spmv<S::partial>(d) : A.spmv<S::absolute>(d);
        } else {
            dim::gpu::spmv_data_t<double> d{{static_cast<const double*>(x), m.cols},
                                            {static_cast<double*>(y), m.rows}, calibrator};
            partial ? A.spmv<S::partial>(d) : A.spmv<S::absolute>(d);
        }
    };
    for (int i = 0; i < 5; ++i) // warm-up (main.cpp:60-62)
        run();
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    HIPCHECK(hipEventRecord(e0, nullptr));
    for (uint64_t i = 0; i < n_runs; ++i)
        run();
    HIPCHECK(hipEventRecord(e1, nullptr));
    HIPCHECK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
    const double avg = n_runs ? ms / 1e3 / n_runs : 0.0;
    const double bytes = static_cast<double>(esz + 4) * m.nnz + static_cast<double>(esz) * (m.rows + m.cols);
    std::printf("[dim] CSR5 SpMV took %.6es on average over %llu runs (%.1f GB/s of algorithmic bytes)\n", avg,
                static_cast<unsigned long long>(n_runs), avg > 0 ? bytes / avg / 1e9 : 0.0);
    if (auto out = a.maybe("output-file")) {
        std::vector<double> yh(m.rows);
        if (f32) {
            std::vector<float> yf(m.rows);
            HIPCHECK(hipMemcpy(yf.data(), y, yf.size() * 4, hipMemcpyDeviceToHost));
            std::copy(yf.begin(), yf.end(), yh.begin());
        } else {
            HIPCHECK(hipMemcpy(yh.data(), y, yh.size() * 8, hipMemcpyDeviceToHost));
        }
        h5check(dim_h5_write_vector(out->c_str(), a.get("vector-dataset", "Y").c_str(), yh.data(), yh.size(),
                                    a.has("overwrite")));
        info("wrote " + a.get("vector-dataset", "Y") + " to " + *out);
    }
    (void)hipFree(x);
    (void)hipFree(y);
    dim_h5_free_csr(&m);
    return 0;
}

int cg(int argc, char** argv) {
    Args a(argc, argv, 2, {{"-g", "--group-name"}}, {"overwrite"});
    if (a.pos.size() != 1)
        die("usage: dim cg <matrix.h5> [-g A] [--num-iters 5000] [--rtol 1e-8] [--precision f64|f32-refined] "
            "[--b-file F --b-dataset b] [--output-file F --x-dataset x --overwrite] [--stats stats.json]");
    const std::string input = a.pos[0];
    const bool mixed = a.get("precision", "f64") == "f32-refined";
    dim_h5_csr m{};
    auto A = load_matrix(input, a.get("group-name", "A"), static_cast<uint32_t>(std::stoul(a.get("sigma", "0"))),
                         mixed, m);
    if (m.rows != m.cols)
        die("CG needs a square matrix");
    // r0 = b - A x0 with x0 = 0 and b = 1 unless given (dim_sp/main.cpp:88-93)
    auto bh = vector_or(a.maybe("b-file") ? a.maybe("b-file") : std::optional<std::string>(input), a.maybe("b-dataset"),
                        m.rows, 1.0);
    double *b = nullptr, *x = nullptr;
    HIPCHECK(hipMalloc(&b, std::max<size_t>(m.rows, 1) * 8));
    HIPCHECK(hipMalloc(&x, std::max<size_t>(m.rows, 1) * 8));
    HIPCHECK(hipMemcpy(b, bh.data(), bh.size() * 8, hipMemcpyHostToDevice));
    HIPCHECK(hipMemset(x, 0, std::max<size_t>(m.rows, 1) * 8));
    dim::gpu::cg_options o;
    o.rtol = std::stod(a.get("rtol", "1e-8"));
    o.max_iters = static_cast<uint32_t>(std::stoul(a.get("num-iters", "5000")));
    o.f32_refined = mixed;
    const auto res = dim::gpu::conjugate_gradient(A, b, x, o);
    std::printf("[dim] CG %s after %u iterations in %.6fs: relative residual %.3e (recomputed %.3e)%s\n",
                res.converged ? "converged" : "stopped", res.iterations, res.seconds, res.relres, res.relres_true,
                mixed ? (" with " + std::to_string(res.refinements) + " FP64 residual replacements").c_str() : "");
    if (auto out = a.maybe("output-file")) {
        std::vector<double> xh(m.rows);
        HIPCHECK(hipMemcpy(xh.data(), x, xh.size() * 8, hipMemcpyDeviceToHost));
        h5check(dim_h5_write_vector(out->c_str(), a.get("x-dataset", "x").c_str(), xh.data(), xh.size(),
                                    a.has("overwrite")));
    }
    // stats.json, as dim_sp writes it (main.cpp:129-137), with the solver's figures
    std::ofstream js(a.get("stats", "stats.json"));
    js << "{\n    \"cg\": {\n        \"num_iters\": " << res.iterations << ",\n        \"total\": " << res.seconds
       << ",\n        \"converged\": " << (res.converged ? "true" : "false") << ",\n        \"relres\": " << res.relres
       << ",\n        \"relres_true\": " << res.relres_true << ",\n        \"refinements\": " << res.refinements
       << "\n    },\n    \"general_info\": {\"rows\": " << m.rows << ", \"nnz\": " << m.nnz
       << ", \"sigma\": " << A.sigma() << ", \"precision\": \"" << (mixed ? "f32-refined" : "f64") << "\"}\n}\n";
    (void)hipFree(b);
    (void)hipFree(x);
    dim_h5_free_csr(&m);
    return res.converged || o.rtol == 0.0 ? 0 : 2;
}

int info_cmd(int argc, char** argv) {
    Args a(argc, argv, 2, {{"-g", "--group-name"}}, {});
    if (a.pos.size() != 1)
        die("usage: dim info <matrix.h5> [-g A]");
    dim_h5_group_kind kind;
    const std::string group = a.get("group-name", "A");
    h5check(dim_h5_group_kind_of(a.pos[0].c_str(), group.c_str(), &kind));
    dim_h5_csr m{};
    h5check(dim_h5_read_csr(a.pos[0].c_str(), group.c_str(), &m));
    uint32_t empty = 0, longest = 0;
    for (uint32_t r = 0; r < m.rows; ++r) {
        const uint32_t len = m.row_ptr[r + 1] - m.row_ptr[r];
        empty += len == 0;
        longest = std::max(longest, len);
    }
    std::printf("group %s: %s, %u x %u, nnz %llu, empty rows %u, longest row %u, mean row %.2f\n", group.c_str(),
                kind == DIM_H5_CSR5 ? "CSR5 (omega 4, sigma 16)" : "MATLAB CSR", m.rows, m.cols,
                static_cast<unsigned long long>(m.nnz), empty, longest,
                m.rows ? static_cast<double>(m.nnz) / m.rows : 0.0);
    dim_h5_free_csr(&m);
    return 0;
}

} // namespace

int main(int argc, char** argv) {
    const std::string usage = "usage: dim <store-matrix|compare-results|spmv|cg|info> ...";
    if (argc < 2)
        die(usage);
    const std::string cmd = argv[1];
    try {
        if (cmd == "store-matrix" || cmd == "store_matrix")
            return store_matrix(argc, argv);
        if (cmd == "compare-results" || cmd == "compare_results")
            return compare_results(argc, argv);
        if (cmd == "spmv")
            return spmv(argc, argv);
        if (cmd == "cg")
            return cg(argc, argv);
        if (cmd == "info")
            return info_cmd(argc, argv);
    } catch (const std::exception& e) {
        die(e.what());
    }
    die(usage);
}
