// Sanitizer driver for the HDF5 matrix I/O (diploma_research_amd/csrc/h5io.cpp).
// Built by tests/test_sanitizers.py with g++ -fsanitize=address,undefined,
// linked against HDF5 alone (the two GPU-library calls the load-to-GPU entry
// points make are stubbed below) and run on the CPU. Reads the committed
// HDF5 fixtures, writes them back in both layouts, reads every partial load
// for 1..13 parts, and does the same for a random matrix with empty rows and
// a row spanning several tiles. Prints the number of partial loads that
// succeeded, then "done, N failures".
//
// usage: san_h5 <golden dir> <scratch dir>
#include "dim_gpu.h"
#include "dim_h5.h"

#include <cstdio>
#include <random>
#include <string>
#include <vector>

// The load-to-GPU entry points of h5io.cpp call these; this driver never
// takes those paths.
extern "C" dim_status dim_csr5_from_csr(uint32_t, uint32_t, uint64_t, const uint32_t*, const uint32_t*,
                                        const double*, const dim_csr5_opts*, dim_csr5**) {
    return DIM_UNSUPPORTED;
}
extern "C" const char* dim_last_error(void) { return "no GPU library in this driver"; }

namespace {

int g_fail = 0;
uint64_t g_parts = 0; // partial loads that succeeded

void expect_ok(dim_status s, const char* what) {
    if (s != DIM_SUCCESS) {
        std::printf("status %d at %s: %s\n", (int)s, what, dim_h5_last_error());
        ++g_fail;
    }
}

// Every part of every split: element counts add up to nnz and parts tile the rows.
void partial_loads(const std::string& path, const char* group, uint32_t max_parts, const dim_h5_csr& whole) {
    for (uint32_t count = 1; count <= max_parts; ++count) {
        uint64_t nnz = 0;
        uint32_t next_row = 0;
        bool ok = true;
        for (uint32_t i = 0; i < count; ++i) {
            dim_h5_csr5_part p{};
            if (dim_h5_load_csr5_partial(path.c_str(), group, i, count, &p) != DIM_SUCCESS) {
                ok = false; // a split into more parts than tiles is refused
                break;
            }
            ++g_parts;
            nnz += p.block.nnz;
            if (p.first_row + p.block.rows < next_row) {
                std::printf("%s/%s part %u of %u goes back to row %u\n", path.c_str(), group, i, count, p.first_row);
                ++g_fail;
            }
            next_row = p.first_row + p.block.rows;
            dim_h5_free_csr5_part(&p);
        }
        if (ok && (nnz < whole.nnz || next_row != whole.rows)) {
            std::printf("%s/%s: %u parts hold %llu elements up to row %u (matrix: %llu, %u rows)\n", path.c_str(),
                        group, count, (unsigned long long)nnz, next_row, (unsigned long long)whole.nnz, whole.rows);
            ++g_fail;
        }
    }
}

} // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::printf("usage: san_h5 <golden dir> <scratch dir>\n");
        return 2;
    }
    const std::string golden = argv[1], scratch = argv[2];
    const std::string rt = scratch + "/san_rt.h5", big = scratch + "/san_big.h5";
    const char* files[][2] = {{"dirty.csr5.h5", "B"}, {"mm.csr5.h5", "A"}, {"mm.csr.h5", "A"}};
    for (auto& f : files) {
        const std::string path = golden + "/" + f[0];
        dim_h5_csr m{};
        expect_ok(dim_h5_read_csr(path.c_str(), f[1], &m), "read fixture");
        partial_loads(path, f[1], 13, m);
        expect_ok(dim_h5_write_csr5(rt.c_str(), "A", 0, &m, nullptr), "write csr5");
        expect_ok(dim_h5_write_matlab_csr(rt.c_str(), "M", 1, &m, nullptr), "write matlab csr");
        dim_h5_csr r{};
        expect_ok(dim_h5_read_csr(rt.c_str(), "A", &r), "read back");
        if (r.nnz != m.nnz || r.rows != m.rows) {
            std::printf("%s: round trip changed the shape\n", f[0]);
            ++g_fail;
        }
        partial_loads(rt, "A", 9, m);
        dim_h5_free_csr(&r);
        dim_h5_free_csr(&m);
    }

    std::mt19937 g(1);
    const uint32_t rows = 3000, cols = 2500;
    std::vector<uint32_t> rp(rows + 1, 0), ci;
    std::vector<double> v;
    for (uint32_t r = 0; r < rows; ++r) {
        const uint32_t len = (r % 97 == 0) ? 0 : (r == 5 ? 2000 : g() % 9);
        uint32_t c = g() % 50;
        for (uint32_t k = 0; k < len && c < cols; ++k, c += 1 + g() % 3) {
            ci.push_back(c);
            v.push_back((double)(g() % 17) - 8.0);
        }
        rp[r + 1] = (uint32_t)ci.size();
    }
    dim_h5_csr m{rows, cols, ci.size(), rp.data(), ci.data(), v.data()};
    expect_ok(dim_h5_write_csr5(big.c_str(), "A", 0, &m, nullptr), "write random");
    partial_loads(big, "A", 13, m);
    std::vector<double> vec(1000, 1.5);
    expect_ok(dim_h5_write_vector(big.c_str(), "Y", vec.data(), vec.size(), 1), "write vector");
    double* back = nullptr;
    uint64_t n = 0;
    expect_ok(dim_h5_read_vector(big.c_str(), "Y", &back, &n), "read vector");
    if (n != vec.size() || (back && back[999] != 1.5)) {
        std::printf("vector round trip\n");
        ++g_fail;
    }
    dim_h5_free(back);
    std::printf("%llu partial loads\n", (unsigned long long)g_parts);
    std::printf("done, %d failures\n", g_fail);
    return g_fail != 0;
}
