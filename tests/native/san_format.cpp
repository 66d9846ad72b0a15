// Sanitizer driver for the CSR -> CSR5 format builder (diploma_research_amd/csrc/format.cpp).
// Built by tests/test_sanitizers.py with g++ -fsanitize=address,undefined (and
// separately -fsanitize=thread for the threaded builds) and run on the CPU.
// Builds the format at every tile height for CSR matrices with the shapes
// the builder branches on (no elements, 1 x 1, empty rows, a row longer than
// many tiles, rows of one element, wide matrices, stencils that get column
// codes, sizes large enough for the multi-threaded passes) and checks the
// invariants the GPU kernels rely on. Prints "done, N failures".
//
// usage: san_format [small|large|all|threads]
//   threads: only the large matrices at tile heights 4, 7 and 8, where they
//   have >= 4096 tiles (the size from which every pass runs on several
//   threads); for ThreadSanitizer, whose runs are slow.
#include "dim_internal.h"

#include <algorithm>
#include <iterator>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

using namespace dimgpu;

namespace {

int g_fail = 0;

#define EXPECT(cond, ...)                                                                                    \
    do {                                                                                                     \
        if (!(cond)) {                                                                                       \
            std::printf(__VA_ARGS__);                                                                        \
            std::printf("\n");                                                                               \
            ++g_fail;                                                                                        \
        }                                                                                                    \
    } while (0)

struct Csr {
    uint32_t rows = 0, cols = 0;
    std::vector<uint32_t> rp, ci;
    std::vector<double> v;
};

// Random rows of up to max_len sorted distinct columns (a fraction empty), an
// optional long row, or a five-diagonal stencil with offsets +-1, +-stencil.
Csr make(uint32_t rows, uint32_t cols, int max_len, double empty_frac, long long long_row, int stencil,
         unsigned seed) {
    std::mt19937 g(seed);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    Csr m;
    m.rows = rows;
    m.cols = cols;
    m.rp.assign(size_t{rows} + 1, 0);
    for (uint32_t r = 0; r < rows; ++r) {
        if (stencil) {
            for (long long d : {-(long long)stencil, -1LL, 0LL, 1LL, (long long)stencil}) {
                const long long c = (long long)r + d;
                if (c >= 0 && c < (long long)cols) {
                    m.ci.push_back((uint32_t)c);
                    m.v.push_back(2.0 + (double)d);
                }
            }
        } else {
            int len = u(g) < empty_frac ? 0 : (int)(g() % (uint32_t)(max_len + 1));
            if ((long long)r == long_row)
                len = (int)std::min<uint32_t>(cols, 5000);
            std::vector<uint32_t> cs;
            for (int k = 0; k < len; ++k)
                cs.push_back(g() % cols);
            std::sort(cs.begin(), cs.end());
            cs.erase(std::unique(cs.begin(), cs.end()), cs.end());
            for (uint32_t c : cs) {
                m.ci.push_back(c);
                m.v.push_back((double)(g() % 9) - 4.0);
            }
        }
        m.rp[r + 1] = (uint32_t)m.ci.size();
    }
    return m;
}

constexpr uint32_t kAllSigmas[] = {4, 5, 6, 7, 8, 10, 12, 14, 16, 20, 24, 28, 32};

void check(const char* name, const Csr& m, std::vector<uint32_t> sigmas = {}) {
    if (sigmas.empty())
        sigmas.assign(std::begin(kAllSigmas), std::end(kAllSigmas));
    uint64_t nonempty = 0;
    double vsum = 0.0;
    uint64_t csum = 0;
    for (uint32_t r = 0; r < m.rows; ++r)
        nonempty += m.rp[r + 1] > m.rp[r];
    for (size_t k = 0; k < m.v.size(); ++k) {
        vsum += m.v[k];
        csum += m.ci[k];
    }
    for (uint32_t sigma : sigmas) {
        HostFormat f;
        const std::string err =
            build_format(m.rows, m.cols, m.ci.size(), m.rp.data(), m.ci.data(), m.v.data(), sigma, f);
        if (m.ci.empty() || m.rows == 0) {
            if (!err.empty())
                continue; // refusing an empty matrix is allowed; it must not touch memory
        }
        EXPECT(err.empty(), "%s sigma %u: %s", name, sigma, err.c_str());
        if (!err.empty())
            continue;
        const uint64_t per_tile = 64ull * sigma;
        EXPECT(f.tiles == (m.ci.size() + per_tile - 1) / per_tile, "%s sigma %u: %u tiles", name, sigma, f.tiles);
        EXPECT(f.tile_ptr.size() == size_t{f.tiles} + 1, "%s sigma %u: tile_ptr size", name, sigma);
        EXPECT(f.tile_ptr.empty() || f.tile_ptr.back() == m.rows, "%s sigma %u: tile_ptr[tiles]", name, sigma);
        for (uint32_t t = 0; t + 1 < f.tile_ptr.size(); ++t)
            EXPECT((f.tile_ptr[t] & 0x7fffffffu) <= (f.tile_ptr[t + 1] & 0x7fffffffu), "%s sigma %u: tile_ptr %u",
                   name, sigma, t);
        EXPECT(f.vals.size() == f.tiles * per_tile && f.col_idx.size() == f.vals.size(), "%s sigma %u: sizes",
               name, sigma);
        double fv = 0.0;
        uint64_t fc = 0;
        for (size_t k = 0; k < f.vals.size(); ++k) {
            fv += f.vals[k];
            fc += f.col_idx[k];
            EXPECT(f.col_idx[k] < std::max(m.cols, 1u), "%s sigma %u: column %u", name, sigma, f.col_idx[k]);
        }
        // padding is (val 0, col 0), so the sums of a permutation agree exactly
        // (small integers in doubles)
        EXPECT(fv == vsum && fc == csum, "%s sigma %u: elements are not a permutation", name, sigma);
        uint64_t starts = 0;
        for (uint32_t w : f.bit_flag)
            starts += (uint64_t)__builtin_popcount(w);
        EXPECT(starts == nonempty, "%s sigma %u: %llu row starts for %llu non-empty rows", name, sigma,
               (unsigned long long)starts, (unsigned long long)nonempty);
        EXPECT(f.empty_rows.size() == m.rows - nonempty, "%s sigma %u: empty_rows", name, sigma);
        EXPECT(f.col_code.empty() || f.col_code.size() == size_t{f.tiles} * 64, "%s sigma %u: col_code size",
               name, sigma);
        if (!f.col_code.empty())
            EXPECT(sigma <= kMaxCodedSigma && f.col_dict.size() <= kMaxDiagonals, "%s sigma %u: col codes", name,
                   sigma);
    }
}

} // namespace

int main(int argc, char** argv) {
    const char* mode = argc > 1 ? argv[1] : "small";
    const bool threads = std::strcmp(mode, "threads") == 0;
    const bool small = !threads && std::strcmp(mode, "large") != 0;
    const bool large = threads || std::strcmp(mode, "small") != 0;
    const std::vector<uint32_t> sigmas = threads ? std::vector<uint32_t>{4, 7, 8} : std::vector<uint32_t>{};
    if (small) {
        check("1x1", make(1, 1, 1, 0.0, -1, 0, 1));
        check("no elements", make(1, 1, 0, 1.0, -1, 0, 2));
        check("0x0", make(0, 0, 0, 0.0, -1, 0, 3));
        check("empty rows", make(100, 90, 3, 0.5, -1, 0, 4));
        check("long row", make(5000, 4000, 12, 0.1, 7, 0, 5));
        check("stencil", make(20000, 20000, 0, 0.0, -1, 100, 6));
        check("wide", make(777, 20000, 40, 0.0, 3, 0, 7));
        check("short rows", make(3000, 3000, 1, 0.0, -1, 0, 8));
    }
    if (large) {
        // about 3M elements: >= 4096 tiles up to tile height 8, so the
        // multi-threaded passes run there
        check("threaded stencil", make(600000, 600000, 0, 0.0, -1, 700, 9), sigmas);
        check("threaded random", make(400000, 400000, 16, 0.05, 11, 0, 10), sigmas);
    }
    std::printf("done, %d failures\n", g_fail);
    return g_fail != 0;
}
