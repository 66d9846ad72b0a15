// Thin extern "C" shim over the original CSR5 code that the reference vendors
// (third-party/csr5_orig/include/anonymouslib_avx2.h, Liu & Vinter's AVX2
// anonymouslib, omega = 4), compiled by oracle/Makefile straight from the
// reference tree into oracle/_ref/libcsr5orig.so.
//
// TEST INFRASTRUCTURE ONLY: a second CPU oracle that cross-checks the
// restatement of the reference (its CSR5 descends from this code; the two
// differ in the tile count, where the original keeps a tail partition, and
// in how the first row of a tile is written). It is never part of the
// product.
#include <anonymouslib_avx2.h>

#include <fcntl.h>
#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <iostream>
#include <vector>

namespace {

// The original prints its conversion timings to stdout; keep them out of the
// test logs.
struct QuietStdout {
    int saved = -1;
    QuietStdout() {
        std::fflush(stdout);
        std::cout.flush();
        const int null_fd = open("/dev/null", O_WRONLY);
        if (null_fd >= 0) {
            saved = dup(1);
            dup2(null_fd, 1);
            close(null_fd);
        }
    }
    ~QuietStdout() {
        std::fflush(stdout);
        std::cout.flush();
        if (saved >= 0) {
            dup2(saved, 1);
            close(saved);
        }
    }
};

} // namespace

extern "C" {

// y = A x with the original CSR5 at tile height sigma. The conversion
// transposes its own copies of the column and value arrays; y is zero-filled
// first because the original never writes empty rows.
// Returns the library's status (0 = success).
int csr5orig_spmv(int rows, int cols, int nnz, const int* row_ptr, const int* col_idx, const double* vals,
                  int sigma, const double* x, double* y) {
    std::vector<int> rp(row_ptr, row_ptr + rows + 1);
    std::vector<int> ci(col_idx, col_idx + nnz);
    std::vector<double> v(vals, vals + nnz);
    anonymouslibHandle<int, unsigned int, double> A(rows, cols);
    int st = A.inputCSR(nnz, rp.data(), ci.data(), v.data());
    if (st != ANONYMOUSLIB_SUCCESS)
        return st;
    A.setSigma(sigma);
    {
        QuietStdout quiet;
        st = A.asCSR5();
    }
    if (st != ANONYMOUSLIB_SUCCESS)
        return st;
    A.setX(const_cast<double*>(x));
    std::fill(y, y + rows, 0.0);
    st = A.spmv(1.0, y);
    A.destroy();
    return st;
}

} // extern "C"
