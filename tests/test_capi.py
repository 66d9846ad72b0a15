"""CPU checks of the C ABI boundary: the library loads, exports every symbol
include/dim_gpu.h declares, and its argument validation behaves like the
reference (no GPU compute is issued here)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from diploma_research_amd import native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "dim_gpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(dim_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_what_the_binding_lists():
    assert _declared() == sorted(N.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = N.lib()
    for name in _declared():
        assert hasattr(L, name), name
    assert L.dim_version() == 1


def test_status_strings():
    L = N.lib()
    assert L.dim_status_string(N.DIM_SUCCESS) == b"success"
    assert L.dim_status_string(N.DIM_BREAKDOWN).startswith(b"breakdown")


def test_invalid_csr_rejected_before_touching_the_gpu():
    L = N.lib()
    h = C.c_void_p()
    rp = np.array([0, 2, 1, 2], np.uint32)  # decreasing
    ci = np.array([0, 1], np.uint32)
    v = np.ones(2)
    st = L.dim_csr5_from_csr(3, 2, 2, rp.ctypes.data, ci.ctypes.data, v.ctypes.data, None, C.byref(h))
    assert st == N.DIM_INVALID_ARGUMENT
    assert b"non-decreasing" in L.dim_last_error()
    rp = np.array([0, 1, 2], np.uint32)
    ci = np.array([0, 5], np.uint32)  # out of range column
    st = L.dim_csr5_from_csr(2, 2, 2, rp.ctypes.data, ci.ctypes.data, v.ctypes.data, None, C.byref(h))
    assert st == N.DIM_INVALID_ARGUMENT
    opts = N.Csr5Opts(9, 0)  # unsupported sigma
    ci = np.array([0, 1], np.uint32)
    st = L.dim_csr5_from_csr(2, 2, 2, rp.ctypes.data, ci.ctypes.data, v.ctypes.data, C.byref(opts), C.byref(h))
    assert st == N.DIM_INVALID_ARGUMENT
    with pytest.raises(ValueError):
        N.check(st)


def test_null_handle_rejected():
    L = N.lib()
    assert L.dim_csr5_spmv(None, 0, None, 0, None, 0, None) == N.DIM_INVALID_ARGUMENT
    assert L.dim_cg(None, None, None, None, None, None) in (N.DIM_INVALID_ARGUMENT,)


def test_cpp_headers_compile_with_the_reference_spelling(tmp_path):
    """Host code spelled with the reference's names (dim::mat::csr5<double>,
    csr5::spmv_strategy::partial, dim::io::h5::load_csr5_partial) compiles
    against include/dim/gpu/*.hpp (syntax and types only)."""
    import shutil
    import subprocess

    cxx = shutil.which("g++")
    if not cxx:
        pytest.skip("no g++")
    src = tmp_path / "use.cpp"
    src.write_text(r'''
#include <dim/gpu/csr5.hpp>
#include <dim/gpu/h5.hpp>
using matrix_type = dim::mat::csr5<double>;
void rank_share(const double* x, double* y) {
    auto A = dim::io::h5::load_csr5_partial("m.csr5.h5", "A", dim::io::h5::partial_identifier_t{1, 8});
    auto calibrator = A.create_calibrator();
    const size_t n = A.last_row_idx() - A.first_row_idx() + 1;
    A.spmv<matrix_type::spmv_strategy::partial>(
        matrix_type::spmv_data_t{{x, A.dimensions().cols}, {y, n}, calibrator});
    dim::mat::csr<double> c{dim::mat::dimensions_t{2, 2}, {1.0, 1.0}, {0, 1, 2}, {0, 1}};
    auto B = matrix_type::from_csr(c);
    auto W = dim::io::h5::load_csr5("m.csr5.h5", "A");
    (void)B; (void)W;
}
''')
    p = subprocess.run([cxx, "-std=c++20", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), str(src)],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert p.returncode == 0, p.stdout
