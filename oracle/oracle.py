"""ctypes wrapper over oracle/liboracle.so (our C restatement) and, when built,
oracle/_ref/libdimref.so (the reference's own CSR5 header compiled unchanged).

TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")

DIRTY_BIT = 0x80000000
IDX_MASK = 0x7FFFFFFF


class _OrcCsr5(C.Structure):
    _fields_ = [
        ("sigma", C.c_uint32), ("omega", C.c_uint32), ("rows", C.c_uint32), ("cols", C.c_uint32),
        ("nnz", C.c_uint64), ("tile_count", C.c_uint64),
        ("tile_ptr", C.POINTER(C.c_uint32)), ("y_offset", C.POINTER(C.c_uint32)),
        ("scansum_offset", C.POINTER(C.c_uint32)), ("bit_flag", C.POINTER(C.c_uint64)),
        ("tile_desc_offset_ptr", C.POINTER(C.c_uint32)), ("tile_desc_offset", C.POINTER(C.c_uint32)),
        ("n_tile_desc_offset", C.c_uint64),
        ("vals", C.POINTER(C.c_double)), ("col_idx", C.POINTER(C.c_uint32)), ("row_ptr", C.POINTER(C.c_uint32)),
    ]


class _CgResult(C.Structure):
    _fields_ = [("iterations", C.c_uint32), ("converged", C.c_int32), ("relres", C.c_double),
                ("relres_true", C.c_double)]


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle not built: {path} (run `make -C oracle`)")
        L = C.CDLL(path)
        L.orc_csr5_from_csr.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64,
                                        _u32p, _u32p, _f64p, C.POINTER(C.POINTER(_OrcCsr5))]
        L.orc_csr5_from_csr.restype = C.c_int
        L.orc_csr5_free.argtypes = [C.POINTER(_OrcCsr5)]
        L.orc_csr5_pack_desc.argtypes = [C.POINTER(_OrcCsr5), C.c_uint64, _u32p]
        L.orc_csr5_spmv.argtypes = [C.POINTER(_OrcCsr5), C.c_uint32, _f64p, C.c_uint64, _f64p, C.c_uint64]
        L.orc_csr_spmv.argtypes = [C.c_uint32, _u32p, _u32p, _f64p, _f64p, _f64p]
        L.orc_csr_spmv.restype = None
        L.orc_csr_block_spmv.argtypes = [C.c_uint32, _u32p, _u32p, _f64p, C.c_uint64, C.c_uint64, _f64p,
                                         C.c_uint32, C.c_uint32, _f64p]
        L.orc_dot.argtypes = [C.c_uint64, _f64p, _f64p]
        L.orc_dot.restype = C.c_double
        L.orc_axpy.argtypes = [C.c_uint64, _f64p, _f64p, C.c_double]
        L.orc_aypx.argtypes = [C.c_uint64, _f64p, _f64p, C.c_double]
        L.orc_cg.argtypes = [C.POINTER(_OrcCsr5), C.c_uint32, _f64p, _f64p, C.c_double, C.c_uint32,
                             C.POINTER(_CgResult), C.c_void_p]
        L.orc_lane_inclusive_scan.argtypes = [C.c_uint32, _f64p, _f64p]
        L.orc_fast_segmented_sum.argtypes = [C.c_uint32, _f64p, _u32p, _f64p]
        L.orc_shuffle_relative.argtypes = [C.c_uint32, _f64p, _u32p, _f64p]
        L.orc_lane_sum.argtypes = [C.c_uint32, _f64p]
        L.orc_lane_sum.restype = C.c_double
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(os.path.join(_HERE, "_ref", "libdimref.so"))


def ref():
    """The reference's own CSR5 code (oracle/_ref/libdimref.so)."""
    global _ref
    if _ref is None:
        path = os.path.join(_HERE, "_ref", "libdimref.so")
        if not os.path.exists(path):
            raise RuntimeError(f"reference build absent: {path}")
        L = C.CDLL(path)
        L.dimref_csr5_from_csr.argtypes = [C.c_uint32, C.c_uint32, C.c_uint64, _u32p, _u32p, _f64p]
        L.dimref_csr5_from_csr.restype = C.c_void_p
        L.dimref_csr5_free.argtypes = [C.c_void_p]
        L.dimref_csr5_tile_count.argtypes = [C.c_void_p]
        L.dimref_csr5_tile_count.restype = C.c_uint64
        L.dimref_csr5_tile_desc_offset_size.argtypes = [C.c_void_p]
        L.dimref_csr5_tile_desc_offset_size.restype = C.c_uint64
        L.dimref_csr5_tail_partition_start.argtypes = [C.c_void_p]
        L.dimref_csr5_tail_partition_start.restype = C.c_uint32
        L.dimref_csr5_copy.argtypes = [C.c_void_p, _u32p, _u32p, _u32p, _u32p, _f64p, _u32p]
        L.dimref_csr5_spmv.argtypes = [C.c_void_p, C.c_int, _f64p, C.c_uint64, _f64p, C.c_uint64]
        L.dimref_csr5_time_spmv.argtypes = [C.c_void_p, C.c_int, _f64p, C.c_uint64, _f64p, C.c_uint64, C.c_int,
                                            C.c_int, _f64p]
        L.dimref_cg.argtypes = [C.c_void_p, C.c_int, _f64p, _f64p, C.c_double, C.c_uint32, _f64p]
        L.dimref_session_new.argtypes = [C.c_void_p, C.c_int]
        L.dimref_session_new.restype = C.c_void_p
        L.dimref_session_free.argtypes = [C.c_void_p]
        L.dimref_session_spmv.argtypes = [C.c_void_p, _f64p, C.c_uint64, _f64p, C.c_uint64]
        L.dimref_csr_from_coo.argtypes = [C.c_uint32, C.c_uint32, C.c_uint64, C.c_int, _u32p, _u32p, _f64p,
                                          _u32p, _u32p, _f64p]
        L.dimref_csr_from_coo.restype = C.c_uint64
        _ref = L
    return _ref


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(int(n),)).copy()


class Csr5:
    """The oracle's CSR5 matrix (reference layout, any sigma/omega <= 64)."""

    def __init__(self, rows, cols, row_ptr, col_idx, vals, sigma=16, omega=4):
        self.row_ptr = np.ascontiguousarray(row_ptr, dtype=np.uint32)
        self.col_idx_csr = np.ascontiguousarray(col_idx, dtype=np.uint32)
        self.vals_csr = np.ascontiguousarray(vals, dtype=np.float64)
        self.rows, self.cols = int(rows), int(cols)
        h = C.POINTER(_OrcCsr5)()
        st = lib().orc_csr5_from_csr(sigma, omega, self.rows, self.cols, len(self.vals_csr), self.row_ptr,
                                     self.col_idx_csr, self.vals_csr, C.byref(h))
        if st != 0:
            raise ValueError(f"orc_csr5_from_csr failed ({st})")
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_csr5_free(self._h)
            self._h = None

    @property
    def info(self):
        m = self._h.contents
        T, w = m.tile_count, m.omega
        return dict(
            tile_count=int(T), sigma=int(m.sigma), omega=int(w),
            tile_ptr=_arr(m.tile_ptr, T + 1, np.uint32),
            y_offset=_arr(m.y_offset, T * w, np.uint32).reshape(-1, w),
            scansum_offset=_arr(m.scansum_offset, T * w, np.uint32).reshape(-1, w),
            bit_flag=_arr(m.bit_flag, T * w, np.uint64).reshape(-1, w),
            tile_desc_offset_ptr=_arr(m.tile_desc_offset_ptr, T + 1, np.uint32),
            tile_desc_offset=_arr(m.tile_desc_offset, m.n_tile_desc_offset, np.uint32),
            vals=_arr(m.vals, m.nnz, np.float64), col_idx=_arr(m.col_idx, m.nnz, np.uint32),
        )

    def packed_desc(self):
        T = self._h.contents.tile_count
        out = np.zeros((T, 4), dtype=np.uint32)
        for t in range(T):
            w = np.zeros(4, dtype=np.uint32)
            lib().orc_csr5_pack_desc(self._h, t, w)
            out[t] = w
        return out

    def spmv(self, x, nthreads=1):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(self.rows, dtype=np.float64)
        st = lib().orc_csr5_spmv(self._h, nthreads, x, len(x), y, len(y))
        if st != 0:
            raise ValueError("can't multiply, vector size doesn't match column count of the matrix")
        return y

    def cg(self, b, x0=None, rtol=1e-8, maxit=5000, nthreads=8, history=False):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.rows) if x0 is None else np.array(x0, dtype=np.float64)
        hist = np.zeros(maxit + 1) if history else None
        res = _CgResult()
        st = lib().orc_cg(self._h, nthreads, b, x, rtol, maxit, C.byref(res),
                          hist.ctypes.data if hist is not None else None)
        out = dict(x=x, iterations=res.iterations, converged=bool(res.converged), relres=res.relres,
                   relres_true=res.relres_true, status=st)
        if hist is not None:
            out["history"] = hist[: res.iterations + 1]
        return out


def csr_spmv(row_ptr, col_idx, vals, x):
    rows = len(row_ptr) - 1
    y = np.empty(rows)
    lib().orc_csr_spmv(rows, np.ascontiguousarray(row_ptr, np.uint32), np.ascontiguousarray(col_idx, np.uint32),
                       np.ascontiguousarray(vals, np.float64), np.ascontiguousarray(x, np.float64), y)
    return y


def csr_block_spmv(row_ptr, col_idx, vals, nz_begin, nz_end, x, first_row, n_out):
    y = np.empty(n_out)
    st = lib().orc_csr_block_spmv(len(row_ptr) - 1, np.ascontiguousarray(row_ptr, np.uint32),
                                  np.ascontiguousarray(col_idx, np.uint32), np.ascontiguousarray(vals, np.float64),
                                  nz_begin, nz_end, np.ascontiguousarray(x, np.float64), first_row, n_out, y)
    if st:
        raise ValueError("bad block")
    return y


class RefCsr5:
    """The reference's csr5<double> (sigma=16, omega=4) built from its own header."""

    def __init__(self, rows, cols, row_ptr, col_idx, vals):
        self.rows, self.cols = int(rows), int(cols)
        rp = np.ascontiguousarray(row_ptr, np.uint32)
        ci = np.ascontiguousarray(col_idx, np.uint32)
        v = np.ascontiguousarray(vals, np.float64)
        self.nnz = len(v)
        self._h = ref().dimref_csr5_from_csr(self.rows, self.cols, self.nnz, rp, ci, v)

    def __del__(self):
        if getattr(self, "_h", None):
            ref().dimref_csr5_free(self._h)
            self._h = None

    @property
    def info(self):
        L = ref()
        T = L.dimref_csr5_tile_count(self._h)
        ntdo = L.dimref_csr5_tile_desc_offset_size(self._h)
        tile_ptr = np.zeros(T + 1, np.uint32)
        desc = np.zeros(max(T, 1) * 4, np.uint32)
        tdo_ptr = np.zeros(T + 1, np.uint32)
        tdo = np.zeros(max(ntdo, 1), np.uint32)
        vals = np.zeros(max(self.nnz, 1))
        cols = np.zeros(max(self.nnz, 1), np.uint32)
        L.dimref_csr5_copy(self._h, tile_ptr, desc, tdo_ptr, tdo, vals, cols)
        return dict(tile_count=int(T), tile_ptr=tile_ptr, tile_desc=desc[: T * 4].reshape(-1, 4),
                    tile_desc_offset_ptr=tdo_ptr, tile_desc_offset=tdo[:ntdo], vals=vals[: self.nnz],
                    col_idx=cols[: self.nnz], tail_partition_start=int(L.dimref_csr5_tail_partition_start(self._h)))

    def spmv(self, x, nthreads=1):
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty(self.rows)
        if ref().dimref_csr5_spmv(self._h, nthreads, x, len(x), y, len(y)):
            raise ValueError("can't multiply, vector size doesn't match column count of the matrix")
        return y

    def time_spmv(self, x, nthreads, warmup=5, runs=20):
        """The reference's benchmark protocol: seconds of each of `runs`
        products after `warmup` (only csr5::spmv is timed; y is zeroed and the
        calibrator fresh before each). Returns (times, y)."""
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty(self.rows)
        t = np.zeros(runs)
        ref().dimref_csr5_time_spmv(self._h, int(nthreads), x, len(x), y, len(y), int(warmup), int(runs), t)
        return t, y

    def cg(self, b, x0=None, rtol=1e-8, maxit=5000, nthreads=8):
        """Textbook CG with the reference's SpMV (dimref_cg): returns dict of
        x, iterations, converged, relres, relres_true and seconds."""
        x = np.zeros(self.rows) if x0 is None else np.array(x0, np.float64)
        out = np.zeros(5)
        ref().dimref_cg(self._h, int(nthreads), np.ascontiguousarray(b, np.float64), x, float(rtol), int(maxit), out)
        return dict(x=x, iterations=int(out[0]), converged=bool(out[1]), relres=float(out[2]),
                    relres_true=float(out[3]), seconds=float(out[4]))

    def session(self, nthreads):
        """The reference's own calling convention: one y and one calibrator reused."""
        return _RefSession(self, nthreads)


class _RefSession:
    def __init__(self, m: RefCsr5, nthreads):
        self.m = m
        self._s = ref().dimref_session_new(m._h, nthreads)

    def spmv(self, x, y):
        return ref().dimref_session_spmv(self._s, x, len(x), y, len(y))

    def __del__(self):
        if getattr(self, "_s", None):
            ref().dimref_session_free(self._s)
            self._s = None


def ref_csr_from_coo(rows, cols, ri, ci, v, symmetric=False):
    nnz = len(v)
    cap = 2 * nnz if symmetric else nnz
    rp = np.zeros(rows + 1, np.uint32)
    oc = np.zeros(max(cap, 1), np.uint32)
    ov = np.zeros(max(cap, 1))
    n = ref().dimref_csr_from_coo(rows, cols, nnz, int(symmetric), np.ascontiguousarray(ri, np.uint32),
                                  np.ascontiguousarray(ci, np.uint32), np.ascontiguousarray(v, np.float64), rp, oc, ov)
    return rp, oc[:n], ov[:n]


# ---------------------------------------------------------------------------
# The reference's own HDF5 I/O (oracle/_ref/libdimref_h5.so, built from
# libs/dim/io/h5.cpp and friends against the HDF5 C library). Used by the tests
# to write files the way the reference's store-matrix does and to load files
# the way its benchmarks do.

_ref_h5 = None

# ---- The original CSR5 that the reference vendors (third-party/csr5_orig:
# Liu & Vinter's AVX2 anonymouslib, omega = 4), built by oracle/Makefile into
# oracle/_ref/libcsr5orig.so: a second CPU oracle for cross-checks.

_orig = None


def orig_available() -> bool:
    return os.path.exists(os.path.join(_HERE, "_ref", "libcsr5orig.so"))


def orig_spmv(rows, cols, row_ptr, col_idx, vals, x, sigma=16):
    """y = A x by the original CSR5 (y zero-filled first: it never writes
    empty rows). Its kernel has sigma = 16 compiled in
    (ANONYMOUSLIB_CSR5_SIGMA, detail/avx2/common_avx2.h:13): other tile
    heights are refused."""
    if sigma != 16:
        raise ValueError("the original CSR5 kernel is compiled for sigma = 16")
    global _orig
    if _orig is None:
        L = C.CDLL(os.path.join(_HERE, "_ref", "libcsr5orig.so"))
        i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        L.csr5orig_spmv.argtypes = [C.c_int, C.c_int, C.c_int, i32p, i32p, _f64p, C.c_int, _f64p, _f64p]
        L.csr5orig_spmv.restype = C.c_int
        _orig = L
    if rows >= 2**31 or cols >= 2**31 or len(vals) >= 2**31:
        raise ValueError("the original CSR5 takes 32-bit signed sizes")
    y = np.empty(rows)
    st = _orig.csr5orig_spmv(rows, cols, len(vals), np.ascontiguousarray(row_ptr, np.int32),
                             np.ascontiguousarray(col_idx, np.int32), np.ascontiguousarray(vals, np.float64),
                             int(sigma), np.ascontiguousarray(x, np.float64), y)
    if st != 0:
        raise RuntimeError(f"original CSR5 returned status {st}")
    return y


def ref_h5_available() -> bool:
    return os.path.exists(os.path.join(_HERE, "_ref", "libdimref_h5.so"))


def _rh5():
    global _ref_h5
    if _ref_h5 is None:
        L = C.CDLL(os.path.join(_HERE, "_ref", "libdimref_h5.so"))
        for name in ("dimref_h5_store_csr5", "dimref_h5_store_matlab"):
            f = getattr(L, name)
            f.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_uint32, C.c_uint32, C.c_uint64, _u32p, _u32p, _f64p]
            f.restype = C.c_int
        L.dimref_h5_load_csr5.argtypes = [C.c_char_p, C.c_char_p]
        L.dimref_h5_load_csr5.restype = C.c_void_p
        L.dimref_h5_csr5_free.argtypes = [C.c_void_p]
        L.dimref_h5_csr5_sizes.argtypes = [C.c_void_p, np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")]
        L.dimref_h5_csr5_copy.argtypes = [C.c_void_p, _u32p, _u32p, _u32p, _u32p, _u32p, _f64p, _u32p]
        L.dimref_h5_load_csr5_partial.argtypes = [C.c_char_p, C.c_char_p, C.c_uint64, C.c_uint64]
        L.dimref_h5_load_csr5_partial.restype = C.c_void_p
        L.dimref_h5_part_info.argtypes = [C.c_void_p, np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")]
        L.dimref_h5_part_copy.argtypes = [C.c_void_p, _u32p, _u32p, _f64p, _u32p]
        L.dimref_h5_part_spmv.argtypes = [C.c_void_p, C.c_int, _f64p, _f64p]
        L.dimref_h5_part_spmv.restype = C.c_int
        L.dimref_h5_read_matlab.argtypes = [C.c_char_p, C.c_char_p]
        L.dimref_h5_read_matlab.restype = C.c_void_p
        L.dimref_h5_csr_free.argtypes = [C.c_void_p]
        L.dimref_h5_csr_sizes.argtypes = [C.c_void_p, np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")]
        L.dimref_h5_csr_copy.argtypes = [C.c_void_p, _u32p, _u32p, _f64p]
        _ref_h5 = L
    return _ref_h5


def _csr_args(rows, cols, rp, ci, v):
    rp = np.ascontiguousarray(rp, dtype=np.uint32)
    ci = np.ascontiguousarray(ci, dtype=np.uint32)
    v = np.ascontiguousarray(v, dtype=np.float64)
    return rows, cols, len(v), rp, ci, v


def ref_h5_store_csr5(path, group, rows, cols, rp, ci, v, append=False):
    """h5::store(csr5::from_csr(csr)) as store_matrix -f csr5 does."""
    if _rh5().dimref_h5_store_csr5(os.fspath(path).encode(), group.encode(), int(append),
                                   *_csr_args(rows, cols, rp, ci, v)):
        raise RuntimeError("reference h5::store failed")


def ref_h5_store_matlab(path, group, rows, cols, rp, ci, v, append=False):
    if _rh5().dimref_h5_store_matlab(os.fspath(path).encode(), group.encode(), int(append),
                                     *_csr_args(rows, cols, rp, ci, v)):
        raise RuntimeError("reference write_matlab_compatible failed")


def ref_h5_load_csr5(path, group="A") -> dict:
    """h5::load_csr5: the arrays exactly as stored (tile-transposed values,
    packed tile_desc words)."""
    L = _rh5()
    h = L.dimref_h5_load_csr5(os.fspath(path).encode(), group.encode())
    if not h:
        raise RuntimeError("reference load_csr5 failed")
    try:
        sz = np.zeros(5, np.uint64)
        L.dimref_h5_csr5_sizes(h, sz)
        rows, cols, nnz, T, ntdo = (int(s) for s in sz)
        out = dict(rows=rows, cols=cols, row_ptr=np.zeros(rows + 1, np.uint32), tile_ptr=np.zeros(T + 1, np.uint32),
                   tile_desc=np.zeros(4 * T, np.uint32), tile_desc_offset_ptr=np.zeros(T + 1, np.uint32),
                   tile_desc_offset=np.zeros(ntdo, np.uint32), vals=np.zeros(nnz), col_idx=np.zeros(nnz, np.uint32))
        L.dimref_h5_csr5_copy(h, out["row_ptr"], out["tile_ptr"], out["tile_desc"], out["tile_desc_offset_ptr"],
                              out["tile_desc_offset"], out["vals"], out["col_idx"])
        return out
    finally:
        L.dimref_h5_csr5_free(h)


def ref_h5_load_csr5_partial(path, group, idx, count, x=None, nthreads=4) -> dict:
    """h5::load_csr5_partial (h5.cpp:137-173) of part `idx` of `count`: the
    part's global dimensions, first_row_idx() / last_row_idx(), val_offset,
    skip_tail, tile count and loaded slabs; with x, also y_block from the
    reference's csr5::spmv<partial> on that part (zero-filled y, fresh
    calibrator)."""
    L = _rh5()
    h = L.dimref_h5_load_csr5_partial(os.fspath(path).encode(), group.encode(), int(idx), int(count))
    if not h:
        raise RuntimeError("reference load_csr5_partial failed")
    try:
        info = np.zeros(10, np.uint64)
        L.dimref_h5_part_info(h, info)
        rows, cols, first, last, val_offset, skip_tail, T, nv, nrp, uncapped = (int(i) for i in info)
        out = dict(rows=rows, cols=cols, first_row=first, last_row=last, val_offset=val_offset,
                   skip_tail=bool(skip_tail), first_tile_uncapped=bool(uncapped), tile_count=T,
                   row_ptr=np.zeros(nrp, np.uint32),
                   tile_ptr=np.zeros(T + 1, np.uint32), vals=np.zeros(nv), col_idx=np.zeros(nv, np.uint32))
        L.dimref_h5_part_copy(h, out["row_ptr"], out["tile_ptr"], out["vals"], out["col_idx"])
        if x is not None:
            y = np.zeros(last - first + 1)
            if L.dimref_h5_part_spmv(h, int(nthreads), np.ascontiguousarray(x, np.float64), y):
                raise RuntimeError("reference spmv<partial> failed")
            out["y_block"] = y
        return out
    finally:
        L.dimref_h5_csr5_free(h)


def ref_h5_read_matlab(path, group="A"):
    L = _rh5()
    h = L.dimref_h5_read_matlab(os.fspath(path).encode(), group.encode())
    if not h:
        raise RuntimeError("reference read_matlab_compatible failed")
    try:
        sz = np.zeros(3, np.uint64)
        L.dimref_h5_csr_sizes(h, sz)
        rows, cols, nnz = (int(s) for s in sz)
        rp, ci, v = np.zeros(rows + 1, np.uint32), np.zeros(nnz, np.uint32), np.zeros(nnz)
        L.dimref_h5_csr_copy(h, rp, ci, v)
        return rows, cols, rp, ci, v
    finally:
        L.dimref_h5_csr_free(h)
