"""ctypes binding of libdim_h5.so (include/dim_h5.h): the reference's HDF5
matrix groups (CSR5 and MATLAB CSR, libs/dim/io/h5.cpp:45-74,
include/dim/io/h5.h:68-97) and FP64 vector datasets.

    from diploma_research_amd import h5
    h5.write_csr5("m.csr5.h5", "A", rows, cols, row_ptr, col_idx, vals)
    rows, cols, rp, ci, v = h5.read_csr("m.csr5.h5", "A")
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdim_h5.so")

CSR5, MATLAB_CSR = 1, 2


class H5Csr(C.Structure):
    _fields_ = [("rows", C.c_uint32), ("cols", C.c_uint32), ("nnz", C.c_uint64),
                ("row_ptr", C.POINTER(C.c_uint32)), ("col_idx", C.POINTER(C.c_uint32)),
                ("vals", C.POINTER(C.c_double))]


class DatasetProps(C.Structure):
    _fields_ = [("chunk", C.c_uint64), ("deflate", C.c_int32)]


class CsrProps(C.Structure):
    _fields_ = [(n, DatasetProps) for n in ("values", "col_idx", "row_idx")]


class Csr5Props(C.Structure):
    _fields_ = [(n, DatasetProps) for n in ("vals", "col_idx", "row_ptr", "tile_ptr", "tile_desc",
                                            "tile_desc_offset", "tile_desc_offset_ptr")]


class H5Csr5Part(C.Structure):
    _fields_ = [("rows", C.c_uint32), ("cols", C.c_uint32), ("first_row", C.c_uint32), ("last_row", C.c_uint32),
                ("tile_begin", C.c_uint64), ("tile_count", C.c_uint64), ("val_offset", C.c_uint64),
                ("skip_tail", C.c_int32), ("block", H5Csr), ("first_tile_uncapped", C.c_int32)]


class H5Error(RuntimeError):
    pass


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libdim_h5.so not built: {LIB_PATH} (run `python -m diploma_research_amd.build`)")
        L = C.CDLL(LIB_PATH)
        L.dim_h5_last_error.restype = C.c_char_p
        L.dim_h5_is_hdf5.argtypes = [C.c_char_p]
        L.dim_h5_group_kind_of.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(C.c_int)]
        L.dim_h5_read_csr.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(H5Csr)]
        L.dim_h5_free_csr.argtypes = [C.POINTER(H5Csr)]
        L.dim_h5_write_matlab_csr.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.POINTER(H5Csr), C.POINTER(CsrProps)]
        L.dim_h5_write_csr5.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.POINTER(H5Csr), C.POINTER(Csr5Props)]
        L.dim_h5_write_vector.argtypes = [C.c_char_p, C.c_char_p, C.c_void_p, C.c_uint64, C.c_int]
        L.dim_h5_read_vector.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(C.POINTER(C.c_double)),
                                         C.POINTER(C.c_uint64)]
        L.dim_h5_free.argtypes = [C.c_void_p]
        L.dim_h5_load_csr5_partial.argtypes = [C.c_char_p, C.c_char_p, C.c_uint32, C.c_uint32,
                                               C.POINTER(H5Csr5Part)]
        L.dim_h5_free_csr5_part.argtypes = [C.POINTER(H5Csr5Part)]
        L.dim_csr5_load_h5.argtypes = [C.c_char_p, C.c_char_p, C.c_void_p, C.POINTER(C.c_void_p)]
        L.dim_csr5_load_h5_partial.argtypes = [C.c_char_p, C.c_char_p, C.c_uint32, C.c_uint32, C.c_void_p,
                                               C.POINTER(C.c_void_p), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        _lib = L
    return _lib


EXPORTS = ["dim_h5_last_error", "dim_h5_is_hdf5", "dim_h5_group_kind_of", "dim_h5_read_csr", "dim_h5_free_csr",
           "dim_h5_write_matlab_csr", "dim_h5_write_csr5", "dim_h5_write_vector", "dim_h5_read_vector",
           "dim_h5_free", "dim_h5_load_csr5_partial", "dim_h5_free_csr5_part", "dim_csr5_load_h5",
           "dim_csr5_load_h5_partial"]


def _check(st):
    if st != 0:
        msg = lib().dim_h5_last_error().decode()
        if st == 1:
            raise ValueError(msg)
        raise H5Error(msg)


def _b(s):
    return os.fspath(s).encode()


def is_hdf5(path) -> bool:
    return lib().dim_h5_is_hdf5(_b(path)) == 1


def group_kind(path, group="A") -> int:
    k = C.c_int(0)
    _check(lib().dim_h5_group_kind_of(_b(path), _b(group), C.byref(k)))
    return k.value


def read_csr(path, group="A"):
    """(rows, cols, row_ptr, col_idx, vals) of a CSR5 or MATLAB CSR group."""
    m = H5Csr()
    _check(lib().dim_h5_read_csr(_b(path), _b(group), C.byref(m)))
    try:
        rp = np.ctypeslib.as_array(m.row_ptr, (m.rows + 1,)).copy()
        ci = np.ctypeslib.as_array(m.col_idx, (m.nnz,)).copy() if m.nnz else np.zeros(0, np.uint32)
        v = np.ctypeslib.as_array(m.vals, (m.nnz,)).copy() if m.nnz else np.zeros(0)
        return m.rows, m.cols, rp, ci, v
    finally:
        lib().dim_h5_free_csr(C.byref(m))


def load_csr5_partial(path, group, part_idx, part_count) -> dict:
    """load_csr5_partial (libs/dim/io/h5.cpp:137-173): part part_idx of
    part_count of a CSR5 group's tiles, as one rank loads it. Returns the
    whole matrix's rows/cols, first_row / last_row (first_row_idx(),
    last_row_idx()), tile_begin, tile_count, val_offset, skip_tail,
    first_tile_uncapped and
    `block` = (n_rows, cols, row_ptr, col_idx, vals): the part's elements as
    CSR over rows first_row .. last_row, whose product is the part's share
    of y (edge rows partial)."""
    m = H5Csr5Part()
    _check(lib().dim_h5_load_csr5_partial(_b(path), _b(group), int(part_idx), int(part_count), C.byref(m)))
    try:
        b = m.block
        rp = np.ctypeslib.as_array(b.row_ptr, (b.rows + 1,)).copy()
        ci = np.ctypeslib.as_array(b.col_idx, (b.nnz,)).copy() if b.nnz else np.zeros(0, np.uint32)
        v = np.ctypeslib.as_array(b.vals, (b.nnz,)).copy() if b.nnz else np.zeros(0)
        return dict(rows=m.rows, cols=m.cols, first_row=m.first_row, last_row=m.last_row, tile_begin=m.tile_begin,
                    tile_count=m.tile_count, val_offset=m.val_offset, skip_tail=bool(m.skip_tail),
                    first_tile_uncapped=bool(m.first_tile_uncapped), block=(int(b.rows), int(b.cols), rp, ci, v))
    finally:
        lib().dim_h5_free_csr5_part(C.byref(m))


def load_csr5_gpu(path, group="A", sigma=16, keep_f32=False):
    """h5::load_csr5 straight to a GPU matrix (dim_csr5_load_h5): a Csr5."""
    from . import native as N
    from .csr5 import Csr5

    opts = N.Csr5Opts(sigma, N.DIM_CSR5_KEEP_F32 if keep_f32 else 0)
    h = C.c_void_p()
    _check(lib().dim_csr5_load_h5(_b(path), _b(group), C.cast(C.byref(opts), C.c_void_p), C.byref(h)))
    return Csr5(h.value)


def load_csr5_partial_gpu(path, group, part_idx, part_count, sigma=16, keep_f32=False):
    """load_csr5_partial straight to a GPU matrix of the part's share
    (dim_csr5_load_h5_partial): returns (first_row, Csr5); its products give
    y rows first_row .. first_row + rows - 1."""
    from . import native as N
    from .csr5 import Csr5

    opts = N.Csr5Opts(sigma, N.DIM_CSR5_KEEP_F32 if keep_f32 else 0)
    h = C.c_void_p()
    first, n = C.c_uint32(), C.c_uint32()
    _check(lib().dim_csr5_load_h5_partial(_b(path), _b(group), int(part_idx), int(part_count),
                                          C.cast(C.byref(opts), C.c_void_p), C.byref(h), C.byref(first),
                                          C.byref(n)))
    return first.value, Csr5(h.value)


def partition_block(rows, cols, row_ptr, col_idx, vals, part_idx, part_count):
    """The share load_csr5_partial returns, computed from an in-memory CSR
    matrix with the same rules (reference tiles of 64 elements, ceil(T /
    part_count) tiles per part, the last part taking the tail): returns
    (first_row, val_offset, block) with block = (n_rows, cols, row_ptr,
    col_idx, vals). For timing a rank's share without writing a file; the
    tests check it against the file loader."""
    rp = np.asarray(row_ptr, dtype=np.int64)
    nnz = len(vals)
    T = nnz // 64
    part = -(-T // part_count)
    t0 = part * part_idx
    last = part_idx + 1 == part_count
    if t0 > T or (not last and t0 == T):
        raise ValueError("part starts beyond the last tile: use fewer parts")
    nt = min(part, T - t0)
    v0 = t0 * 64
    v1 = nnz if last else v0 + nt * 64
    first = int(np.searchsorted(rp, v0, side="right") - 1)
    last_row = rows - 1 if last else int(np.searchsorted(rp, v1 - 1, side="right") - 1)
    n = last_row - first + 1
    brp = np.clip(rp[first:first + n + 1], v0, v1) - v0
    brp[-1] = v1 - v0
    return first, v0, (n, cols, brp.astype(np.uint32), np.asarray(col_idx[v0:v1], np.uint32),
                       np.asarray(vals[v0:v1], np.float64))


def _csr(rows, cols, row_ptr, col_idx, vals):
    rp = np.ascontiguousarray(row_ptr, dtype=np.uint32)
    ci = np.ascontiguousarray(col_idx, dtype=np.uint32)
    v = np.ascontiguousarray(vals, dtype=np.float64)
    m = H5Csr(rows, cols, len(v), rp.ctypes.data_as(C.POINTER(C.c_uint32)),
              ci.ctypes.data_as(C.POINTER(C.c_uint32)), v.ctypes.data_as(C.POINTER(C.c_double)))
    return m, (rp, ci, v)


def _props(cls, chunk, deflate):
    p = cls()
    for name, _ in cls._fields_:
        setattr(p, name, DatasetProps(chunk.get(name, 0) if chunk else 0, deflate.get(name, -1) if deflate else -1))
    return p


def write_matlab_csr(path, group, rows, cols, row_ptr, col_idx, vals, append=False, chunk=None, deflate=None):
    m, keep = _csr(rows, cols, row_ptr, col_idx, vals)
    _check(lib().dim_h5_write_matlab_csr(_b(path), _b(group), int(append), C.byref(m),
                                         C.byref(_props(CsrProps, chunk, deflate))))


def write_csr5(path, group, rows, cols, row_ptr, col_idx, vals, append=False, chunk=None, deflate=None):
    m, keep = _csr(rows, cols, row_ptr, col_idx, vals)
    _check(lib().dim_h5_write_csr5(_b(path), _b(group), int(append), C.byref(m),
                                   C.byref(_props(Csr5Props, chunk, deflate))))


def write_vector(path, name, v, overwrite=False):
    a = np.ascontiguousarray(v, dtype=np.float64)
    _check(lib().dim_h5_write_vector(_b(path), _b(name), a.ctypes.data, len(a), int(overwrite)))


def read_vector(path, name) -> np.ndarray:
    p = C.POINTER(C.c_double)()
    n = C.c_uint64(0)
    _check(lib().dim_h5_read_vector(_b(path), _b(name), C.byref(p), C.byref(n)))
    try:
        return np.ctypeslib.as_array(p, (n.value,)).copy() if n.value else np.zeros(0)
    finally:
        lib().dim_h5_free(p)
