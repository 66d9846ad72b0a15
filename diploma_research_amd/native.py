"""ctypes binding of libdim_gpu.so (the C ABI declared in include/dim_gpu.h).

The library is built in-tree by ``python -m diploma_research_amd.build``.
Loading fails loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdim_gpu.so")

DIM_SUCCESS = 0
DIM_INVALID_ARGUMENT = 1
DIM_HIP_ERROR = 2
DIM_OUT_OF_MEMORY = 3
DIM_NOT_CONVERGED = 4
DIM_BREAKDOWN = 5
DIM_IO_ERROR = 6
DIM_UNSUPPORTED = 7

DIM_F64 = 0
DIM_F32 = 1
DIM_CSR5_KEEP_F32 = 1
DIM_CSR5_XCD_SLICES = 2
DIM_CSR5_PLAIN_COLUMNS = 4
DIM_CG_F64 = 0
DIM_CG_F32_REFINED = 1
DIM_CG_TRANSPOSE = 1
DIM_CG_PERSISTENT = 2
DIM_CG_NOALLOC_MAX_ROWS = 2097152

# Every symbol include/dim_gpu.h declares (checked by the CPU test suite).
EXPORTS = [
    "dim_last_error", "dim_status_string", "dim_version",
    "dim_csr5_from_csr", "dim_csr5_from_csr_device", "dim_csr5_destroy", "dim_csr5_get_info",
    "dim_csr5_copy_format", "dim_csr5_spmv", "dim_csr5_spmv_host", "dim_csr5_spmv_timed", "dim_csr5_block_rows", "dim_csr5_spmv_block", "dim_csr5_spmm",
    "dim_csr5_spmm_colmajor", "dim_csr5_spmv_transpose", "dim_csr5_sampled_outer", "dim_cg", "dim_cg_host", "dim_cg_noalloc",
    "dim_cg_multi", "dim_csr5_from_csr_batched",
    "dim_cg_batched", "dim_csr5_n_systems", "dim_csr5_set_values", "dim_stream_read",
]


class Csr5Opts(C.Structure):
    _fields_ = [("sigma", C.c_uint32), ("flags", C.c_uint32)]


class Csr5Info(C.Structure):
    _fields_ = [("rows", C.c_uint32), ("cols", C.c_uint32), ("nnz", C.c_uint64), ("sigma", C.c_uint32),
                ("omega", C.c_uint32), ("tile_count", C.c_uint32), ("tail_partition_start", C.c_uint32),
                ("n_empty_rows", C.c_uint32), ("n_dirty_tiles", C.c_uint32), ("has_f32", C.c_uint32),
                ("device_bytes", C.c_uint64), ("xcd_slices", C.c_uint32), ("xcd_sigma", C.c_uint32),
                ("column_codes", C.c_uint32), ("reserved", C.c_uint32)]


class CgOpts(C.Structure):
    _fields_ = [("rtol", C.c_double), ("max_iters", C.c_uint32), ("precision", C.c_uint32),
                ("check_every", C.c_uint32), ("refine_delta", C.c_double), ("history", C.c_void_p),
                ("flags", C.c_uint32)]


class CgResult(C.Structure):
    _fields_ = [("iterations", C.c_uint32), ("converged", C.c_int32), ("relres", C.c_double),
                ("relres_true", C.c_double), ("refinements", C.c_uint32), ("seconds", C.c_double)]


class DimError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{message} (status {status})")
        self.status = status


_lib = None


def lib():
    """Loads libdim_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libdim_gpu.so not built: {LIB_PATH} (run `python -m diploma_research_amd.build`)")
    L = C.CDLL(LIB_PATH)
    vp, u32, u64, i32, f64 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int, C.c_double
    P = C.POINTER
    sig = {
        "dim_last_error": ([], C.c_char_p),
        "dim_status_string": ([i32], C.c_char_p),
        "dim_version": ([], i32),
        "dim_csr5_from_csr": ([u32, u32, u64, vp, vp, vp, P(Csr5Opts), P(vp)], i32),
        "dim_csr5_from_csr_device": ([u32, u32, u64, vp, vp, vp, P(Csr5Opts), vp, P(vp)], i32),
        "dim_csr5_destroy": ([vp], i32),
        "dim_csr5_get_info": ([vp, P(Csr5Info)], i32),
        "dim_csr5_copy_format": ([vp, vp, vp, vp, vp, vp, P(u64), vp], i32),
        "dim_csr5_spmv": ([vp, i32, vp, u64, vp, u64, vp], i32),
        "dim_csr5_spmv_host": ([vp, i32, vp, u64, vp, u64, vp], i32),
        "dim_csr5_spmv_timed": ([vp, i32, vp, u64, vp, u64, vp, P(C.c_float)], i32),
        "dim_csr5_block_rows": ([vp, u32, u32, P(u32), P(u32)], i32),
        "dim_csr5_spmv_block": ([vp, i32, u32, u32, vp, u64, vp, u64, vp], i32),
        "dim_csr5_spmm": ([vp, i32, u32, vp, vp, vp], i32),
        "dim_csr5_spmm_colmajor": ([vp, i32, u32, vp, u64, vp, u64, vp], i32),
        "dim_csr5_spmv_transpose": ([vp, i32, vp, vp, vp], i32),
        "dim_csr5_sampled_outer": ([vp, f64, vp, vp, vp, vp], i32),
        "dim_cg": ([vp, vp, vp, P(CgOpts), P(CgResult), vp], i32),
        "dim_cg_host": ([vp, vp, vp, P(CgOpts), P(CgResult), vp], i32),
        "dim_cg_noalloc": ([vp, vp, vp, P(CgOpts), P(CgResult), vp], i32),
        "dim_cg_multi": ([vp, u32, vp, vp, P(CgOpts), P(CgResult), vp], i32),
        "dim_csr5_from_csr_batched": ([u32, vp, vp, vp, vp, vp, P(Csr5Opts), P(vp)], i32),
        "dim_cg_batched": ([vp, vp, vp, P(CgOpts), P(CgResult), vp], i32),
        "dim_csr5_n_systems": ([vp], u32),
        "dim_csr5_set_values": ([vp, vp, vp], i32),
        "dim_stream_read": ([vp, u64, vp, vp], i32),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def check(status, allow=()):
    """Raises DimError for a non-success status (statuses in `allow` pass)."""
    if status == DIM_SUCCESS or status in allow:
        return status
    msg = lib().dim_last_error().decode(errors="replace")
    if status == DIM_INVALID_ARGUMENT:
        err = ValueError(msg)
        err.status = status
        raise err
    raise DimError(status, msg)
