"""Python mirror of dim's CSR5 interface, running on the MI355X through libdim_gpu.

Names and argument meaning follow the reference C++ surface:

    dim::mat::coo / csr / csr::from_coo   include/dim/mat/storage_formats/{coo,csr}.h
    dim::mat::csr5::from_csr               csr5.h:791-803
    csr5::spmv(spmv_data_t{x, y, calib})   csr5.h:280-285, 818-829
    csr5::create_calibrator                csr5.h:809
    csr5::tail_partition_start             csr5.h:806
    csr5::first_row_idx / last_row_idx     csr5.h:849-855
    csr5::non_zero_count                   csr5.h:880

Vectors are torch tensors on the GPU (float64, or float32 for the FP32
products). A size mismatch raises ValueError, as the reference throws
std::invalid_argument. y is overwritten (empty rows get 0).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import native as N


def _torch():
    import torch

    return torch


def _stream_ptr(stream=None):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _dtype_code(t):
    name = t.dtype.name if isinstance(t, np.ndarray) else str(t.dtype).replace("torch.", "")
    if name == "float64":
        return N.DIM_F64
    if name == "float32":
        return N.DIM_F32
    raise TypeError(f"unsupported dtype {t.dtype}")


# the reference's message for an x that does not match the column count
# (include/dim/mat/storage_formats/csr5.h:820-821)
_X_COLS = "can't multiply, vector size doesn't match column count of the matrix"


def _check_tensor(t, name, numel, dtypes=("float64",), len_msg=None):
    """Validates a tensor before its pointer crosses the C ABI, which takes no
    lengths for some arguments and cannot tell host from device memory: a GPU
    tensor of one of `dtypes` (torch dtype names), contiguous, with `numel`
    entries. Raises TypeError (not a tensor, wrong dtype, not on the GPU) or
    ValueError (not contiguous, wrong length)."""
    torch = _torch()
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch tensor, got {type(t).__name__}")
    if str(t.dtype).replace("torch.", "") not in dtypes:
        raise TypeError(f"{name} must have dtype {' or '.join(dtypes)}, got {t.dtype}")
    if not t.is_cuda:
        raise TypeError(f"{name} must be on the GPU, got device {t.device}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if t.numel() != numel:
        raise ValueError(len_msg or f"{name} must have {numel} entries, got {t.numel()}")
    return t


def _check_host(a, name, numel, dtypes=("float64",), len_msg=None):
    """Validates a host (numpy) vector for the host-pointer entry points: an
    ndarray of one of `dtypes`, C-contiguous, one-dimensional, with `numel`
    entries. Raises TypeError or ValueError as _check_tensor does."""
    if not isinstance(a, np.ndarray):
        raise TypeError(f"{name} must be a numpy array, got {type(a).__name__}")
    if a.dtype.name not in dtypes:
        raise TypeError(f"{name} must have dtype {' or '.join(dtypes)}, got {a.dtype}")
    if a.ndim != 1 or not a.flags.c_contiguous:
        raise ValueError(f"{name} must be a contiguous vector")
    if a.size != numel:
        raise ValueError(len_msg or f"{name} must have {numel} entries, got {a.size}")
    return a


@dataclass
class Dimensions:
    """dim::mat::dimensions_t (storage_formats/base.h:10-14)."""

    rows: int
    cols: int


@dataclass
class Coo:
    """dim::mat::coo (coo.h:17-47): a symmetric COO also yields (col, row) for off-diagonal entries."""

    dimensions: Dimensions
    values: np.ndarray
    row_indices: np.ndarray
    col_indices: np.ndarray
    symmetric: bool = False

    def iterate(self):
        r, c, v = self.row_indices, self.col_indices, self.values
        if not self.symmetric:
            return r, c, v
        off = r != c
        # the reference yields (v, row, col) then (v, col, row) per entry
        rr = np.empty(len(r) + int(off.sum()), dtype=r.dtype)
        cc = np.empty_like(rr)
        vv = np.empty(len(rr), dtype=v.dtype)
        pos = np.arange(len(r)) + np.concatenate([[0], np.cumsum(off)[:-1]])
        rr[pos], cc[pos], vv[pos] = r, c, v
        pos2 = pos[off] + 1
        rr[pos2], cc[pos2], vv[pos2] = c[off], r[off], v[off]
        return rr, cc, vv


@dataclass
class Csr:
    """dim::mat::csr (csr.h:19-78) with uint32 indices."""

    dimensions: Dimensions
    values: np.ndarray
    row_start_offsets: np.ndarray
    col_indices: np.ndarray

    @staticmethod
    def from_coo(coo: Coo) -> "Csr":
        """csr::from_coo (csr.h:42-77): count entries per row, exclusive scan,
        scatter in COO order, then sort every row by column (values follow)."""
        r, c, v = coo.iterate()
        rows = coo.dimensions.rows
        counts = np.bincount(np.asarray(r, np.int64), minlength=rows)
        row_ptr = np.zeros(rows + 1, dtype=np.uint32)
        np.cumsum(counts, out=row_ptr[1:])
        # scatter in COO order == stable sort by row; then per-row sort by column
        order = np.lexsort((np.asarray(c, np.int64), np.asarray(r, np.int64)))
        return Csr(coo.dimensions, np.asarray(v, np.float64)[order], row_ptr,
                   np.asarray(c, np.uint32)[order])

    @property
    def nnz(self):
        return len(self.values)


class Calibrator:
    """csr5::create_calibrator() (csr5.h:809). The GPU path keeps its carries
    in per-matrix device scratch, so this is a token kept for source
    compatibility; calls on one matrix must still not overlap."""


@dataclass
class SpmvData:
    """csr5::spmv_data_t (csr5.h:280-285)."""

    x: object
    y: object
    calibrator: Optional[Calibrator] = None


class Csr5:
    """dim::mat::csr5 on the GPU (omega = 64 lanes, sigma from the autotuner or given)."""

    omega = 64

    def __init__(self, handle, keepalive=None):
        self._h = C.c_void_p(handle)
        self._keepalive = keepalive
        info = N.Csr5Info()
        N.check(N.lib().dim_csr5_get_info(self._h, C.byref(info)))
        self.info = info
        self.dimensions = Dimensions(info.rows, info.cols)
        self.sigma = info.sigma

    # ---------------------------------------------------------- creation --
    @classmethod
    def from_csr(cls, csr, sigma: int = 16, keep_f32: bool = False, xcd_slices: bool = False,
                 plain_columns: bool = False) -> "Csr5":
        """csr5::from_csr (csr5.h:791-803). `csr` is a Csr or a (rows, cols,
        row_ptr, col_idx, vals) tuple of host arrays. sigma=0 runs the
        autotuner (times the candidate sigmas on this matrix, and the
        column-sliced form for large scattered x). xcd_slices=True forces the
        column-sliced form for spmv. plain_columns=True leaves out the
        column codes (3 or 4 bits per element) that stencil-like matrices
        otherwise get."""
        if isinstance(csr, Csr):
            rows, cols = csr.dimensions.rows, csr.dimensions.cols
            rp, ci, v = csr.row_start_offsets, csr.col_indices, csr.values
        else:
            rows, cols, rp, ci, v = csr
        rp = np.ascontiguousarray(rp, dtype=np.uint32)
        ci = np.ascontiguousarray(ci, dtype=np.uint32)
        v = np.ascontiguousarray(v, dtype=np.float64)
        if len(rp) != rows + 1:
            raise ValueError("row_ptr must have rows + 1 entries")
        flags = ((N.DIM_CSR5_KEEP_F32 if keep_f32 else 0) | (N.DIM_CSR5_XCD_SLICES if xcd_slices else 0) |
                 (N.DIM_CSR5_PLAIN_COLUMNS if plain_columns else 0))
        opts = N.Csr5Opts(sigma, flags)
        h = C.c_void_p()
        N.check(N.lib().dim_csr5_from_csr(rows, cols, len(v), rp.ctypes.data, ci.ctypes.data, v.ctypes.data,
                                          C.byref(opts), C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_csr_batched(cls, systems, sigma: int = 0, keep_f32: bool = False) -> "Csr5":
        """A block-diagonal batch of square systems, each (rows, row_ptr,
        col_idx, vals) on the host; sigma=0 autotunes on the whole batch."""
        n = len(systems)
        keep = []
        rows = (C.c_uint32 * n)()
        nnz = (C.c_uint64 * n)()
        rps = (C.c_void_p * n)()
        cis = (C.c_void_p * n)()
        vs = (C.c_void_p * n)()
        for i, (r, rp, ci, v) in enumerate(systems):
            rp = np.ascontiguousarray(rp, dtype=np.uint32)
            ci = np.ascontiguousarray(ci, dtype=np.uint32)
            v = np.ascontiguousarray(v, dtype=np.float64)
            keep += [rp, ci, v]
            rows[i], nnz[i] = int(r), len(v)
            rps[i], cis[i], vs[i] = rp.ctypes.data, ci.ctypes.data, v.ctypes.data
        opts = N.Csr5Opts(sigma, N.DIM_CSR5_KEEP_F32 if keep_f32 else 0)
        h = C.c_void_p()
        N.check(N.lib().dim_csr5_from_csr_batched(n, rows, nnz, rps, cis, vs, C.byref(opts), C.byref(h)))
        A = cls(h.value)
        A.system_rows = np.array([rows[i] for i in range(n)], dtype=np.int64)
        A.system_nnz = np.array([nnz[i] for i in range(n)], dtype=np.int64)
        return A

    @property
    def n_systems(self) -> int:
        return int(N.lib().dim_csr5_n_systems(self._h))

    def set_values(self, vals, stream=None):
        """Replaces the stored values (same pattern) from a float64 GPU tensor
        in the CSR order the matrix was built from."""
        _check_tensor(vals, "vals", self.non_zero_count())
        N.check(N.lib().dim_csr5_set_values(self._h, _ptr(vals), _stream_ptr(stream)))

    @classmethod
    def from_csr_device(cls, rows, cols, row_ptr, col_idx, vals, sigma=16, keep_f32=False, stream=None):
        """Same, with the CSR arrays already on the GPU (torch tensors:
        int32/uint32 row_ptr and col_idx, float64 vals); conversion runs on
        the GPU."""
        _check_tensor(row_ptr, "row_ptr", rows + 1, ("int32", "uint32"))
        _check_tensor(vals, "vals", vals.numel())
        _check_tensor(col_idx, "col_idx", vals.numel(), ("int32", "uint32"))
        opts = N.Csr5Opts(sigma, N.DIM_CSR5_KEEP_F32 if keep_f32 else 0)
        h = C.c_void_p()
        N.check(N.lib().dim_csr5_from_csr_device(rows, cols, vals.numel(), _ptr(row_ptr), _ptr(col_idx), _ptr(vals),
                                                 C.byref(opts), _stream_ptr(stream), C.byref(h)))
        return cls(h.value)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                N.lib().dim_csr5_destroy(h)
            except Exception:
                pass
            self._h = None

    # ------------------------------------------------------------ queries --
    def non_zero_count(self) -> int:
        return int(self.info.nnz)

    def tail_partition_start(self) -> int:
        """The GPU format pads its last tile, so no tail partition remains."""
        return int(self.info.tail_partition_start)

    @property
    def tile_count(self) -> int:
        return int(self.info.tile_count)

    def first_row_idx(self) -> int:
        return 0

    def last_row_idx(self) -> int:
        return int(self.info.rows) - 1

    def create_calibrator(self) -> Calibrator:
        return Calibrator()

    def format_arrays(self):
        """The device format copied to host numpy arrays (for tests)."""
        T = self.tile_count
        tile_ptr = np.zeros(T + 1, np.uint32)
        bit_flag = np.zeros(max(T * 64, 1), np.uint32)
        tile_flags = np.zeros(T + 1, np.uint8)
        tdo_ptr = np.zeros(T + 1, np.uint32)
        n = C.c_uint64()
        N.check(N.lib().dim_csr5_copy_format(self._h, None, None, None, None, None, C.byref(n), None))
        tdo = np.zeros(max(n.value, 1), np.uint32)
        empty = np.zeros(max(self.info.n_empty_rows, 1), np.uint32)
        N.check(N.lib().dim_csr5_copy_format(self._h, tile_ptr.ctypes.data, bit_flag.ctypes.data,
                                             tile_flags.ctypes.data, tdo_ptr.ctypes.data, tdo.ctypes.data,
                                             C.byref(n), empty.ctypes.data))
        return dict(tile_ptr=tile_ptr, bit_flag=bit_flag[: T * 64].reshape(T, 64), tile_flags=tile_flags,
                    tile_desc_offset_ptr=tdo_ptr, tile_desc_offset=tdo[: n.value],
                    empty_rows=empty[: self.info.n_empty_rows])

    # ----------------------------------------------------------- products --
    def spmv(self, data, y=None, calibrator=None, stream=None):
        """csr5::spmv: y = A*x. Accepts SpmvData(x, y, calibrator) or (x, y);
        with y omitted a new tensor is returned. x and y are GPU tensors, or
        numpy arrays in host memory (dim_csr5_spmv_host: copied to the device
        and back, y returned once written)."""
        if isinstance(data, SpmvData):
            x, y = data.x, data.y
        else:
            x = data
        if isinstance(x, np.ndarray):
            return self._spmv_host(x, y, stream)
        torch = _torch()
        _check_tensor(x, "x", self.dimensions.cols, ("float64", "float32"), _X_COLS)
        if y is None:
            y = torch.empty(self.dimensions.rows, dtype=x.dtype, device=x.device)
        if x.dtype != y.dtype:
            raise TypeError("x and y must have the same dtype")
        _check_tensor(y, "y", self.dimensions.rows, ("float64", "float32"))
        N.check(N.lib().dim_csr5_spmv(self._h, _dtype_code(x), _ptr(x), x.numel(), _ptr(y), y.numel(),
                                      _stream_ptr(stream)))
        return y

    def _spmv_host(self, x, y, stream):
        _check_host(x, "x", self.dimensions.cols, ("float64", "float32"), _X_COLS)
        if y is None:
            y = np.empty(self.dimensions.rows, dtype=x.dtype)
        if x.dtype != getattr(y, "dtype", None):
            raise TypeError("x and y must have the same dtype")
        _check_host(y, "y", self.dimensions.rows, ("float64", "float32"))
        N.check(N.lib().dim_csr5_spmv_host(self._h, _dtype_code(x), x.ctypes.data, x.size, y.ctypes.data, y.size,
                                           _stream_ptr(stream)))
        return y

    def spmv_timed(self, x, y, stream=None):
        """y = A*x with HIP events around each kernel; returns (tiles_ms, fixup_ms)."""
        _check_tensor(x, "x", self.dimensions.cols, ("float64", "float32"), _X_COLS)
        _check_tensor(y, "y", self.dimensions.rows, (str(x.dtype).replace("torch.", ""),))
        ms = (C.c_float * 2)()
        N.check(N.lib().dim_csr5_spmv_timed(self._h, _dtype_code(x), _ptr(x), x.numel(), _ptr(y), y.numel(),
                                            _stream_ptr(stream), ms))
        return float(ms[0]), float(ms[1])

    def block_rows(self, tile_begin, tile_end):
        first = C.c_uint32()
        n = C.c_uint32()
        N.check(N.lib().dim_csr5_block_rows(self._h, tile_begin, tile_end, C.byref(first), C.byref(n)))
        return first.value, n.value

    def spmv_block(self, x, tile_begin, tile_end, y=None, stream=None):
        """csr5::spmv<partial> on tiles [tile_begin, tile_end): returns
        (first_row, y_block) with this block's contributions."""
        torch = _torch()
        _check_tensor(x, "x", self.dimensions.cols, ("float64", "float32"), _X_COLS)
        first, n = self.block_rows(tile_begin, tile_end)
        if y is None:
            y = torch.empty(n, dtype=x.dtype, device=x.device)
        _check_tensor(y, "y", n, (str(x.dtype).replace("torch.", ""),))
        N.check(N.lib().dim_csr5_spmv_block(self._h, _dtype_code(x), tile_begin, tile_end, _ptr(x), x.numel(),
                                            _ptr(y), y.numel(), _stream_ptr(stream)))
        return first, y

    def spmm(self, X, Y=None, stream=None):
        """Y = A*X for X of shape (cols, k), k in 1..16.

        Column-major X (stride (1, ld), e.g. the transpose of a contiguous
        (k, cols) tensor) runs the column-major product, the faster layout;
        a new Y then has the same layout. Anything else is made row-major."""
        torch = _torch()
        if not isinstance(X, torch.Tensor) or X.dim() != 2 or X.shape[0] != self.dimensions.cols:
            raise ValueError("X must be a tensor of shape (cols, k)")
        if str(X.dtype).replace("torch.", "") not in ("float64", "float32") or not X.is_cuda:
            raise TypeError("X must be a float64 or float32 GPU tensor")
        if Y is not None and (not isinstance(Y, torch.Tensor) or not Y.is_cuda):
            raise TypeError("Y must be a GPU tensor")
        rows, k = self.dimensions.rows, X.shape[1]
        colmajor = k > 1 and X.stride(0) == 1 and X.stride(1) >= max(X.shape[0], 1)
        if not colmajor:
            X = X.contiguous()
        if Y is None:
            Y = (torch.empty((k, rows), dtype=X.dtype, device=X.device).t() if colmajor else
                 torch.empty((rows, k), dtype=X.dtype, device=X.device))
        if Y.shape != (rows, k) or Y.dtype != X.dtype:
            raise ValueError("Y must have shape (rows, k) and the dtype of X")
        if colmajor:
            if Y.stride(0) != 1 or Y.stride(1) < max(rows, 1):
                raise ValueError("Y must be column-major like X")
            N.check(N.lib().dim_csr5_spmm_colmajor(self._h, _dtype_code(X), k, _ptr(X), X.stride(1), _ptr(Y),
                                                   Y.stride(1), _stream_ptr(stream)))
        else:
            if not Y.is_contiguous():
                raise ValueError("Y must be row-major (contiguous) like X")
            N.check(N.lib().dim_csr5_spmm(self._h, _dtype_code(X), k, _ptr(X), _ptr(Y), _stream_ptr(stream)))
        return Y

    def spmv_transpose(self, x, y=None, stream=None):
        """y = A^T x."""
        torch = _torch()
        _check_tensor(x, "x", self.dimensions.rows, ("float64", "float32"))
        if y is None:
            y = torch.empty(self.dimensions.cols, dtype=x.dtype, device=x.device)
        _check_tensor(y, "y", self.dimensions.cols, (str(x.dtype).replace("torch.", ""),))
        N.check(N.lib().dim_csr5_spmv_transpose(self._h, _dtype_code(x), _ptr(x), _ptr(y), _stream_ptr(stream)))
        return y

    def sampled_outer(self, u, v, scale=1.0, out=None, stream=None):
        """out[e] = scale * u[row(e)] * v[col(e)] over the stored elements, CSR order."""
        torch = _torch()
        _check_tensor(u, "u", self.dimensions.rows)
        _check_tensor(v, "v", self.dimensions.cols)
        if out is None:
            out = torch.empty(self.non_zero_count(), dtype=torch.float64, device=u.device)
        _check_tensor(out, "out", self.non_zero_count())
        N.check(N.lib().dim_csr5_sampled_outer(self._h, float(scale), _ptr(u), _ptr(v), _ptr(out),
                                               _stream_ptr(stream)))
        return out


@dataclass
class CgResult:
    iterations: int
    converged: bool
    relres: float
    relres_true: float
    refinements: int = 0
    seconds: float = 0.0
    history: Optional[np.ndarray] = field(default=None, repr=False)


def _cg_opts(rtol, max_iters, precision, check_every, refine_delta, hist, flags=0):
    prec = {"f64": N.DIM_CG_F64, "fp64": N.DIM_CG_F64, "f32-refined": N.DIM_CG_F32_REFINED,
            "fp32-ir": N.DIM_CG_F32_REFINED}[precision]
    return N.CgOpts(rtol, max_iters, prec, check_every, refine_delta,
                    hist.ctypes.data if hist is not None else None, flags)


def _result(r, hist=None):
    h = None if hist is None else hist[: r.iterations + 1].copy()
    return CgResult(r.iterations, bool(r.converged), r.relres, r.relres_true, r.refinements, r.seconds, h)


def conjugate_gradient(A: Csr5, b, x=None, rtol=1e-8, max_iters=5000, precision="f64", check_every=0,
                       refine_delta=0.0, history=False, stream=None):
    """conjugate_gradient(A, b, x) (thesis.tex:2133-2160): solves A x = b.

    b, x: float64 GPU tensors of length rows (x holds x0 and receives the
    solution; zeros when omitted), or numpy arrays in host memory (dim_cg_host:
    the vectors are copied to the device and the solution back). precision:
    "f64" or "f32-refined" (FP32 iteration inside FP64 residual replacement;
    needs keep_f32). Returns (x, CgResult); a breakdown (non-SPD matrix)
    raises DimError."""
    n = A.dimensions.rows
    if isinstance(b, np.ndarray):
        _check_host(b, "b", n)
        x = np.zeros(n) if x is None else x
        _check_host(x, "x", n)
        hist = np.zeros(max_iters + 1) if history else None
        opts = _cg_opts(rtol, max_iters, precision, check_every, refine_delta, hist)
        res = N.CgResult()
        N.check(N.lib().dim_cg_host(A._h, b.ctypes.data, x.ctypes.data, C.byref(opts), C.byref(res),
                                    _stream_ptr(stream)), allow=(N.DIM_NOT_CONVERGED,))
        return x, _result(res, hist)
    torch = _torch()
    _check_tensor(b, "b", n)
    if x is None:
        x = torch.zeros_like(b)
    _check_tensor(x, "x", n)
    hist = np.zeros(max_iters + 1) if history else None
    opts = _cg_opts(rtol, max_iters, precision, check_every, refine_delta, hist)
    res = N.CgResult()
    N.check(N.lib().dim_cg(A._h, _ptr(b), _ptr(x), C.byref(opts), C.byref(res), _stream_ptr(stream)),
            allow=(N.DIM_NOT_CONVERGED,))
    return x, _result(res, hist)


def conjugate_gradient_multi(A: Csr5, B, X=None, rtol=1e-8, max_iters=5000, precision="f64", check_every=0,
                             refine_delta=0.0, stream=None):
    """k independent right-hand sides (B: rows x k) iterated with one SpMM per step."""
    torch = _torch()
    if not isinstance(B, torch.Tensor) or B.dim() != 2:
        raise ValueError("B must be a tensor of shape (rows, k)")
    n, k = A.dimensions.rows, B.shape[1]
    if B.shape[0] != n or not 1 <= k <= 16:
        raise ValueError(f"B must have shape ({n}, k) with k in 1..16, got {tuple(B.shape)}")
    _check_tensor(B, "B", n * k)
    if X is None:
        X = torch.zeros_like(B)
    if not isinstance(X, torch.Tensor) or tuple(X.shape) != tuple(B.shape):
        raise ValueError("X must have the shape of B")
    _check_tensor(X, "X", n * k)
    opts = _cg_opts(rtol, max_iters, precision, check_every, refine_delta, None)
    res = (N.CgResult * k)()
    N.check(N.lib().dim_cg_multi(A._h, k, _ptr(B), _ptr(X), C.byref(opts), res, _stream_ptr(stream)),
            allow=(N.DIM_NOT_CONVERGED,))
    return X, [_result(r) for r in res]


def conjugate_gradient_batched(A: Csr5, B, X=None, rtol=1e-8, max_iters=5000, check_every=0, transpose=False,
                               stream=None, persistent=False):
    """Solves every system of a batch (A from Csr5.from_csr_batched; B, X:
    float64 GPU tensors holding the systems' vectors back to back). With
    transpose=True the products use A^T (the adjoint solve). persistent=True
    runs the iteration loop in one cooperative kernel with grid barriers
    (not with transpose). Returns (X, [CgResult per system])."""
    torch = _torch()
    n = A.dimensions.rows  # the systems' rows together
    _check_tensor(B, "B", n)
    if X is None:
        X = torch.zeros_like(B)
    _check_tensor(X, "X", n)
    flags = (N.DIM_CG_TRANSPOSE if transpose else 0) | (N.DIM_CG_PERSISTENT if persistent else 0)
    opts = _cg_opts(rtol, max_iters, "f64", check_every, 0.0, None, flags)
    S = A.n_systems
    res = (N.CgResult * S)()
    N.check(N.lib().dim_cg_batched(A._h, _ptr(B), _ptr(X), C.byref(opts), res, _stream_ptr(stream)),
            allow=(N.DIM_NOT_CONVERGED,))
    return X, [_result(r) for r in res]
