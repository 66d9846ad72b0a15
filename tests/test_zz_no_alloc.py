"""dim_cg_noalloc, the allocation-free CG entry point, against dim_cg through
the C ABI: bit-identical solutions and the same iteration counts on two
calls in a row (the second on a smaller system, so the static scratch holds
the first call's larger contents), in FP64 and FP32-refined mode, and its
limits rejected before any launch."""
import ctypes as C

import numpy as np
import pytest

from gpu_util import to_dev
from diploma_research_amd import Csr5, matgen
from diploma_research_amd import native as N

pytestmark = pytest.mark.gpu


def _solve(entry, A, b, precision, rtol=1e-10, max_iters=5000):
    import torch

    x = torch.zeros_like(b)
    opts = N.CgOpts(rtol, max_iters, precision, 0, 0.0, None, 0)
    res = N.CgResult()
    st = getattr(N.lib(), entry)(A._h, C.c_void_p(b.data_ptr()), C.c_void_p(x.data_ptr()), C.byref(opts),
                                 C.byref(res), C.c_void_p(torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    return st, x, res


@pytest.mark.parametrize("precision", [N.DIM_CG_F64, N.DIM_CG_F32_REFINED])
def test_noalloc_equals_dim_cg_on_two_calls(cuda, precision):
    for side, seed in ((160, 1), (72, 2)):  # the second call is smaller
        rp, ci, v = matgen.varcoef((side, side), 0.5, seed=seed)
        n = side * side
        A = Csr5.from_csr((n, n, rp, ci, v), sigma=8, keep_f32=True)
        b = to_dev(np.random.default_rng(seed).uniform(-1, 1, n))
        print(f"noalloc case: side {side}, precision {precision}", flush=True)
        st0, x0, r0 = _solve("dim_cg", A, b, precision)
        st1, x1, r1 = _solve("dim_cg_noalloc", A, b, precision)
        assert st0 == N.DIM_SUCCESS and st1 == N.DIM_SUCCESS, (st0, st1, N.lib().dim_last_error())
        assert r1.converged and r1.iterations == r0.iterations, (r0.iterations, r1.iterations)
        assert r1.refinements == r0.refinements
        assert r1.relres_true == r0.relres_true and r1.relres_true <= 1.01e-10
        assert bool((x0 == x1).all()), float((x0 - x1).abs().max())


def test_noalloc_rejects_history_and_oversized_systems(cuda):
    import torch

    rp, ci, v = matgen.poisson2d(32)
    A = Csr5.from_csr((1024, 1024, rp, ci, v), sigma=8)
    b = torch.ones(1024, dtype=torch.float64, device="cuda")
    x = torch.zeros_like(b)
    hist = np.zeros(11)
    opts = N.CgOpts(1e-8, 10, N.DIM_CG_F64, 0, 0.0, hist.ctypes.data, 0)
    res = N.CgResult()
    st = N.lib().dim_cg_noalloc(A._h, C.c_void_p(b.data_ptr()), C.c_void_p(x.data_ptr()), C.byref(opts),
                                C.byref(res), None)
    assert st == N.DIM_INVALID_ARGUMENT and b"history" in N.lib().dim_last_error()
    n = N.DIM_CG_NOALLOC_MAX_ROWS + 1
    D = Csr5.from_csr((n, n, np.arange(n + 1, dtype=np.uint32), np.arange(n, dtype=np.uint32), np.full(n, 2.0)),
                      sigma=16)
    bd = torch.ones(n, dtype=torch.float64, device="cuda")
    xd = torch.zeros_like(bd)
    opts = N.CgOpts(1e-8, 10, N.DIM_CG_F64, 0, 0.0, None, 0)
    st = N.lib().dim_cg_noalloc(D._h, C.c_void_p(bd.data_ptr()), C.c_void_p(xd.data_ptr()), C.byref(opts),
                                C.byref(res), None)
    assert st == N.DIM_INVALID_ARGUMENT and b"at most" in N.lib().dim_last_error()
    assert float(xd.abs().max()) == 0.0  # nothing ran
