"""CPU checks of the Python binding's argument validation: every wrapper that
passes a tensor's pointer over the C ABI first checks its dtype, device,
contiguity and length, and raises before any native call (the C ABI takes no
lengths for CG vectors and cannot tell host from device memory)."""
import ctypes as C

import pytest

torch = pytest.importorskip("torch")

from diploma_research_amd import csr5 as M  # noqa: E402
from diploma_research_amd import native as N  # noqa: E402


class _Trap:
    """Stands in for the native library: any call is a test failure."""

    def __getattr__(self, name):
        raise AssertionError(f"native call {name} reached before validation")


@pytest.fixture()
def fake(monkeypatch):
    """A Csr5 of 10 rows and 12 columns with 30 stored entries that was never
    built on a GPU, with the native library replaced by a trap."""
    monkeypatch.setattr(N, "lib", lambda: _Trap())
    A = M.Csr5.__new__(M.Csr5)
    A._h = C.c_void_p(0)
    A.info = N.Csr5Info(rows=10, cols=12, nnz=30, sigma=16, omega=64, tile_count=1)
    A.dimensions = M.Dimensions(10, 12)
    A.sigma = 16
    return A


def _cpu(n, dtype=torch.float64):
    return torch.zeros(n, dtype=dtype)


def test_tensor_check_messages():
    with pytest.raises(TypeError, match="must be a torch tensor"):
        M._check_tensor([1.0], "b", 1)
    with pytest.raises(TypeError, match="dtype float64"):
        M._check_tensor(_cpu(3, torch.float32), "b", 3)
    with pytest.raises(TypeError, match="on the GPU"):
        M._check_tensor(_cpu(3), "b", 3)


def test_cg_wrappers_reject_bad_vectors_before_any_call(fake):
    A = fake
    with pytest.raises(TypeError, match="b must have dtype float64"):
        M.conjugate_gradient(A, _cpu(10, torch.float32))
    with pytest.raises(TypeError, match="b must be on the GPU"):
        M.conjugate_gradient(A, _cpu(10))
    with pytest.raises(ValueError, match="B must have shape"):
        M.conjugate_gradient_multi(A, torch.zeros(10, 17, dtype=torch.float64))
    with pytest.raises(ValueError, match="B must have shape"):
        M.conjugate_gradient_multi(A, torch.zeros(9, 2, dtype=torch.float64))
    with pytest.raises(TypeError, match="B must be on the GPU"):
        M.conjugate_gradient_multi(A, torch.zeros(10, 2, dtype=torch.float64))
    with pytest.raises(TypeError, match="B must be on the GPU"):
        M.conjugate_gradient_batched(A, _cpu(10))
    with pytest.raises(TypeError, match="B must have dtype float64"):
        M.conjugate_gradient_batched(A, _cpu(10, torch.float32))


def test_products_and_values_reject_bad_tensors_before_any_call(fake):
    A = fake
    with pytest.raises(TypeError, match="vals must be on the GPU"):
        A.set_values(_cpu(30))
    with pytest.raises(TypeError, match="vals must have dtype float64"):
        A.set_values(_cpu(30, torch.float32))
    with pytest.raises(TypeError, match="x must be on the GPU"):
        A.spmv(_cpu(12))
    with pytest.raises(TypeError, match="x must have dtype"):
        A.spmv(_cpu(12, torch.int32))
    with pytest.raises(TypeError, match="u must be on the GPU"):
        A.sampled_outer(_cpu(10), _cpu(12))
    with pytest.raises(TypeError, match="x must be on the GPU"):
        A.spmv_transpose(_cpu(10))
    with pytest.raises(TypeError, match="X must be a float64 or float32 GPU tensor"):
        A.spmm(torch.zeros(12, 3, dtype=torch.float64))
    with pytest.raises(ValueError, match="X must be a tensor of shape"):
        A.spmm(torch.zeros(11, 3, dtype=torch.float64))
    with pytest.raises(TypeError, match="row_ptr must be on the GPU"):
        M.Csr5.from_csr_device(10, 12, torch.zeros(11, dtype=torch.int32), torch.zeros(30, dtype=torch.int32),
                               _cpu(30))
