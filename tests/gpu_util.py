"""Shared helpers of the GPU parity tests."""
import numpy as np

U = np.finfo(np.float64).eps / 2


def to_dev(a, dtype=None):
    import torch

    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


def assert_spmv_close(y, y_ref, row_ptr, col_idx, vals, x, rel=1e-13):
    """FP64 parity gate: normwise ||dy|| / ||y|| <= rel (summation order only),
    plus a componentwise rounding bound 2*(len_i + 2)*u*sum_j |a_ij x_j| as a
    bug detector (both sides may sum a row in any order)."""
    y = np.asarray(y, dtype=np.float64)
    d = y - y_ref
    ny = np.linalg.norm(y_ref)
    assert np.linalg.norm(d) <= rel * max(ny, 1e-300) + 1e-300, (np.linalg.norm(d), ny)
    lens = np.diff(row_ptr.astype(np.int64))
    rows = np.repeat(np.arange(len(lens)), lens)
    absum = np.bincount(rows, weights=np.abs(vals * x[col_idx]), minlength=len(lens))
    bound = 2.0 * (lens + 2) * U * absum
    bad = np.abs(d) > bound
    assert not bad.any(), f"{bad.sum()} rows over the rounding bound, first {np.nonzero(bad)[0][:5]}"
