"""The dim command-line tool (diploma_research_amd/dim): store-matrix,
compare-results and info on the CPU; spmv and cg on the GPU against the
oracle on the same HDF5 file (the 2-D 5-point Laplacian on a 256 x 256 grid,
x ~ U[-1,1])."""
import os
import subprocess

import numpy as np
import pytest

from diploma_research_amd import h5, matgen
from diploma_research_amd.csr5 import Csr
from diploma_research_amd.mmio import read_coo
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
DIM = os.path.join(ROOT, "diploma_research_amd", "dim")

pytestmark = pytest.mark.skipif(not os.path.exists(DIM), reason="dim CLI not built")


def run(*args, check=True):
    p = subprocess.run([DIM, *map(str, args)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=300)
    if check and p.returncode != 0:
        raise AssertionError(f"dim {' '.join(map(str, args))} -> {p.returncode}\n{p.stdout}")
    return p


def _mm_csr():
    coo = read_coo(os.path.join(GOLDEN, "mm.mtx"))
    c = Csr.from_coo(coo)
    return coo.dimensions.rows, coo.dimensions.cols, c.row_start_offsets, c.col_indices, c.values


def test_store_matrix_from_matrix_market(tmp_path):
    rows, cols, rp, ci, v = _mm_csr()
    for fmt in ("csr", "csr5"):
        out = tmp_path / f"mm.{fmt}.h5"
        run("store-matrix", os.path.join(GOLDEN, "mm.mtx"), "-o", out, "-f", fmt)
        assert h5.group_kind(out, "A") == (h5.CSR5 if fmt == "csr5" else h5.MATLAB_CSR)
        r = h5.read_csr(out, "A")
        assert (r[0], r[1]) == (rows, cols)
        np.testing.assert_array_equal(r[2], rp)
        np.testing.assert_array_equal(r[3], ci)
        np.testing.assert_array_equal(r[4], v)
    # the CSR5 group equals the reference's own fixture, array for array
    if O.ref_h5_available():
        a = O.ref_h5_load_csr5(tmp_path / "mm.csr5.h5")
        b = O.ref_h5_load_csr5(os.path.join(GOLDEN, "mm.csr5.h5"))
        for k in b:
            if isinstance(b[k], np.ndarray):
                np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_store_matrix_default_name_group_config_and_append(tmp_path):
    src = tmp_path / "mm.mtx"
    src.write_text(open(os.path.join(GOLDEN, "mm.mtx")).read())
    cfg = tmp_path / "config.yaml"
    cfg.write_text("csr:\n  values:\n    chunk_size: 100\n    compression_level: 4\n  col_idx:\n    chunk_size: null\n"
                   "csr5:\n  vals:\n    chunk_size: 64\n    compression_level: null\n  tile_desc:\n    chunk_size: 3\n")
    run("store-matrix", src, "-f", "csr5", "-g", "M", "-c", cfg)
    out = tmp_path / "mm.csr5.h5"  # every extension replaced by .csr5.h5
    assert out.exists()
    rows, cols, rp, ci, v = _mm_csr()
    np.testing.assert_array_equal(h5.read_csr(out, "M")[4], v)
    # MATLAB CSR input group -> second group of the same file
    csr_file = tmp_path / "in.h5"
    h5.write_matlab_csr(csr_file, "A", rows, cols, rp, ci, v)
    run("store-matrix", csr_file, "--in-group-name", "A", "-o", out, "-g", "N", "-f", "csr5", "--append", "-c", cfg)
    np.testing.assert_array_equal(h5.read_csr(out, "N")[3], ci)
    np.testing.assert_array_equal(h5.read_csr(out, "M")[3], ci)
    p = run("info", out, "-g", "N")
    assert f"{rows} x {cols}, nnz {len(v)}" in p.stdout


def test_compare_results(tmp_path):
    y = np.random.default_rng(1).uniform(-1, 1, 500)
    h5.write_vector(tmp_path / "a.h5", "Y", y)
    h5.write_vector(tmp_path / "b.h5", "Y", y * (1 + 1e-3))
    h5.write_vector(tmp_path / "b.h5", "grp/Z", y * (1 + 0.05))
    assert run("compare-results", tmp_path / "a.h5", tmp_path / "b.h5").returncode == 0  # within 1%
    assert run("compare-results", tmp_path / "a.h5", tmp_path / "b.h5", "--rtol", "1e-4", check=False).returncode == 1
    p = run("compare-results", tmp_path / "a.h5", tmp_path / "b.h5", "--rhs-group", "grp", "--rhs-dataset", "Z",
            check=False)
    assert p.returncode == 1 and "not equal" in p.stdout
    assert run("compare-results", tmp_path / "a.h5", tmp_path / "b.h5", "--normwise", "--rtol", "2e-3").returncode == 0


def test_bad_input_fails_loudly(tmp_path):
    p = run("store-matrix", tmp_path / "none.mtx", check=False)
    assert p.returncode != 0 and "error" in p.stdout
    p = run("info", os.path.join(GOLDEN, "mm.csr5.h5"), "-g", "Q", check=False)
    assert p.returncode != 0 and "no group" in p.stdout


# ------------------------------------------------------------------ GPU ---

def _laplacian_file(tmp_path):
    rp, ci, v = matgen.poisson2d(256)
    n = 256 * 256
    x = np.random.default_rng(2024).uniform(-1, 1, n)
    src = tmp_path / "lap.h5"
    h5.write_matlab_csr(src, "A", n, n, rp, ci, v)
    h5.write_vector(src, "X", x)
    return src, n, rp, ci, v, x


@pytest.mark.gpu
def test_cli_spmv_on_gpu_matches_cpu_reference_path(tmp_path, cuda):
    src, n, rp, ci, v, x = _laplacian_file(tmp_path)
    m5 = tmp_path / "lap.csr5.h5"
    run("store-matrix", src, "-o", m5, "-f", "csr5")
    run("store-matrix", src, "-o", m5, "-f", "csr5", "-g", "B", "--append")
    out = tmp_path / "y.h5"
    p = run("spmv", m5, "-n", "20", "--x-file", src, "--x-dataset", "X", "--output-file", out)
    assert "CSR5 SpMV took" in p.stdout
    y = h5.read_vector(out, "Y")
    # the CPU reference path on the same file: the reference's CSR5 product
    # when built here, else the oracle's restatement of it
    rows, cols, rp2, ci2, v2 = h5.read_csr(m5, "A")
    M = O.RefCsr5(rows, cols, rp2, ci2, v2) if O.ref_available() else O.Csr5(rows, cols, rp2, ci2, v2)
    yref = M.spmv(x, nthreads=4)
    assert np.linalg.norm(y - yref) <= 1e-13 * np.linalg.norm(yref)
    h5.write_vector(out, "Yref", yref)
    run("compare-results", out, "--rhs-dataset", "Yref", "--normwise", "--rtol", "1e-13")
    # x = 1 by default (csr5_single_node): row sums
    run("spmv", m5, "-g", "B", "-n", "3", "--output-file", out, "--vector-dataset", "Y1")
    np.testing.assert_allclose(h5.read_vector(out, "Y1"), O.csr_spmv(rp, ci, v, np.ones(n)), rtol=0, atol=1e-14)
    # existing dataset without --overwrite is refused
    assert run("spmv", m5, "-n", "1", "--output-file", out, check=False).returncode != 0
    run("spmv", m5, "-n", "1", "--output-file", out, "--overwrite", "--dtype", "f32", "--x-file", src,
        "--x-dataset", "X")
    assert np.linalg.norm(h5.read_vector(out, "Y") - yref) <= 1e-6 * np.linalg.norm(yref)  # FP32 product


@pytest.mark.gpu
def test_cli_cg_on_gpu(tmp_path, cuda):
    src, n, rp, ci, v, x = _laplacian_file(tmp_path)
    m5 = tmp_path / "lap.csr5.h5"
    run("store-matrix", src, "-o", m5, "-f", "csr5")
    out = tmp_path / "x.h5"
    stats = tmp_path / "stats.json"
    p = run("cg", m5, "--rtol", "1e-10", "--num-iters", "5000", "--output-file", out, "--stats", stats)
    assert "converged" in p.stdout
    import json

    js = json.load(open(stats))
    oc = O.Csr5(n, n, rp, ci, v).cg(np.ones(n), rtol=1e-10, maxit=5000)
    assert abs(js["cg"]["num_iters"] - oc["iterations"]) <= 2
    assert js["cg"]["relres_true"] <= 1.01e-10
    np.testing.assert_allclose(h5.read_vector(out, "x"), oc["x"], rtol=1e-7, atol=1e-9)
    p = run("cg", m5, "--rtol", "1e-10", "--precision", "f32-refined", "--stats", stats)
    assert json.load(open(stats))["cg"]["relres_true"] <= 1.01e-10
