#!/usr/bin/env python3
"""Benchmark of the MI355X CSR5 path.

A step is one FP64 CSR5 SpMV y = A*x over the 3-D 7-point Poisson operator on
a 128^3 grid (n = 2,097,152, nnz = 14,581,760), x ~ U[-1,1] from a seed, with
every input already resident in HBM. The headline value is the algorithmic
bandwidth of that product, (8+4)*nnz + 8*(n_rows + n_cols) bytes per step,
with the 256 MB Infinity Cache defeated by rotating over three copies of the
matrix and vectors (625 MB). The same JSON line carries:

  * roofline: the tile kernel's algorithmic rate against the 8 TB/s peak, and
    its HBM bytes per launch from rocprofv3 FETCH_SIZE / WRITE_SIZE;
  * the same product with the cache warm (one copy), with 32-bit column
    indices (no column codes), and the measured stream-read ceiling;
  * cpu_baseline: the reference's own CSR5 AVX2/OpenMP SpMV and a textbook CG
    over it, timed in a child process on this host's cores with OpenMP
    threads bound (OMP_PROC_BIND=spread, OMP_PLACES=cores);
  * extras: FP64 CG on the same operator, one rank's share (part 3 of 8) of
    the 128^3 and power-law matrices, the power-law SpMV in FP64 and FP32,
    FP64 against FP32-refined multi-RHS CG on the 192^3 variable-coefficient
    operator, and the batched CG with its adjoint gradient; every FP32 result
    with its error against the FP64 oracle.

    python bench.py [--gpus 1] [--steps 50] [--warmup 5] [--no-extras] [--extras cg,parts,...]
    python bench.py --traffic-sweep [powerlaw|spmm|spmm-all]   # the products rocprofv3 counter runs profile
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "SpMV GB/s (% HBM peak); CG secs to rel-res 1e-10, FP32-refined vs FP64"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def spmv_bytes(nnz, rows, cols, vb=8, ib=4, k=1):
    return (vb + ib) * nnz + k * vb * (rows + cols)


# ----------------------------------------------------------------- GPU legs --

def time_spmv(mats, xs, ys, steps, warmup):
    """ms per step over `steps` products rotating through the copies."""
    import torch

    for i in range(warmup):
        j = i % len(mats)
        mats[j].spmv(xs[j], ys[j])
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for i in range(steps):
        j = i % len(mats)
        mats[j].spmv(xs[j], ys[j])
    e1.record(s)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def kernel_times(mats, xs, ys, reps):
    """Per-launch durations of the tile kernel and the fix-up kernel (HIP
    events recorded around each launch on the stream they run on)."""
    tiles, fix = [], []
    for i in range(reps):
        j = i % len(mats)
        a, b = mats[j].spmv_timed(xs[j], ys[j])
        tiles.append(a)
        fix.append(b)
    return float(np.mean(tiles)), float(np.mean(fix))


def stream_read_ceiling(gib=4, reps=5):
    """GB/s of dim_stream_read over a `gib` GiB buffer (the best of `reps`)."""
    import ctypes as C

    import torch

    from diploma_research_amd import native as N

    n = gib * (1 << 30) // 8
    buf = torch.ones(n, dtype=torch.float64, device="cuda")
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    s = torch.cuda.current_stream()
    best = 0.0
    for _ in range(reps + 1):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(s)
        N.check(N.lib().dim_stream_read(C.c_void_p(buf.data_ptr()), n * 8, C.c_void_p(out.data_ptr()),
                                        C.c_void_p(s.cuda_stream)))
        e1.record(s)
        torch.cuda.synchronize()
        best = max(best, n * 8 / e0.elapsed_time(e1) / 1e6)
    assert float(out) == float(n)  # every word was read
    del buf
    return round(best, 1)


def traffic_for(sigma, codes):
    """HBM bytes per launch of the headline kernel at `sigma` from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs (profiles/)."""
    tfile = os.path.join(ROOT, "profiles", "spmv_tiles_f64_poisson128_traffic.json")
    if not os.path.exists(tfile):
        return None
    with open(tfile) as f:
        t = json.load(f)
    key = f"{sigma}{'c' if codes else ''}"
    entry = t.get("by_kernel", {}).get(key) or t.get("by_sigma", {}).get(str(sigma))
    return entry.get("hbm_bytes_per_launch") if entry else None


def bench_cg(A, rows, rp, ci, v):
    """FP64 CG on the 128^3 operator: b = A x*, x* ~ U[-1,1], x0 = 0, rtol 1e-8."""
    import torch

    from diploma_research_amd import conjugate_gradient

    xs = np.random.default_rng(42).uniform(-1, 1, rows)
    xd = torch.from_numpy(xs).cuda()
    b = torch.empty_like(xd)
    A.spmv(xd, b)
    conjugate_gradient(A, b, rtol=1e-8, max_iters=5000)  # warm-up
    x, res = conjugate_gradient(A, b, rtol=1e-8, max_iters=5000)
    err = float(torch.linalg.vector_norm(x - xd) / torch.linalg.vector_norm(xd))
    nnz = len(v)
    it_bytes = 12 * nnz + 11 * 8 * rows  # per iteration: matrix once + 11 vector passes
    sec = res.seconds
    return {"workload": "3-D 7-pt Poisson 128^3, FP64, b = A x*, rtol 1e-8, max 5000", "seconds": round(sec, 5),
            "iterations": res.iterations, "relres_true": res.relres_true, "rel_err_vs_xstar": err,
            "ms_per_iteration": round(1e3 * sec / max(res.iterations, 1), 4),
            "GB/s_per_iteration": round(it_bytes * res.iterations / max(sec, 1e-12) / 1e9, 1)}


def bench_parts(copies_c2=8, copies_c3=3, steps=40, part=3, count=8):
    """One rank's share: part `part` of `count` of the reference's tile
    partition (ceil(T/count) tiles of 64 elements, libs/dim/io/h5.cpp:95-103)
    of the 128^3 and power-law matrices, converted as a rank converts its
    loaded part (autotuned) and multiplied with the cache defeated by
    rotating copies. Bytes: 12 per element + 8 per row of the share + 8 per
    distinct column its elements read (x read once)."""
    import torch

    from diploma_research_amd import Csr5, h5, matgen

    out = {}
    for name, gen, n, copies in (("poisson128", lambda: matgen.poisson3d(128), 128 ** 3, copies_c2),
                                 ("powerlaw4m", lambda: matgen.powerlaw(4_000_000, seed=2024), 4_000_000, copies_c3)):
        rp, ci, v = gen()
        first, v0, (nr, nc, brp, bci, bv) = h5.partition_block(n, n, rp, ci, v, part, count)
        del rp, ci, v
        ucols = int(np.unique(bci).size)
        A0 = Csr5.from_csr((nr, nc, brp, bci, bv), sigma=0)
        sliced = bool(A0.info.xcd_slices)
        mats = [A0] + [Csr5.from_csr((nr, nc, brp, bci, bv), sigma=A0.sigma, xcd_slices=sliced)
                       for _ in range(copies - 1)]
        x = torch.from_numpy(np.random.default_rng(3).uniform(-1, 1, n)).cuda()
        xs = [x.clone() for _ in mats]
        ys = [torch.empty(nr, dtype=torch.float64, device="cuda") for _ in mats]
        ms = time_spmv(mats, xs, ys, steps, 5)
        byts = 12 * len(bv) + 8 * (nr + ucols)
        out[name] = {"part": f"{part} of {count}", "first_row": first, "rows": nr, "nnz": len(bv),
                     "distinct_cols": ucols, "sigma": A0.sigma, "xcd_slices": A0.info.xcd_slices,
                     "ms": round(ms, 5), "GB/s": round(byts / ms / 1e6, 1),
                     "frac_hbm_peak": round(byts / ms / 1e6 / HBM_PEAK_GBS, 4), "copies_rotated": copies}
        del mats, xs, ys, A0
        torch.cuda.synchronize()
    return out


def bench_powerlaw(steps, keep):
    """Power-law SpMV (n = 4M), FP64 and FP32, autotuned format; keeps both
    results in `keep` for the comparison with the FP64 oracle."""
    import torch

    from diploma_research_amd import Csr5, matgen

    n = 4_000_000
    rp, ci, v = matgen.powerlaw(n, seed=2024)
    nnz = len(v)
    out = {"n": n, "nnz": nnz}
    A = Csr5.from_csr((n, n, rp, ci, v), sigma=0, keep_f32=True)
    out["sigma"] = A.sigma
    # the autotuner's choice of the column-sliced form (8 slices, one per XCD)
    out["xcd_slices"] = A.info.xcd_slices
    out["xcd_sigma"] = A.info.xcd_sigma
    del rp, ci, v
    xh = np.random.default_rng(7).uniform(-1, 1, n)
    for dt, vb in ((torch.float64, 8), (torch.float32, 4)):
        x = torch.from_numpy(xh).to(dt).cuda()
        y = torch.empty(n, dtype=dt, device="cuda")
        ms = time_spmv([A], [x], [y], steps, 3)
        gbs = spmv_bytes(nnz, n, n, vb=vb) / ms / 1e6
        key = "fp64" if vb == 8 else "fp32"
        t_tiles, t_rest = kernel_times([A], [x], [y], 3)
        out[key] = {"ms": round(ms, 4), "GB/s": round(gbs, 1), "frac_hbm_peak": round(gbs / HBM_PEAK_GBS, 3),
                    "tile_kernel_ms": round(t_tiles, 4), "rest_ms": round(t_rest, 4)}
        keep[f"powerlaw_y_{key}"] = y.double().cpu().numpy()
    return out


def bench_mixed(keep, side=192, k=8, rtol=1e-10):
    """FP64 CG against FP32 CG inside FP64 residual replacement on the 3-D
    variable-coefficient operator (lognormal(0,1) diffusivity, harmonic-mean
    faces), k right-hand sides ~ U[-1,1] solved together (one SpMM per
    iteration), x0 = 0, stop at relative residual rtol (FP64 true residual).
    Keeps column 0 of both solutions for the comparison with the oracle."""
    import torch

    from diploma_research_amd import Csr5, conjugate_gradient_multi, matgen

    n = side ** 3
    rp, ci, v = matgen.varcoef((side, side, side), 1.0, seed=4)
    A = Csr5.from_csr((n, n, rp, ci, v), sigma=0, keep_f32=True)
    nnz = len(v)
    del rp, ci, v
    B = torch.from_numpy(np.random.default_rng(5).uniform(-1, 1, (n, k))).cuda()
    out = {"workload": f"3-D variable-coefficient operator {side}^3 (n={n}, nnz={nnz}), lognormal(0,1), "
                       f"k={k} RHS, rtol {rtol:g}", "sigma": A.sigma, "column_codes": bool(A.info.column_codes)}
    X = {}
    for prec in ("f64", "f32-refined"):
        Xp, res = conjugate_gradient_multi(A, B, rtol=rtol, max_iters=20000, precision=prec)
        torch.cuda.synchronize()
        X[prec] = Xp
        sec = max(r.seconds for r in res)
        its = max(r.iterations for r in res)
        vb = 8 if prec == "f64" else 4
        it_bytes = (vb + 4) * nnz + k * 11 * vb * n
        out[prec] = {"seconds": round(sec, 4), "iterations": [r.iterations for r in res],
                     "relres_true_max": max(r.relres_true for r in res),
                     "converged": all(r.converged for r in res), "refinements": res[0].refinements,
                     "GB/s_per_iteration": round(it_bytes * its / sec / 1e9, 1)}
        keep[f"mixed_x0_{prec}"] = Xp[:, 0].cpu().numpy()
    err = torch.linalg.vector_norm(X["f32-refined"] - X["f64"], dim=0) / torch.linalg.vector_norm(X["f64"], dim=0)
    out["f32-refined"]["max_rel_err_vs_fp64_gpu_solution"] = float(err.max())
    out["speedup_f32_refined_vs_f64"] = round(out["f64"]["seconds"] / out["f32-refined"]["seconds"], 3)
    return out


def bench_batched_adjoint(n_sys=32, lo=100, hi=700, n_fd=64, h=1e-6, h4=1e-3, rtol=1e-13):
    """Uneven batch of 2-D variable-coefficient systems (sides uniform in
    [lo, hi], lognormal(0, 0.5)) solved in one batched FP64 CG launch
    sequence; L = 0.5 ||x||^2; dL/db and dL/dvals from the adjoint (batched CG
    with A's own product, A being symmetric, + sampled outer product) through
    torch.autograd; central differences on n_fd random stored entries, each a
    set of re-solves of its system: 2nd order with step h (evaluated through
    x(+h) - x(-h) = -2h A(+h)^-1 E x(-h), and directly), and 4th order with
    step h4."""
    import torch

    from diploma_research_amd import Csr5, conjugate_gradient, matgen
    from diploma_research_amd.autograd import solve
    from diploma_research_amd.csr5 import conjugate_gradient_batched

    rng = np.random.default_rng(2024)
    sides = rng.integers(lo, hi + 1, size=n_sys)
    systems = []
    for i, s in enumerate(sides):
        rp, ci, v = matgen.varcoef((int(s), int(s)), 0.5, seed=100 + i)
        systems.append((int(s) * int(s), rp, ci, v))
    A = Csr5.from_csr_batched(systems, sigma=0)
    vals_np = np.concatenate([v for *_, v in systems])
    bs = [rng.uniform(-1, 1, n) for n, *_ in systems]
    vals = torch.from_numpy(vals_np.copy()).cuda().requires_grad_(True)
    b = torch.from_numpy(np.concatenate(bs)).cuda().requires_grad_(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x = solve(A, vals, b, rtol=rtol, max_iters=50000)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    L = 0.5 * (x * x).sum()
    L.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    gv = vals.grad.cpu().numpy()
    # the forward solve alone in both forms: the captured launch sequence
    # (default) and the persistent kernel with grid barriers
    bd = b.detach()
    forms = {}
    for name, pers in (("launch_sequence", False), ("persistent", True)):
        torch.cuda.synchronize()
        ta = time.perf_counter()
        xf, rf = conjugate_gradient_batched(A, bd, rtol=rtol, max_iters=50000, persistent=pers)
        torch.cuda.synchronize()
        forms[name] = (time.perf_counter() - ta, xf, max(r.iterations for r in rf))

    # central differences, each perturbing a_ij and a_ji together (symmetric
    # systems: the derivative of that parametrisation is g_ij + g_ji)
    starts = np.concatenate([[0], np.cumsum([len(v) for *_, v in systems])])
    picks = rng.choice(len(vals_np), size=n_fd, replace=False)
    singles = {}
    errs, errs_direct, errs4 = [], [], []
    t3 = time.perf_counter()
    for e in picks:
        s = int(np.searchsorted(starts, e, side="right") - 1)
        n, rp, ci, v = systems[s]
        if s not in singles:
            singles[s] = Csr5.from_csr((n, n, rp, ci, v), sigma=A.sigma)
        M = singles[s]
        kk = int(e - starts[s])
        i = int(np.searchsorted(rp.astype(np.int64), kk, side="right") - 1)
        j = int(ci[kk])
        mirror = [kk]
        if i != j:
            seg = slice(int(rp[j]), int(rp[j + 1]))
            mirror.append(int(rp[j]) + int(np.nonzero(ci[seg] == i)[0][0]))
        bd = torch.from_numpy(bs[s]).cuda()
        xs = {}
        for step in (2 * h4, h4, -h4, -2 * h4, -h, h):
            vv = v.copy()
            for m in mirror:
                vv[m] += step
            M.set_values(torch.from_numpy(vv).cuda())
            xs[step], _ = conjugate_gradient(M, bd, rtol=1e-15, max_iters=50000)
        # M holds A(+h) now. x(+h) - x(-h) = -2h A(+h)^-1 E x(-h) with E the
        # unit perturbation (e_i e_j^T + e_j e_i^T, or e_i e_i^T): d solves
        # A(+h) d = E x(-h), so the central difference needs no subtraction
        # of two nearly equal solutions
        w = torch.zeros_like(bd)
        xm = xs[-h]
        if i != j:
            w[i], w[j] = xm[j], xm[i]
        else:
            w[i] = xm[i]
        d, _ = conjugate_gradient(M, w, rtol=1e-15, max_iters=50000)

        def diff(a, b):  # L(a) - L(b) without cancelling 0.5||x||^2 terms
            return 0.5 * float(torch.dot(xs[a] - xs[b], xs[a] + xs[b]))

        # [L(+h) - L(-h)] / 2h = -0.5 d.(x(+h) + x(-h)), exactly
        fd2 = -0.5 * float(torch.dot(d, xs[h] + xs[-h]))
        fd2_direct = diff(h, -h) / (2 * h)
        fd4 = (8.0 * diff(h4, -h4) - diff(2 * h4, -2 * h4)) / (12.0 * h4)
        an = float(sum(gv[starts[s] + m] for m in mirror))
        errs.append(abs(fd2 - an) / max(abs(an), 1e-300))
        errs_direct.append(abs(fd2_direct - an) / max(abs(an), 1e-300))
        errs4.append(abs(fd4 - an) / max(abs(an), 1e-300))
    t4 = time.perf_counter()
    errs, errs4 = np.array(errs + [0.0] * (not errs)), np.array(errs4 + [0.0] * (not errs4))
    errs_direct = np.array(errs_direct + [0.0] * (not errs_direct))
    return {"workload": f"{n_sys} 2-D variable-coefficient systems, sides {int(sides.min())}..{int(sides.max())} "
                        f"(n={int(sum(n for n, *_ in systems))}, nnz={len(vals_np)}), lognormal(0,0.5), FP64, "
                        f"rtol {rtol:g}",
            "sigma": A.sigma, "forward_seconds": round(t1 - t0, 4), "backward_seconds": round(t2 - t1, 4),
            "cg_forms": {k: {"seconds": round(v[0], 4), "max_iterations": v[2]} for k, v in forms.items()},
            "cg_forms_bitwise_equal": bool(torch.equal(forms["launch_sequence"][1], forms["persistent"][1])),
            "fd": {"entries": n_fd, "h": h, "stencil": "central, 2nd order",
                   "evaluation": "[L(+h) - L(-h)] / 2h = -0.5 d.(x(+h) + x(-h)), A(+h) d = E x(-h)",
                   "seconds": round(t4 - t3, 2),
                   "rel_err_median": float(np.median(errs)), "rel_err_max": float(errs.max()),
                   "within_1e-6": int((errs <= 1e-6).sum()),
                   "direct": {"evaluation": "0.5 (x(+h) - x(-h)).(x(+h) + x(-h)) / 2h",
                              "rel_err_median": float(np.median(errs_direct)),
                              "rel_err_max": float(errs_direct.max()),
                              "within_1e-6": int((errs_direct <= 1e-6).sum())}},
            "fd4": {"entries": n_fd, "h": h4, "stencil": "central, 4th order",
                    "rel_err_median": float(np.median(errs4)), "rel_err_max": float(errs4.max()),
                    "within_1e-6": int((errs4 <= 1e-6).sum())}}


# ------------------------------------------------------- CPU baseline leg --

def host_cores():
    """(threads to use, physical cores lscpu reports, CPUs this process may
    run on). The thread count is OMP_NUM_THREADS when the host sets it (the
    GPU box sets its share), else the physical cores, within the affinity."""
    phys = None
    try:
        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], stdout=subprocess.PIPE, text=True, timeout=10).stdout
        pairs = {tuple(l.split(",")) for l in out.splitlines() if l and not l.startswith("#")}
        phys = len(pairs) or None
    except Exception:
        pass
    avail = len(os.sched_getaffinity(0))
    want = int(os.environ.get("OMP_NUM_THREADS", "0")) or (phys or avail)
    return max(1, min(want, avail)), phys, avail


def cpu_leg(outdir, with_c4=True):
    """The CPU baseline leg, run in a child process whose OpenMP threads are
    bound (OMP_PROC_BIND=spread, OMP_PLACES=cores, set before the runtime
    starts). Times the reference's CSR5 SpMV (oracle/_ref, built from its own
    header; the oracle's C restatement when that build is absent) on the
    128^3 operator with the reference's benchmark protocol (5 warm-up + 20
    timed products), and a textbook CG over it to rtol 1e-8. As the checker
    of the GPU results it also writes the FP64 oracle's power-law product
    and, with with_c4, the oracle's CG solution of column 0 of the 192^3
    variable-coefficient system to `outdir`."""
    from diploma_research_amd import matgen
    from oracle import oracle as O

    threads = int(os.environ["OMP_NUM_THREADS"])
    log(f"CPU leg: {threads} threads")
    n = 128 ** 3
    rp, ci, v = matgen.poisson3d(128)
    x = np.random.default_rng(1).uniform(-1, 1, n)
    kind = "reference" if O.ref_available() else "port"
    res = {"threads": threads, "kind": kind}
    if kind == "reference":
        M = O.RefCsr5(n, n, rp, ci, v)
        ts, y = M.time_spmv(x, threads, warmup=5, runs=20)
    else:
        M = O.Csr5(n, n, rp, ci, v)
        ts = []
        for i in range(25):
            t0 = time.perf_counter()
            y = M.spmv(x, nthreads=threads)
            if i >= 5:
                ts.append(time.perf_counter() - t0)
        ts = np.array(ts)
    np.save(os.path.join(outdir, "c2_y.npy"), y)
    med = float(np.median(ts))
    res["spmv"] = {"median_ms": round(med * 1e3, 4), "mean_ms": round(float(np.mean(ts)) * 1e3, 4),
                   "GB/s": round(spmv_bytes(len(v), n, n) / med / 1e9, 2)}
    xs = np.random.default_rng(42).uniform(-1, 1, n)
    b = O.csr_spmv(rp, ci, v, xs)
    if kind == "reference":
        cg = M.cg(b, rtol=1e-8, maxit=5000, nthreads=threads)
    else:
        t0 = time.perf_counter()
        cg = M.cg(b, rtol=1e-8, maxit=5000, nthreads=threads)
        cg["seconds"] = time.perf_counter() - t0
    res["cg"] = {"seconds": round(cg["seconds"], 4), "iterations": cg["iterations"],
                 "relres_true": cg["relres_true"], "converged": cg["converged"]}
    del M
    log("CPU leg: SpMV and CG timed; oracle products for the checks")
    # checker outputs (not timed)
    m = 4_000_000
    rp, ci, v = matgen.powerlaw(m, seed=2024)
    xp = np.random.default_rng(7).uniform(-1, 1, m)
    np.save(os.path.join(outdir, "powerlaw_y.npy"), O.csr_spmv(rp, ci, v, xp))
    lens = np.diff(rp.astype(np.int64))
    np.save(os.path.join(outdir, "powerlaw_absum.npy"),
            np.bincount(np.repeat(np.arange(m), lens), weights=np.abs(v * xp[ci]), minlength=m))
    np.save(os.path.join(outdir, "powerlaw_lens.npy"), lens)
    del rp, ci, v
    if with_c4:
        side = 192
        nc = side ** 3
        rp, ci, v = matgen.varcoef((side, side, side), 1.0, seed=4)
        b0 = np.random.default_rng(5).uniform(-1, 1, (nc, 8))[:, 0].copy()
        log("CPU leg: oracle CG of column 0 of the 192^3 system (about a minute)")
        t0 = time.perf_counter()
        Mc = O.RefCsr5(nc, nc, rp, ci, v) if kind == "reference" else O.Csr5(nc, nc, rp, ci, v)
        oc = Mc.cg(b0, rtol=1e-10, maxit=20000, nthreads=threads)
        np.save(os.path.join(outdir, "c4_x0.npy"), oc["x"])
        res["c4_oracle_cg"] = {"iterations": oc["iterations"], "relres_true": oc["relres_true"],
                               "seconds": round(time.perf_counter() - t0, 2)}
    print(json.dumps(res), flush=True)


def run_cpu_leg(with_c4):
    """Starts the CPU leg in a child process with bound OpenMP threads;
    returns (its JSON, the folder holding its checker outputs)."""
    threads, phys, avail = host_cores()
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="spread", OMP_PLACES="cores")
    outdir = tempfile.mkdtemp(prefix="dim_cpu_leg_")
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-leg", outdir] + ([] if with_c4 else ["--no-oracle-c4"])
    p = subprocess.run(cmd, env=env, stdout=subprocess.PIPE, text=True, timeout=1500)
    if p.returncode != 0:
        raise RuntimeError(f"CPU leg failed with status {p.returncode}")
    res = json.loads(p.stdout.strip().splitlines()[-1])
    res["binding"] = {"OMP_NUM_THREADS": threads, "OMP_PROC_BIND": "spread", "OMP_PLACES": "cores",
                      "lscpu_physical_cores": phys, "cpus_in_affinity": avail}
    return res, outdir


# -------------------------------------------------------- traffic sweep --

def traffic_sweep(which="all", copies=3, reps=6):
    """The products that the rocprofv3 counter runs profile: the 128^3
    Poisson SpMV at every tile height (column-coded and 32-bit-column kernels
    where both exist), `reps` launches each rotating over `copies` copies,
    then the power-law product (autotuned, or with which="powerlaw" alone in
    the column-sliced form at the fixed tile height 4, so that every counter
    pass profiles the same kernel), and a 256 MiB device copy of known size
    that scales the counters. which="spmm" runs only the multi-RHS products
    of the 192^3 variable-coefficient operator at k = 8, 2 and 1 (FP64);
    which="spmm-all" at k = 16, 8, 4, 2, 1 in FP64 and FP32."""
    import torch

    from diploma_research_amd import Csr5, matgen

    if which in ("spmm", "spmm-all"):
        # the multi-RHS product of the 192^3 variable-coefficient operator
        # (column-major blocks, tile height 7, column codes) at k = 8, 2, 1
        # in FP64 (spmm-all: k = 16, 8, 4, 2, 1 in FP64 and FP32)
        side = 192
        n = side ** 3
        rp, ci, v = matgen.varcoef((side, side, side), 1.0, seed=4)
        A = Csr5.from_csr((n, n, rp, ci, v), sigma=7, keep_f32=which == "spmm-all")
        del rp, ci, v
        ks = (16, 8, 4, 2, 1) if which == "spmm-all" else (8, 2, 1)
        dts = (torch.float64, torch.float32) if which == "spmm-all" else (torch.float64,)
        for dt in dts:
            for k in ks:
                X = torch.from_numpy(np.random.default_rng(1).uniform(-1, 1, (k, n))).to("cuda", dt).t()
                Y = torch.empty((k, n), dtype=dt, device="cuda").t()
                for _ in range(reps + 2):
                    if k > 1:
                        A.spmm(X, Y)
                    else:
                        A.spmv(X[:, 0].contiguous(), Y[:, 0])
                torch.cuda.synchronize()
                log(f"spmm {dt} k={k} done")
                del X, Y
        return
    if which == "all":
        n = 128 ** 3
        rp, ci, v = matgen.poisson3d(128)
        x = torch.from_numpy(np.random.default_rng(1).uniform(-1, 1, n)).cuda()
        for sigma in (4, 5, 6, 7, 8, 10, 12, 14, 16, 20, 24, 28, 32):
            for plain in ((False, True) if sigma <= 8 else (False,)):
                mats = [Csr5.from_csr((n, n, rp, ci, v), sigma=sigma, plain_columns=plain) for _ in range(copies)]
                xs = [x.clone() for _ in mats]
                ys = [torch.empty(n, dtype=torch.float64, device="cuda") for _ in mats]
                time_spmv(mats, xs, ys, reps, 0)
                log(f"sigma {sigma} {'plain' if plain else 'coded' if mats[0].info.column_codes else ''} done")
                del mats, xs, ys
        del rp, ci, v
    m = 4_000_000
    rp, ci, v = matgen.powerlaw(m, seed=2024)
    if which == "powerlaw":
        A = Csr5.from_csr((m, m, rp, ci, v), sigma=4, xcd_slices=True)
    else:
        A = Csr5.from_csr((m, m, rp, ci, v), sigma=0, keep_f32=True)
    log(f"power-law: sigma {A.sigma}, slices {A.info.xcd_slices} (stacked sigma {A.info.xcd_sigma})")
    xp = torch.from_numpy(np.random.default_rng(7).uniform(-1, 1, m)).cuda()
    yp = torch.empty(m, dtype=torch.float64, device="cuda")
    for _ in range(reps):
        A.spmv(xp, yp)
    torch.cuda.synchronize()
    src = torch.ones(256 * 1024 * 1024 // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    log("traffic sweep done")


# ------------------------------------------------------------------ main --

def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-extras", action="store_true", help="SpMV headline and CPU baseline only")
    ap.add_argument("--extras", default="cg,parts,powerlaw,mixed,batched",
                    help="comma-separated extras after the headline (cg, parts, powerlaw, mixed, batched)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-oracle-c4", action="store_true",
                    help="skip the oracle's CG solve of the 192^3 system (about a minute of CPU time)")
    ap.add_argument("--fd-entries", type=int, default=64,
                    help="stored entries the batched adjoint is checked on by finite differences (0: none)")
    ap.add_argument("--traffic-sweep", nargs="?", const="all", choices=("all", "powerlaw", "spmm", "spmm-all"),
                    help="run the products the rocprofv3 counter runs profile, and nothing else "
                         "(powerlaw: only the power-law product, column-sliced at tile height 4; "
                         "spmm: the 192^3 multi-RHS products at k = 8, 2, 1 in FP64; spmm-all: k = 16..1, "
                         "FP64 and FP32)")
    ap.add_argument("--cpu-leg", metavar="DIR", help=argparse.SUPPRESS)
    a = ap.parse_args()
    if a.cpu_leg:
        cpu_leg(a.cpu_leg, with_c4=not a.no_oracle_c4)
        return
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the CSR5 path runs on a single MI355X here "
              "(multi-GPU sharding is a design note, not a measured configuration)", file=sys.stderr)
        sys.exit(2)

    import torch

    from diploma_research_amd import Csr5, matgen

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if a.traffic_sweep:
        traffic_sweep(a.traffic_sweep)
        return
    side = 128
    n = side ** 3
    log("generating the 128^3 7-point Poisson operator")
    rp, ci, v = matgen.poisson3d(side)
    nnz = len(v)
    log("autotuning sigma and converting")
    A0 = Csr5.from_csr((n, n, rp, ci, v), sigma=0)
    sigma = A0.sigma
    codes = bool(A0.info.column_codes)  # 4-bit column codes (7 diagonals)
    mats = [A0] + [Csr5.from_csr((n, n, rp, ci, v), sigma=sigma) for _ in range(2)]
    x_host = np.random.default_rng(1).uniform(-1, 1, n)
    xs = [torch.from_numpy(x_host).to(dev) for _ in mats]
    ys = [torch.empty(n, dtype=torch.float64, device=dev) for _ in mats]
    bytes_step = spmv_bytes(nnz, n, n)

    log(f"timing {a.steps} steps (sigma={sigma})")
    torch.cuda.synchronize()
    ms = time_spmv(mats, xs, ys, a.steps, a.warmup)
    gbs = bytes_step / ms / 1e6
    # The look-ahead across chunk boundaries leaves this matrix without a
    # fix-up kernel: each step is then exactly one launch of the tile kernel,
    # and the timed region's HIP events give that kernel's average duration.
    t_tiles, t_fix = kernel_times(mats, xs, ys, 3)
    single_kernel = t_fix == 0.0
    kernel_ms = ms if single_kernel else t_tiles
    achieved = bytes_step / kernel_ms / 1e6
    y_gpu = ys[(a.steps - 1) % len(mats)].cpu().numpy()
    gbs_warm = round(bytes_step / time_spmv(mats[:1], xs[:1], ys[:1], a.steps, a.warmup) / 1e6, 2)
    del mats, xs, ys
    # the same product with 32-bit column indices (no column codes), the form
    # every matrix that is not a stencil uses
    P0 = Csr5.from_csr((n, n, rp, ci, v), sigma=0, plain_columns=True)
    pmats = [P0] + [Csr5.from_csr((n, n, rp, ci, v), sigma=P0.sigma, plain_columns=True) for _ in range(2)]
    pxs = [torch.from_numpy(x_host).to(dev) for _ in pmats]
    pys = [torch.empty(n, dtype=torch.float64, device=dev) for _ in pmats]
    ms_plain = time_spmv(pmats, pxs, pys, a.steps, a.warmup)
    plain = {"sigma": P0.sigma, "ms": round(ms_plain, 5), "GB/s": round(bytes_step / ms_plain / 1e6, 2),
             "frac_hbm_peak": round(bytes_step / ms_plain / 1e6 / HBM_PEAK_GBS, 4),
             "traffic": traffic_for(P0.sigma, False)}
    del pmats, pxs, pys, P0
    ceiling = stream_read_ceiling()
    traffic = traffic_for(sigma, codes)
    # the same product with x and y in pageable host memory
    # (dim_csr5_spmv_host: x over PCIe to the device and y back, every call)
    y_np = np.empty(n)
    for _ in range(2):
        A0.spmv(x_host, y_np)
    t0 = time.perf_counter()
    for _ in range(10):
        A0.spmv(x_host, y_np)
    ms_host = (time.perf_counter() - t0) / 10 * 1e3
    host_ptrs = {"ms": round(ms_host, 4), "GB/s": round(bytes_step / ms_host / 1e6, 1),
                 "note": "x and y in pageable host memory, copied over PCIe on every product; not the headline"}

    extras = {}
    keep = {}
    want = set() if a.no_extras else {w.strip() for w in a.extras.split(",") if w.strip()}
    runs = [("cg", "cg_fp64", "FP64 CG on the 128^3 operator", lambda: bench_cg(A0, n, rp, ci, v)),
            ("parts", "rank_share", "one rank's share (part 3 of 8) of the 128^3 and power-law matrices",
             bench_parts),
            ("powerlaw", "powerlaw_spmv", "power-law SpMV (n = 4M)",
             lambda: bench_powerlaw(max(10, a.steps // 2), keep)),
            ("mixed", "cg_mixed_vs_fp64", "FP64 vs FP32-refined CG, 192^3 variable coefficients, k = 8",
             lambda: bench_mixed(keep)),
            ("batched", "batched_adjoint", "batched CG + adjoint, 32 uneven systems",
             lambda: bench_batched_adjoint(n_fd=a.fd_entries))]
    for key, name, what, fn in runs:
        if key not in want:
            continue
        log(what)
        try:
            extras[name] = fn()
        except Exception as e:  # report, never hide
            extras[name] = {"error": f"{type(e).__name__}: {e}"}
    del A0
    torch.cuda.synchronize()

    cpu = None
    parity = None
    if not a.no_cpu:
        log("CPU leg: the reference CSR5 SpMV and CG on host cores, then the oracle's checker outputs")
        with_c4 = "mixed" in want and not a.no_oracle_c4
        try:
            cres, outdir = run_cpu_leg(with_c4)
            y_cpu = np.load(os.path.join(outdir, "c2_y.npy"))
            parity = float(np.linalg.norm(y_gpu - y_cpu) / np.linalg.norm(y_cpu))
            sp = cres["spmv"]
            cpu = {"value": sp["GB/s"], "unit": "GB/s", "cores": cres["threads"], "kind": cres["kind"],
                   "sample": f"full 128^3 Poisson FP64 SpMV, 5 warm-up + 20 timed products (only csr5::spmv "
                             f"timed; y zero-filled and a fresh calibrator before each), median "
                             f"{sp['median_ms']} ms, mean {sp['mean_ms']} ms",
                   "binding": cres["binding"],
                   "cg": dict(cres["cg"], workload="textbook CG over the reference's SpMV on the 128^3 operator, "
                                                   "b = A x*, rtol 1e-8, max 5000")}
            if "cg_fp64" in extras and "iterations" in extras["cg_fp64"]:
                g = extras["cg_fp64"]
                g["cpu_seconds"] = cres["cg"]["seconds"]
                g["cpu_iterations"] = cres["cg"]["iterations"]
                g["speedup_vs_cpu"] = round(cres["cg"]["seconds"] / g["seconds"], 1)
            pl = extras.get("powerlaw_spmv", {})
            if "powerlaw_y_fp64" in keep:
                ref = np.load(os.path.join(outdir, "powerlaw_y.npy"))
                absum = np.load(os.path.join(outdir, "powerlaw_absum.npy"))
                lens = np.load(os.path.join(outdir, "powerlaw_lens.npy"))
                u32 = np.finfo(np.float32).eps / 2
                for key in ("fp64", "fp32"):
                    yk = keep[f"powerlaw_y_{key}"]
                    entry = {"normwise_rel_err_vs_fp64_oracle": float(np.linalg.norm(yk - ref) / np.linalg.norm(ref))}
                    if key == "fp32":
                        entry["rows_within_fp32_bound"] = bool(
                            (np.abs(yk - ref) <= (lens + 4) * u32 * absum + 1e-30).all())
                    pl[key].update(entry)
            mx = extras.get("cg_mixed_vs_fp64", {})
            if "c4_oracle_cg" in cres and "mixed_x0_f64" in keep:
                xo = np.load(os.path.join(outdir, "c4_x0.npy"))
                mx["oracle_cg_column0"] = cres["c4_oracle_cg"]
                for prec in ("f64", "f32-refined"):
                    mx[prec]["column0_rel_err_vs_fp64_oracle_x"] = float(
                        np.linalg.norm(keep[f"mixed_x0_{prec}"] - xo) / np.linalg.norm(xo))
        except Exception as e:  # report, never hide
            cpu = {"error": f"{type(e).__name__}: {e}"}

    line = {
        "metric": METRIC,
        "value": round(gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "FP64 CSR5 SpMV, 3-D 7-point Poisson 128^3 (n=2097152, nnz=14581760), "
                               "x~U[-1,1]; Infinity Cache defeated by rotating 3 matrix copies",
                   "sigma": sigma, "omega": 64, "bytes_per_step": bytes_step, "column_codes": codes,
                   "frac_hbm_peak": round(gbs / HBM_PEAK_GBS, 4), "warm_GB/s": gbs_warm,
                   "plain_columns": plain, "host_pointers": host_ptrs, "stream_read_ceiling_GB/s": ceiling,
                   "frac_stream_ceiling": round(gbs / ceiling, 4) if ceiling else None,
                   "parity_rel_err_vs_cpu_baseline": parity},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": f"k_csr5_tiles<double,{sigma},1,false,{'true' if codes else 'false'}>",
                     "kernel_ms": round(kernel_ms, 5),
                     "kernels_per_step": 1 if single_kernel else 2,
                     "kernel_ms_source": "HIP events over the timed region, one launch per step" if single_kernel
                     else "HIP events around each tile-kernel launch", "fixup_kernel_ms": round(t_fix, 5)},
        "cpu_baseline": cpu,
    }
    line.update(extras)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
