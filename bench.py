#!/usr/bin/env python3
"""Benchmark of the MI355X CSR5 path.

A step is one FP64 CSR5 SpMV y = A*x over the 3-D 7-point Poisson operator on
a 128^3 grid (n = 2,097,152, nnz = 14,581,760), x ~ U[-1,1] from a seed, with
every input already resident in HBM. The headline value is the algorithmic
bandwidth of that product, (8+4)*nnz + 8*(n_rows + n_cols) bytes per step,
with the 256 MB Infinity Cache defeated by rotating over three copies of the
matrix and vectors (625 MB); --warm adds the rate on one copy. The same JSON
line carries the tile kernel's roofline figures, the reference CSR5
AVX2/OpenMP path timed on this host's cores, FP64 CG on the same operator, the
power-law SpMV, FP64 against FP32-refined multi-RHS CG on the 192^3
variable-coefficient operator, and the batched CG with its adjoint gradient.

    python bench.py [--gpus 1] [--steps 50] [--warmup 5] [--no-extras] [--warm]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def spmv_bytes(nnz, rows, cols, vb=8, ib=4, k=1):
    return (vb + ib) * nnz + k * vb * (rows + cols)


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


def cpu_baseline(rows, cols, rp, ci, v, x):
    """The reference's CSR5 AVX2/OpenMP SpMV (oracle/_ref, built from the
    reference's own header) on this host, 5 warm-ups + 20 timed calls; falls
    back to the oracle's C restatement when the reference build is absent."""
    from oracle import oracle as O

    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
    if O.ref_available():
        M = O.RefCsr5(rows, cols, rp, ci, v)
        kind = "reference"
    else:
        M = O.Csr5(rows, cols, rp, ci, v)
        kind = "port"
    y = np.empty(rows)
    for _ in range(5):
        y = M.spmv(x, nthreads=threads)
    ts = []
    for _ in range(20):
        t0 = time.perf_counter()
        y = M.spmv(x, nthreads=threads)
        ts.append(time.perf_counter() - t0)
    med = float(np.median(ts))
    return dict(value=round(spmv_bytes(len(v), rows, cols) / med / 1e9, 2), unit="GB/s", cores=threads, kind=kind,
                sample=f"full 128^3 Poisson FP64 SpMV, 5 warm-up + 20 timed calls, median {med * 1e3:.3f} ms "
                       f"(y zero-filled and fresh calibrator per call, not timed)"), y


def bench_powerlaw(steps):
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
    del ci
    g = torch.Generator().manual_seed(7)
    for dt, vb in ((torch.float64, 8), (torch.float32, 4)):
        x = (torch.rand(n, generator=g, dtype=torch.float64) * 2 - 1).to(dt).cuda()
        y = torch.empty(n, dtype=dt, device="cuda")
        ms = time_spmv([A], [x], [y], steps, 3)
        gbs = spmv_bytes(nnz, n, n, vb=vb) / ms / 1e6
        key = "fp64" if vb == 8 else "fp32"
        out[key] = {"ms": round(ms, 4), "GB/s": round(gbs, 1), "frac_hbm_peak": round(gbs / HBM_PEAK_GBS, 3)}
    return out


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
    return {"workload": "3-D 7-pt Poisson 128^3, FP64, rtol 1e-8, max 5000", "seconds": round(res.seconds, 5),
            "iterations": res.iterations, "relres_true": res.relres_true, "rel_err_vs_xstar": err,
            "ms_per_iteration": round(1e3 * res.seconds / max(res.iterations, 1), 4)}


def bench_mixed(side=192, k=8, rtol=1e-10):
    """FP64 CG against FP32 CG inside FP64 residual replacement on the 3-D
    variable-coefficient operator (lognormal(0,1) diffusivity, harmonic-mean
    faces), k right-hand sides ~ U[-1,1] solved together (one SpMM per
    iteration), x0 = 0, stop at relative residual rtol (FP64 true residual)."""
    import torch

    from diploma_research_amd import Csr5, conjugate_gradient_multi, matgen

    n = side ** 3
    rp, ci, v = matgen.varcoef((side, side, side), 1.0, seed=4)
    A = Csr5.from_csr((n, n, rp, ci, v), sigma=0, keep_f32=True)
    nnz = len(v)
    del rp, ci, v
    B = torch.from_numpy(np.random.default_rng(5).uniform(-1, 1, (n, k))).cuda()
    out = {"workload": f"3-D variable-coefficient operator {side}^3 (n={n}, nnz={nnz}), lognormal(0,1), "
                       f"k={k} RHS, rtol {rtol:g}", "sigma": A.sigma}
    X = {}
    for prec in ("f64", "f32-refined"):
        Xp, res = conjugate_gradient_multi(A, B, rtol=rtol, max_iters=20000, precision=prec)
        torch.cuda.synchronize()
        X[prec] = Xp
        out[prec] = {"seconds": round(max(r.seconds for r in res), 4),
                     "iterations": [r.iterations for r in res],
                     "relres_true_max": max(r.relres_true for r in res),
                     "converged": all(r.converged for r in res),
                     "refinements": res[0].refinements}
    err = torch.linalg.vector_norm(X["f32-refined"] - X["f64"], dim=0) / torch.linalg.vector_norm(X["f64"], dim=0)
    out["f32-refined"]["max_rel_err_vs_fp64_solution"] = float(err.max())
    out["speedup_f32_refined_vs_f64"] = round(out["f64"]["seconds"] / out["f32-refined"]["seconds"], 3)
    return out


def bench_batched_adjoint(n_sys=32, lo=100, hi=700, n_fd=64, h=1e-6, h4=1e-3, rtol=1e-13):
    """Uneven batch of 2-D variable-coefficient systems (sides uniform in
    [lo, hi], lognormal(0, 0.5)) solved in one batched FP64 CG launch
    sequence; L = 0.5 ||x||^2; dL/db and dL/dvals from the adjoint (batched CG
    with A's own product, A being symmetric, + sampled outer product) through
    torch.autograd; central differences
    on n_fd random stored entries, each a set of re-solves of its system:
    2nd order with step h, and 4th order with step h4. With h = 1e-6 the
    solves' rounding (rtol 1e-15) shows in the quotient at about 1e-6 for
    some entries; with h4 = 1e-3 it stays far below, and the 4th-order
    stencil's truncation error (of order h4^4) too."""
    import torch

    from diploma_research_amd import Csr5, conjugate_gradient, matgen
    from diploma_research_amd.autograd import solve

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

    # central differences, each perturbing a_ij and a_ji together (symmetric
    # systems: the derivative of that parametrisation is g_ij + g_ji)
    starts = np.concatenate([[0], np.cumsum([len(v) for *_, v in systems])])
    picks = rng.choice(len(vals_np), size=n_fd, replace=False)
    singles = {}
    errs, errs4 = [], []
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
        for step in (2 * h4, h4, -h4, -2 * h4, h, -h):
            vv = v.copy()
            for m in mirror:
                vv[m] += step
            M.set_values(torch.from_numpy(vv).cuda())
            xs[step], _ = conjugate_gradient(M, bd, rtol=1e-15, max_iters=50000)

        def diff(a, b):  # L(a) - L(b) without cancelling 0.5||x||^2 terms
            return 0.5 * float(torch.dot(xs[a] - xs[b], xs[a] + xs[b]))

        fd2 = diff(h, -h) / (2 * h)
        fd4 = (8.0 * diff(h4, -h4) - diff(2 * h4, -2 * h4)) / (12.0 * h4)
        an = float(sum(gv[starts[s] + m] for m in mirror))
        errs.append(abs(fd2 - an) / max(abs(an), 1e-300))
        errs4.append(abs(fd4 - an) / max(abs(an), 1e-300))
    t4 = time.perf_counter()
    errs, errs4 = np.array(errs), np.array(errs4)
    return {"workload": f"{n_sys} 2-D variable-coefficient systems, sides {int(sides.min())}..{int(sides.max())} "
                        f"(n={int(sum(n for n, *_ in systems))}, nnz={len(vals_np)}), lognormal(0,0.5), FP64, "
                        f"rtol {rtol:g}",
            "sigma": A.sigma, "forward_seconds": round(t1 - t0, 4), "backward_seconds": round(t2 - t1, 4),
            "fd": {"entries": n_fd, "h": h, "stencil": "central, 2nd order", "seconds": round(t4 - t3, 2),
                   "rel_err_median": float(np.median(errs)), "rel_err_max": float(errs.max()),
                   "within_1e-6": int((errs <= 1e-6).sum())},
            "fd4": {"entries": n_fd, "h": h4, "stencil": "central, 4th order",
                    "rel_err_median": float(np.median(errs4)), "rel_err_max": float(errs4.max()),
                    "within_1e-6": int((errs4 <= 1e-6).sum())}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-extras", action="store_true", help="SpMV headline only")
    ap.add_argument("--extras", default="cg,powerlaw,mixed,batched",
                    help="comma-separated extras to run after the headline (cg, powerlaw, mixed, batched)")
    ap.add_argument("--warm", action="store_true",
                    help="also time the product on one matrix copy over and over (Infinity Cache warm)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 256 MiB device-to-device copy (3 times), the known-size reference for "
                         "scaling rocprofv3 FETCH_SIZE / WRITE_SIZE counters")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the CSR5 path runs on a single MI355X here "
              "(multi-GPU sharding is a design note, not a measured configuration)", file=sys.stderr)
        sys.exit(2)

    import torch

    from diploma_research_amd import Csr5, matgen

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
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
    gbs_warm = None
    if a.warm:
        gbs_warm = round(bytes_step / time_spmv(mats[:1], xs[:1], ys[:1], a.steps, a.warmup) / 1e6, 2)


    # HBM bytes per launch of this kernel from the committed rocprofv3
    # FETCH_SIZE / WRITE_SIZE runs (scaled by a copy of known size)
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "spmv_tiles_f64_poisson128_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("by_sigma", {}).get(str(sigma), {}).get("hbm_bytes_per_launch")

    if a.copy_calibration:
        src = torch.ones(256 * 1024 * 1024 // 8, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    log("CPU baseline (reference CSR5 on host cores)")
    cpu, y_cpu = cpu_baseline(n, n, rp, ci, v, x_host)
    # parity of the timed product against the CPU baseline's own result
    y = ys[0].cpu().numpy()
    parity = float(np.linalg.norm(y - y_cpu) / np.linalg.norm(y_cpu))

    extras = {}
    want = set() if a.no_extras else {w.strip() for w in a.extras.split(",") if w.strip()}
    runs = [("cg", "cg_fp64", "FP64 CG on the 128^3 operator", lambda: bench_cg(mats[0], n, rp, ci, v)),
            ("powerlaw", "powerlaw_spmv", "power-law SpMV (n = 4M)", lambda: bench_powerlaw(max(10, a.steps // 2))),
            ("mixed", "cg_mixed_vs_fp64", "FP64 vs FP32-refined CG, 192^3 variable coefficients, k = 8", bench_mixed),
            ("batched", "batched_adjoint", "batched CG + adjoint, 32 uneven systems", bench_batched_adjoint)]
    for key, name, what, fn in runs:
        if key not in want:
            continue
        log(what)
        try:
            extras[name] = fn()
        except Exception as e:  # report, never hide
            extras[name] = {"error": f"{type(e).__name__}: {e}"}

    line = {
        "metric": "SpMV GB/s (% HBM peak); CG secs to rel-res 1e-10, FP32-refined vs FP64",
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
