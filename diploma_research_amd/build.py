"""Build recipe for libdim_gpu.so (gfx950), libdim_h5.so and the dim CLI.

    python -m diploma_research_amd.build          # libraries + CLI
    python -m diploma_research_amd.build --lib    # libdim_gpu.so only

Every HIP/C++ source is compiled by hipcc for gfx950 in parallel into
_build/, then linked in-tree as diploma_research_amd/libdim_gpu.so.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import re
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(PKG, "libdim_gpu.so")
CLI = os.path.join(PKG, "dim")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = ["format.cpp", "capi.cpp", "spmv.hip", "spmv_k2.hip", "spmv_k4.hip", "spmv_k8.hip", "spmv_k16.hip",
           "cg.hip", "batched.hip", "convert.hip", "probe.hip"]

# Packed FP32 instructions are switched off for all device code (some of their
# forms have given wrong results on these machines).
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-Wall", "-Wno-unused-command-line-argument", "-I", CSRC,
             "-I", os.path.join(ROOT, "include")]

# HDF5 (C API) for the CLI: the image ships it under /opt/conda.
HDF5_ROOT = os.environ.get("DIM_HDF5_ROOT", "/opt/conda")


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l and "target-feature" not in l)
    return p.returncode, out


_INCLUDE = re.compile(r'^\s*#\s*include\s+"([^"]+)"', re.M)


def _deps(path, seen=None):
    """`path` and every file it includes with #include "..." (resolved against
    the including file's folder, then csrc/ and include/), recursively."""
    seen = set() if seen is None else seen
    if path in seen or not os.path.exists(path):
        return seen
    seen.add(path)
    with open(path) as f:
        text = f.read()
    for inc in _INCLUDE.findall(text):
        for d in (os.path.dirname(path), CSRC, os.path.join(ROOT, "include")):
            cand = os.path.normpath(os.path.join(d, inc))
            if os.path.exists(cand):
                _deps(cand, seen)
                break
    return seen


def _compile(src, force=False):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    path = os.path.join(CSRC, src)
    deps = sorted(_deps(path))
    if not force and os.path.exists(obj) and os.path.getmtime(obj) >= max(os.path.getmtime(d) for d in deps):
        return src, 0, "(up to date)"
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    rc, out = _run([HIPCC, *HIP_FLAGS, *lang, "-c", path, "-o", obj])
    return src, rc, out


def build_lib(jobs=None, verbose=False, force=False):
    """Compiles every source (force) or the out-of-date ones, then links."""
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    objs = []
    failed = []
    with cf.ThreadPoolExecutor(jobs) as ex:
        for src, rc, out in ex.map(lambda f: _compile(f, force), SOURCES):
            objs.append(os.path.join(BUILD, os.path.basename(src) + ".o"))
            if rc != 0:
                failed.append((src, out))
            elif verbose and out.strip():
                print(f"[{src}] {out}")
    if failed:
        for src, out in failed:
            print(f"--- compile failed: {src}\n{out}", file=sys.stderr)
        raise RuntimeError("libdim_gpu build failed: " + ", ".join(s for s, _ in failed))
    tmp = LIB + ".tmp"
    rc, out = _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"])
    if rc != 0:
        raise RuntimeError("libdim_gpu link failed:\n" + out)
    os.replace(tmp, LIB)
    return LIB


H5LIB = os.path.join(PKG, "libdim_h5.so")
CXX = os.environ.get("CXX", "g++")


def build_h5():
    """libdim_h5.so (include/dim_h5.h): HDF5 matrix/vector files (host code),
    and their loading straight to GPU matrices through libdim_gpu.so."""
    inc = os.path.join(HDF5_ROOT, "include")
    libdir = os.path.join(HDF5_ROOT, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        print(f"HDF5 headers not found under {HDF5_ROOT}; skipping libdim_h5 and the dim CLI", file=sys.stderr)
        return None
    tmp = H5LIB + ".tmp"
    # links libdim_gpu.so (in this folder) for the load-to-GPU calls
    cmd = [CXX, "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I", inc, "-I", os.path.join(ROOT, "include"),
           os.path.join(CSRC, "h5io.cpp"), "-o", tmp, "-L", libdir, "-lhdf5", "-L", PKG, "-ldim_gpu",
           f"-Wl,-rpath,{libdir}", "-Wl,-rpath,$ORIGIN"]
    rc, out = _run(cmd)
    if rc != 0:
        raise RuntimeError("libdim_h5 build failed:\n" + out)
    os.replace(tmp, H5LIB)
    return H5LIB


def build_cli():
    """The dim command-line tool (links libdim_gpu and libdim_h5 in-tree)."""
    src = os.path.join(PKG, "cli", "dim_cli.cpp")
    if not os.path.exists(src) or build_h5() is None:
        return None
    tmp = CLI + ".tmp"
    cmd = [HIPCC, "-O2", "-std=c++20", "-Wall", f"--offload-arch={ARCH}", "-I", os.path.join(ROOT, "include"), src,
           "-o", tmp, "-L", PKG, "-ldim_gpu", "-ldim_h5", "-Wl,-rpath,$ORIGIN"]
    rc, out = _run(cmd)
    if rc != 0:
        raise RuntimeError("dim CLI build failed:\n" + out)
    os.replace(tmp, CLI)
    return CLI


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true", help="library only")
    ap.add_argument("--clean", action="store_true", help="recompile every source")
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args(argv)
    print(build_lib(a.j, a.v, force=a.clean))
    if not a.lib:
        print(build_cli())


if __name__ == "__main__":
    main()
