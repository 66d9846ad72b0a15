"""Host code under AddressSanitizer, UndefinedBehaviorSanitizer and
ThreadSanitizer (CPU only).

The CSR -> CSR5 format builder (csrc/format.cpp, multi-threaded) and the
HDF5 matrix I/O (csrc/h5io.cpp) are plain host C++; the drivers under
tests/native/ compile them with g++ and the sanitizers, run the shapes they
branch on, and check the format's invariants and the partial loads' row
coverage. Any sanitizer report makes the driver exit non-zero
(-fno-sanitize-recover=all, halt_on_error=1)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "diploma_research_amd", "csrc")
INCLUDE = os.path.join(ROOT, "include")
NATIVE = os.path.join(HERE, "native")
HDF5_ROOT = os.environ.get("DIM_HDF5_ROOT", "/opt/conda")
CXX = os.environ.get("CXX", "g++")

pytestmark = pytest.mark.skipif(shutil.which(CXX) is None, reason=f"{CXX} not found")

ENV = dict(os.environ,
           ASAN_OPTIONS="detect_leaks=1:halt_on_error=1:abort_on_error=0",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1",
           TSAN_OPTIONS="halt_on_error=1")


def _build(tmp_path, name, sanitize, sources, extra=()):
    exe = str(tmp_path / name)
    cmd = [CXX, "-O1", "-g", "-fno-omit-frame-pointer", f"-fsanitize={sanitize}", "-fno-sanitize-recover=all",
           "-std=c++17", "-Wall", "-I", CSRC, "-I", INCLUDE, *sources, "-o", exe, "-pthread", *extra]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0 and "cannot find" in r.stderr and "san" in r.stderr:
        pytest.skip(f"{sanitize} runtime not installed")
    assert r.returncode == 0, r.stderr[-4000:]
    return exe


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True, env=ENV, timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-6000:]
    assert "done, 0 failures" in r.stdout, out[-6000:]
    return r.stdout


def test_format_builder_address_and_ub_sanitizers(tmp_path):
    exe = _build(tmp_path, "san_format_asan", "address,undefined",
                 [os.path.join(NATIVE, "san_format.cpp"), os.path.join(CSRC, "format.cpp")])
    _run([exe, "all"])


def test_format_builder_thread_sanitizer(tmp_path):
    """Matrices of about 3M elements at tile heights 4, 7 and 8 (>= 4096
    tiles), so every pass of the builder runs on several threads."""
    exe = _build(tmp_path, "san_format_tsan", "thread",
                 [os.path.join(NATIVE, "san_format.cpp"), os.path.join(CSRC, "format.cpp")])
    _run([exe, "threads"])


@pytest.mark.skipif(not os.path.exists(os.path.join(HDF5_ROOT, "include", "hdf5.h")),
                    reason=f"HDF5 headers not found under {HDF5_ROOT}")
def test_h5_io_address_and_ub_sanitizers(tmp_path):
    lib = os.path.join(HDF5_ROOT, "lib")
    # libstdc++ is linked statically: the HDF5 folder may hold an older one
    # that the run path would otherwise pick up.
    exe = _build(tmp_path, "san_h5", "address,undefined",
                 [os.path.join(NATIVE, "san_h5.cpp"), os.path.join(CSRC, "h5io.cpp")],
                 extra=["-I", os.path.join(HDF5_ROOT, "include"), os.path.join(lib, "libhdf5.so"),
                        f"-Wl,-rpath,{lib}", "-static-libstdc++"])
    out = _run([exe, os.path.join(HERE, "golden"), str(tmp_path)])
    loads = int(out.split(" partial loads")[0].split()[-1])
    assert loads > 100, out
