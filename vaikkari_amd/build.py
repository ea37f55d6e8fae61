"""Build recipe of the native library (libvaikkari_gpe.so) for gfx950.

Compiles every HIP/C++ source under vaikkari_amd/csrc with hipcc into one
shared library next to this file.  Run as ``python vaikkari_amd/build.py``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import time

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG_DIR, "csrc")
LIB_NAME = "libvaikkari_gpe.so"
LIB_PATH = os.path.join(PKG_DIR, LIB_NAME)
ARCH = os.environ.get("GPE_OFFLOAD_ARCH", "gfx950")

SOURCES = ["gpe_solver.hip", "gpe_batch.hip", "gpe_output.hip", "gpe_primal.hip", "diag.hip", "host_setup.cpp", "integrate.cpp",
           "gpe_serve.cpp", "vortex_state.cpp", "checkpoint.cpp"]

FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    f"--offload-arch={ARCH}",
    # FP arithmetic is the sequence written (FP64 parity with the oracle).
    "-ffp-contract=off",
    # packed FP32 instructions stay off on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libvaikkari_gpe.so")


def _compile(src: str, obj: str) -> subprocess.CompletedProcess:
    flags = [f for f in FLAGS if f != "-shared"]
    return subprocess.run([hipcc(), *flags, "-x", "hip", "-c", src, "-o", obj], capture_output=True, text=True)


def build(verbose: bool = True, jobs: int = 0) -> str:
    """Compiles every source (in parallel, one hipcc per file) and links the library.
    Every call compiles everything: a built library is never taken as current."""
    from concurrent.futures import ThreadPoolExecutor

    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    objdir = os.path.join(PKG_DIR, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = [os.path.join(objdir, os.path.basename(s) + ".o") for s in srcs]
    t0 = time.time()
    if verbose:
        print(f"[build] hipcc {' '.join(FLAGS)} -c <{len(srcs)} sources>", flush=True)
    jobs = jobs or min(len(srcs), max(1, os.cpu_count() or 1), 16)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(_compile, srcs, objs))
    failed = False
    for src, res in zip(srcs, results):
        if res.returncode != 0:
            sys.stderr.write(f"[build] {os.path.basename(src)} failed:\n" + res.stdout + res.stderr)
            failed = True
        elif verbose and res.stderr.strip():
            # warnings, e.g. the host pass not knowing the packed-fp32 feature
            lines = [ln for ln in res.stderr.splitlines() if "packed-fp32-ops" not in ln]
            if lines:
                sys.stderr.write("\n".join(lines[:40]) + "\n")
    if failed:
        raise RuntimeError("hipcc failed")
    tmp = LIB_PATH + ".tmp"
    res = subprocess.run([hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp],
                         capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"link failed with exit code {res.returncode}")
    os.replace(tmp, LIB_PATH)
    if verbose:
        print(f"[build] {LIB_PATH} ({time.time() - t0:.1f}s)", flush=True)
    return LIB_PATH


if __name__ == "__main__":
    build()
