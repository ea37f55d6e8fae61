"""The C ABI from C++, the reference's own language: tests/cpp/binding_check.cpp includes every
public header, compiles with g++ -std=c++17 -Wall -Wextra -Werror, links against
vaikkari_amd/libvaikkari_gpe.so and runs host-side calls only (stencil tables, the reference grid
rule of GPUGrossPitaevskii_BCC.cu:223-227, a VortexState saved and loaded in the format of
VortexState.cpp:49-96).  No GPU is needed."""
import os
import shutil
import subprocess

import pytest

import vaikkari_amd  # noqa: F401  (fails loudly when the library is not built)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "vaikkari_amd")


def test_cpp_caller_compiles_links_and_runs(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = tmp_path / "binding_check"
    cmd = [gxx, "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "binding_check.cpp"), "-L", PKG, "-lvaikkari_gpe",
           f"-Wl,-rpath,{PKG}", "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    r = subprocess.run([str(exe), str(tmp_path / "state.dat")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip() == "dims 15 15 29"
