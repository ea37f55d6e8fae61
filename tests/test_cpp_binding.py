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


@pytest.mark.gpu
def test_cpp_drop_in_on_gpu(tmp_path):
    """tests/cpp/drop_in_gpu.cpp: a C++ caller hands a VortexState to the library in memory
    (gpe_vortex_state_from_table) and runs integrateInTime through gpe_integrate_in_time on the
    GPU (2 iterations, FP64).  Its final fields equal the oracle's setup and leapfrog on the same
    state (which the program also saved in the reference's file format) bit for bit."""
    import ctypes as C

    import numpy as np

    from vaikkari_amd import _lib
    from oracle_lib import Oracle

    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    print("[gpu-case] C++ drop-in caller on the GPU", flush=True)
    exe = tmp_path / "drop_in_gpu"
    cmd = [gxx, "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "drop_in_gpu.cpp"), "-L", PKG, "-lvaikkari_gpe",
           f"-Wl,-rpath,{PKG}", "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    prefix = str(tmp_path / "run")
    r = subprocess.run([str(exe), prefix], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    words = r.stdout.split()
    dims, steps = tuple(int(x) for x in words[1:4]), int(words[5])
    ge, go = np.fromfile(prefix + ".even"), np.fromfile(prefix + ".odd")

    orc = Oracle()
    h = orc.o.or_vs_load((prefix + ".vs").encode())
    assert h
    eps = 1e-5 * orc.o.or_vs_search_max(h)
    mr, mz = orc.o.or_vs_search_max_r(h, eps), orc.o.or_vs_search_max_z(h, eps)
    minp, maxp = (-mr, -mr, -mz), (mr, mr, mz)
    d = (C.c_uint32 * 3)()
    p0 = (C.c_double * 3)()
    orc.o.or_ref_dims(0.15, _lib.d3(minp), _lib.d3(maxp), d, p0)
    assert tuple(d) == dims
    n = dims[0] * dims[1] * dims[2] * 12
    ev, od, pot, sc = np.zeros(2 * n), np.zeros(2 * n), np.zeros(n), np.zeros(6)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    assert orc.o.or_ref_setup(h, 0.15, _lib.d3(minp), _lib.d3(maxp), 0.05, 0.7, 54363, dp(ev), dp(od), dp(pot),
                              dp(sc)) == 0
    orc.o.or_vs_free(h)
    assert int(sc[0]) == steps
    oe, oo = orc.leapfrog(ev, od, pot, dims, sc[2], sc[3], sc[4], 2 * steps)
    np.testing.assert_array_equal(ge, oe)
    np.testing.assert_array_equal(go, oo)
