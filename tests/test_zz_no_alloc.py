"""gpe_observables_noalloc, the allocation-free variant of gpe_observables: through the C ABI,
on fields in a solver's own device buffer and in caller-owned device memory, two calls in a row
(the second on a smaller grid, other precision, streamed potential), each equal to
gpe_observables of the same field bit for bit (same kernels, same fixed-order FP64 reduction).
Device memory for the caller-owned buffers comes from the HIP runtime directly (hipMalloc through
ctypes), as a C caller would provide it."""
import ctypes as C
import math

import numpy as np
import pytest

import vaikkari_amd as va
from vaikkari_amd import _lib

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


H2D, D2H = 1, 2


class DeviceBuffer:
    def __init__(self, nbytes):
        self.p = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.p), int(nbytes)) == 0
        self.nbytes = int(nbytes)

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.nbytes
        assert hip().hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0

    def download(self, dtype, n):
        out = np.empty(n, dtype)
        assert hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, out.nbytes, D2H) == 0
        return out

    @property
    def addr(self):
        return int(self.p.value)

    def free(self):
        if self.p:
            hip().hipFree(self.p)
            self.p = C.c_void_p()


def announce(msg):
    print(f"[gpu-case] {msg}", flush=True)


def reference_obs(s):
    out = np.zeros(4)
    _lib.check(_lib.lib.gpe_observables(s._h, 0, _lib.dptr(out)), "gpe_observables")
    return out


def test_noalloc_observables_two_calls_equal_gpe_observables():
    rng = np.random.default_rng(17)
    out = DeviceBuffer(4 * 8)
    try:
        # call 1: 24 x 20 x 18 cells, FP64, in-kernel trap, the solver's own buffer
        dims, h, g, dt, ratio = (24, 20, 18), 0.4, 600.0, 1e-3, 1.3
        announce(f"noalloc observables call 1 {dims} fp64 trap")
        b = h / math.sqrt(8)
        s = va.Solver(dims, b, dt, g, None, "fp64", "real", "trap", ratio)
        psi = 0.05 * rng.uniform(-1, 1, 2 * va.nodes_of(dims))
        s.set_state(psi, psi)
        va.observables_noalloc(s.cfg, s.device_field(0), out.addr)
        assert hip().hipDeviceSynchronize() == 0
        got = out.download(np.float64, 4)
        np.testing.assert_array_equal(got, reference_obs(s))
        s.close()

        # call 2: a smaller grid, FP32, streamed potential, caller-owned buffers
        dims, h, g, dt, ratio = (10, 9, 7), 0.35, 900.0, 7e-4, 2.1
        announce(f"noalloc observables call 2 {dims} fp32 array")
        b = h / math.sqrt(8)
        origin = va.centered_origin(dims, b)
        V = va.trap_potential(dims, origin, b, ratio) + rng.uniform(0, 1, va.nodes_of(dims))
        s = va.Solver(dims, b, dt, g, origin, "fp32", "real", "array", ratio)
        s.set_potential(V)
        psi = 0.05 * rng.uniform(-1, 1, 2 * va.nodes_of(dims))
        s.set_state(psi, psi)
        L = va.field_layout(dims, "fp32")
        field = DeviceBuffer(L["elems"] * 8)
        field.upload(va.pack_field(dims, psi, "fp32"))
        pot = np.zeros(L["elems"], np.float32)
        k, j, i, l = np.meshgrid(*(np.arange(n) for n in (dims[2], dims[1], dims[0], 12)), indexing="ij")
        pot[L["origin"] + k * L["sz"] + j * L["sy"] + l * L["xp"] + i] = (V * dt).astype(np.float32).reshape(k.shape)
        potb = DeviceBuffer(pot.nbytes)
        potb.upload(pot)
        out.upload(np.full(4, np.nan))
        va.observables_noalloc(s.cfg, field.addr, out.addr, pot_dt=potb.addr)
        assert hip().hipDeviceSynchronize() == 0
        got = out.download(np.float64, 4)
        np.testing.assert_array_equal(got, reference_obs(s))
        s.close()
        field.free()
        potb.free()
    finally:
        out.free()


def test_noalloc_refuses_what_the_scratch_cannot_hold():
    """More blocks than the scratch holds, or a missing argument: GPE_EINVAL before any launch."""
    s = va.Solver((8, 8, 8), 0.1, 1e-4, 10.0, None, "fp64", "real", "trap", 1.0)
    cfg = _lib.GpeConfig()
    C.memmove(C.byref(cfg), C.byref(s.cfg), C.sizeof(cfg))
    cfg.nx, cfg.ny, cfg.nz = 512, 512, 512  # 8 x 128 x 512 blocks > 2^18
    with pytest.raises(_lib.GpeError) as ei:
        va.observables_noalloc(cfg, s.device_field(0), s.device_field(1))
    assert ei.value.status == _lib.EINVAL
    with pytest.raises(_lib.GpeError) as ei:
        va.observables_noalloc(s.cfg, s.device_field(0), 0)
    assert ei.value.status == _lib.EINVAL
    s.close()
