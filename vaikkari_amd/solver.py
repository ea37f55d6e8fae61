"""Host-side mirror of the reference's solver interface over the C ABI.

The reference drives its stepper through ``integrateInTime`` and a
``VortexState`` (koodit/GrossPitaevskii/GPUGrossPitaevskii_BCC.cu:218,
koodit/GrossPitaevskii/VortexState.hpp:16-76).  This module exposes the same
names with the same argument meaning (``VortexState`` with ``load``/``save``,
``setKappa``, ``setG``, ``setDimensions``, ``setRange``, ``initialize``,
``iterateSolution``, ``getPsi``, ``getPhase``, ``getFunction``, the searches and
integrals; ``integrateInTime``) plus ``Solver``, the object form of the stepper
(gpe_create ... gpe_destroy), and the reference's picture / volume-map output
(computed on the device, written in the reference's BMP and MetaImage formats).
Fields are flat float64 arrays of 2*nodes values in the host order
[k][j][i][l], complex interleaved.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import check, dptr, lib

PRECISIONS = {"fp64": _lib.FP64, "fp32": _lib.FP32_BULK, "fp32_bulk": _lib.FP32_BULK}
MODES = {"real": _lib.REAL_TIME, "imag": _lib.IMAG_TIME}
POTENTIALS = {"array": _lib.POT_ARRAY, "trap": _lib.POT_TRAP}
KERNELS = {"auto": _lib.KERNEL_AUTO, "gather": _lib.KERNEL_GATHER, "zmarch": _lib.KERNEL_ZMARCH,
           "fused": _lib.KERNEL_FUSED}
KERNEL_NAMES = {v: k for k, v in KERNELS.items()}

PIx2 = 6.2831853071795864770
HODGE = 1.5


def nodes_of(dims: Sequence[int]) -> int:
    return int(dims[0]) * int(dims[1]) * int(dims[2]) * _lib.GPE_CLASSES


def node_measure(block_scale: float) -> float:
    """w = sqrt(32/9) b^3 = a^3/12 (GPUGrossPitaevskii_BCC.cu:71)."""
    return math.sqrt(32.0 / 9.0) * block_scale ** 3


def as_complex(a: np.ndarray) -> np.ndarray:
    return a.view(np.complex128)


class Solver:
    """One condensate resident on the current GPU (gpe_solver)."""

    def __init__(self, dims: Sequence[int], block_scale: float, dt: float, g: float,
                 origin: Optional[Sequence[float]] = None, precision: str = "fp64", mode: str = "real",
                 potential: str = "trap", trap_ratio: float = 1.0, kernel: str = "auto", tile: int = 0):
        self.dims = tuple(int(d) for d in dims)
        self.block_scale = float(block_scale)
        if origin is None:
            origin = centered_origin(self.dims, block_scale)
        cfg = _lib.GpeConfig()
        cfg.nx, cfg.ny, cfg.nz = self.dims
        cfg.precision = PRECISIONS[precision]
        cfg.block_scale = float(block_scale)
        cfg.dt = float(dt)
        cfg.g = float(g)
        cfg.origin = _lib.d3(origin)
        cfg.potential = POTENTIALS[potential]
        cfg.mode = MODES[mode]
        cfg.trap_ratio = float(trap_ratio)
        cfg.kernel = KERNELS[kernel]
        cfg.tile = int(tile)
        self.cfg = cfg
        self.origin = tuple(float(o) for o in origin)
        self._h = C.c_void_p()
        check(lib.gpe_create(C.byref(cfg), C.byref(self._h)), "gpe_create")
        self.nodes = nodes_of(self.dims)

    # -- checkpoint and resume --------------------------------------------
    def save_checkpoint(self, path: str) -> None:
        """Writes the run's state (configuration in force, steps taken, both time levels, a
        streamed potential) to `path` (gpe_checkpoint_save)."""
        check(lib.gpe_checkpoint_save(self._h, os.fsencode(path)), "gpe_checkpoint_save")

    @classmethod
    def load_checkpoint(cls, path: str) -> "Solver":
        """A solver resuming the run saved at `path` (gpe_checkpoint_load): stepping it continues
        the saved run bit for bit."""
        h = C.c_void_p()
        check(lib.gpe_checkpoint_load(os.fsencode(path), C.byref(h)), "gpe_checkpoint_load")
        s = cls.__new__(cls)
        s._h = h
        cfg = _lib.GpeConfig()
        check(lib.gpe_get_config(h, C.byref(cfg)), "gpe_get_config")
        s.cfg = cfg
        s.dims = (int(cfg.nx), int(cfg.ny), int(cfg.nz))
        s.block_scale = float(cfg.block_scale)
        s.origin = tuple(float(o) for o in cfg.origin)
        s.nodes = nodes_of(s.dims)
        return s

    def config(self) -> dict:
        """The configuration in force (gpe_get_config)."""
        cfg = _lib.GpeConfig()
        check(lib.gpe_get_config(self._h, C.byref(cfg)), "gpe_get_config")
        out = {k: getattr(cfg, k) for k, _ in _lib.GpeConfig._fields_}
        out["origin"] = tuple(cfg.origin)
        return out

    # -- lifecycle -------------------------------------------------------
    def close(self) -> None:
        if self._h:
            lib.gpe_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state -----------------------------------------------------------
    def set_potential(self, V: np.ndarray) -> None:
        V = np.ascontiguousarray(V, dtype=np.float64)
        if V.size != self.nodes:
            raise ValueError(f"potential has {V.size} values, expected {self.nodes}")
        check(lib.gpe_set_potential(self._h, dptr(V)), "gpe_set_potential")

    def set_state(self, even: np.ndarray, odd: Optional[np.ndarray] = None) -> None:
        even = self._field(even)
        odd = None if odd is None else self._field(odd)
        check(lib.gpe_set_state(self._h, dptr(even), dptr(odd)), "gpe_set_state")

    def get_state(self, odd: bool = True):
        e = np.empty(2 * self.nodes)
        o = np.empty(2 * self.nodes) if odd else None
        check(lib.gpe_get_state(self._h, dptr(e), dptr(o)), "gpe_get_state")
        return (e, o) if odd else e

    # -- output (GPUGrossPitaevskii_BCC.cu:424-468), reduced on the device --
    def volume_map(self, unit: float, reference_class0: bool = False) -> np.ndarray:
        """uint16 [nz+2, ny+2, nx+2]: min(65535, trunc(unit * sum_l |psi_l|^2)) of the even field over
        the padded extent; reference_class0 reproduces the reference's |psi_0|^2 added 12 times."""
        nx, ny, nz = self.dims
        out = np.empty((nz + 2, ny + 2, nx + 2), np.uint16)
        check(lib.gpe_volume_map(self._h, float(unit), 1 if reference_class0 else 0,
                                 out.ctypes.data_as(C.POINTER(C.c_uint16))), "gpe_volume_map")
        return out

    def slice_picture(self, k: int, gain: float = 5.0, middle: bool = False) -> np.ndarray:
        """uint8 [ny+2, nx+2, 3] (blue, green, red) of padded z layer k, as Picture::save stores it."""
        nx, ny, _ = self.dims
        out = np.empty((ny + 2, nx + 2, 3), np.uint8)
        check(lib.gpe_slice_picture(self._h, int(k), float(gain), 1 if middle else 0,
                                    out.ctypes.data_as(C.POINTER(C.c_uint8))), "gpe_slice_picture")
        return out

    def set_mode(self, mode: str, dt: float) -> None:
        check(lib.gpe_set_mode(self._h, MODES[mode], float(dt)), "gpe_set_mode")
        self.cfg.mode = MODES[mode]
        self.cfg.dt = float(dt)

    def init_odd(self, mu: float) -> None:
        check(lib.gpe_init_odd(self._h, float(mu)), "gpe_init_odd")

    def init_gaussian(self, sigma: float = 1.5, noise: float = 1e-3, seed: int = 0, mu: float = float("nan")) -> None:
        """Even field = the device Gaussian recipe normalised on the grid (no host array); in real
        time odd = even exp(i mu dt/2), mu NaN = the state's own chemical potential."""
        check(lib.gpe_init_gaussian(self._h, float(sigma), float(noise), int(seed), float(mu)), "gpe_init_gaussian")

    # -- stepping --------------------------------------------------------
    def step(self, nsteps: int) -> None:
        check(lib.gpe_step(self._h, int(nsteps)), "gpe_step")

    def time_steps(self, nsteps: int):
        """Enqueue nsteps and return (device ms, hot-kernel launches)."""
        ms = C.c_double()
        nl = C.c_uint64()
        check(lib.gpe_time_steps(self._h, int(nsteps), C.byref(ms), C.byref(nl)), "gpe_time_steps")
        return ms.value, nl.value

    def synchronize(self) -> None:
        check(lib.gpe_synchronize(self._h), "gpe_synchronize")

    def observables(self, field: int = 0) -> dict:
        out = np.zeros(4)
        check(lib.gpe_observables(self._h, int(field), dptr(out)), "gpe_observables")
        N, K, P, I = (float(x) for x in out)
        return {"N": N, "K": K, "P": P, "I": I, "E": K + P + 0.5 * I, "mu": (K + P + I) / N if N else float("nan")}

    def step_observe(self, nsteps: int) -> dict:
        """nsteps real-time steps; returns the observables of the even field as the last step read
        it (after nsteps - 1 steps), reduced inside that step's first half-update."""
        out = np.zeros(4)
        check(lib.gpe_step_observe(self._h, int(nsteps), dptr(out)), "gpe_step_observe")
        N, K, P, I = (float(x) for x in out)
        return {"N": N, "K": K, "P": P, "I": I, "E": K + P + 0.5 * I, "mu": (K + P + I) / N if N else float("nan")}

    def compare(self, other: "Solver", field: int = 0, other_field: int = 0) -> dict:
        """Distance of this solver's field to `other`'s on the same grid, reduced on the device in
        FP64 (neither field leaves the GPU): rel_l2 = ||self - other|| / ||other||."""
        out = np.zeros(3)
        check(lib.gpe_compare_fields(self._h, int(field), other._h, int(other_field), dptr(out)),
              "gpe_compare_fields")
        d2, b2, a2 = (float(x) for x in out)
        return {"diff2": d2, "other2": b2, "self2": a2, "rel_l2": math.sqrt(d2 / b2) if b2 > 0 else float("nan")}

    def device_field(self, field: int = 0) -> int:
        """Device address of field 0 (even) or 1 (odd) in the layout field_layout() describes."""
        p = C.c_void_p()
        check(lib.gpe_get_field_device(self._h, int(field), C.byref(p)), "gpe_get_field_device")
        return int(p.value or 0)

    def set_variant(self, kernel: str, tile: int) -> None:
        check(lib.gpe_set_variant(self._h, KERNELS[kernel], int(tile)), "gpe_set_variant")

    def info(self) -> dict:
        inf = _lib.GpeSolverInfo()
        check(lib.gpe_get_info(self._h, C.byref(inf)), "gpe_get_info")
        return {"nodes": inf.nodes, "device_bytes": inf.device_bytes, "steps_done": inf.steps_done,
                "kernel": KERNEL_NAMES.get(inf.kernel, str(inf.kernel)), "tile": inf.tile,
                "launches_per_step": inf.launches_per_step, "bytes_per_node_step": inf.bytes_per_node_step}

    def _field(self, a: np.ndarray) -> np.ndarray:
        a = np.ascontiguousarray(a)
        if a.dtype == np.complex128:
            a = a.view(np.float64)
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        if a.size != 2 * self.nodes:
            raise ValueError(f"field has {a.size} doubles, expected {2 * self.nodes}")
        return a


def observables_noalloc(cfg, field: int, out: int, pot_dt: int = 0, stream: int = 0) -> None:
    """gpe_observables_noalloc: enqueue the observables {N, K, P, I} of the device field at
    address `field` (described by the GpeConfig cfg) into the 4 device doubles at `out` on
    `stream`, allocating nothing.  Asynchronous: synchronise the stream before reading out."""
    check(lib.gpe_observables_noalloc(C.byref(cfg), C.c_void_p(field), C.c_void_p(pot_dt or None),
                                      C.c_void_p(out), C.c_void_p(stream or None)), "gpe_observables_noalloc")


def field_layout(dims: Sequence[int], precision: str = "fp64") -> dict:
    """Device layout of one field: node (i, j, k, l) is complex element
    origin + k*sz + j*sy + l*xp + i of a buffer of `elems` elements of `elem_bytes` bytes."""
    L = _lib.GpeFieldLayout()
    check(lib.gpe_get_field_layout(*(int(d) for d in dims), PRECISIONS[precision], C.byref(L)),
          "gpe_get_field_layout")
    return {k: int(getattr(L, k)) for k, _ in _lib.GpeFieldLayout._fields_ if k != "reserved"}


def pack_field(dims: Sequence[int], psi: np.ndarray, precision: str = "fp64") -> np.ndarray:
    """Host field ([k][j][i][l], re/im interleaved) in the device layout, as a flat array of
    float64 (fp64) or float32 (fp32) re/im pairs; zero outside the box."""
    L = field_layout(dims, precision)
    nx, ny, nz = (int(d) for d in dims)
    out = np.zeros(2 * L["elems"], dtype=np.float64 if precision == "fp64" else np.float32)
    c = np.asarray(psi, dtype=np.float64).reshape(nz, ny, nx, 12, 2)
    k, j, i, l = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), np.arange(12), indexing="ij")
    e = L["origin"] + k * L["sz"] + j * L["sy"] + l * L["xp"] + i
    out[2 * e] = c[..., 0]
    out[2 * e + 1] = c[..., 1]
    return out


def variants():
    """All (kernel family, tile, shape) combinations the library can run."""
    out = []
    for name in ("gather", "zmarch", "fused"):
        k = KERNELS[name]
        for t in range(1, lib.gpe_variant_tiles(k) + 1):
            sh = (C.c_int32 * 3)()
            lib.gpe_variant_shape(k, t, sh)
            out.append({"kernel": name, "tile": t, "waves": sh[0], "cz": sh[1],
                        "remap": 2 if sh[2] & 4 else sh[2] & 1, "nt": (sh[2] >> 1) & 1,
                        "row": (sh[2] >> 3) & 1, "node": (sh[2] >> 4) & 3})
    return out


def stream_bandwidth(kind: str, nbytes: int = 1 << 30, iters: int = 20) -> float:
    """GB/s of a streaming probe over buffers of nbytes: "copy" (read + write), "rmw" (read prev,
    non-temporal read-modify-write of next: the half-update's traffic) or "read"."""
    kinds = {"copy": (0, 2), "rmw": (1, 3), "read": (2, 1)}
    k, moved = kinds[kind]
    ms = C.c_double()
    check(lib.gpe_diag_stream_bandwidth(k, int(nbytes), int(iters), C.byref(ms)), "gpe_diag_stream_bandwidth")
    return moved * nbytes / (ms.value * 1e-3) / 1e9


def copy_bandwidth(nbytes: int = 1 << 30, iters: int = 20) -> float:
    """GB/s of a plain device-to-device streaming copy (read + write bytes)."""
    ms = C.c_double()
    check(lib.gpe_diag_copy_bandwidth(int(nbytes), int(iters), C.byref(ms)), "gpe_diag_copy_bandwidth")
    return 2.0 * nbytes / (ms.value * 1e-3) / 1e9


# ---------------------------------------------------------------------------
# geometry
# ---------------------------------------------------------------------------
def bcc_positions() -> np.ndarray:
    out = (C.c_int32 * 36)()
    lib.gpe_bcc_positions(out)
    return np.array(out, dtype=np.int32).reshape(12, 3)


def bcc_stencil() -> np.ndarray:
    out = (C.c_int32 * 192)()
    lib.gpe_bcc_stencil(out)
    return np.array(out, dtype=np.int32).reshape(12, 4, 4)


def bcc_stencil_derived():
    out = (C.c_int32 * 192)()
    n = lib.gpe_bcc_stencil_derived(out)
    return n, np.array(out, dtype=np.int32).reshape(12, 4, 4)


# ---------------------------------------------------------------------------
# reference initial state
# ---------------------------------------------------------------------------
class VortexState:
    """Mirror of the reference VortexState (koodit/GrossPitaevskii/VortexState.hpp:16-76),
    loaded from its binary file format (VortexState.cpp:49-72)."""

    def __init__(self):
        # VortexState() (VortexState.cpp:5-10): kappa 2, g 0, a 100 x 1 table.
        self._h = C.c_void_p()
        check(lib.gpe_vortex_state_create(C.byref(self._h)), "gpe_vortex_state_create")

    def load(self, path: str) -> bool:
        h = C.c_void_p()
        st = lib.gpe_vortex_state_load(str(path).encode(), C.byref(h))
        if st != _lib.OK:
            return False
        self.close()
        self._h = h
        return True

    @classmethod
    def from_table(cls, kappa: int, g: float, rplug: float, rstep: float, zstep: float, mu: float,
                   f: np.ndarray) -> "VortexState":
        """A state from memory: f is the rsize x zsize table f(r, z) (r-major), the rest the
        reference object's members (VortexState.hpp:16-76)."""
        f = np.ascontiguousarray(f, dtype=np.float64)
        if f.ndim != 2:
            raise ValueError("f must be an rsize x zsize array")
        h = C.c_void_p()
        check(lib.gpe_vortex_state_from_table(int(kappa), float(g), f.shape[0], f.shape[1], float(rplug),
                                              float(rstep), float(zstep), float(mu), dptr(f), C.byref(h)),
              "gpe_vortex_state_from_table")
        vs = cls.__new__(cls)
        vs._h = h
        return vs

    def table(self) -> np.ndarray:
        """The rsize x zsize table f(r, z)."""
        p = self._params()
        f = np.empty((int(p[2]), int(p[3])))
        check(lib.gpe_vortex_state_table(self._h, dptr(f)), "gpe_vortex_state_table")
        return f

    def close(self):
        if self._h:
            lib.gpe_vortex_state_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _params(self):
        out = np.zeros(8)
        lib.gpe_vortex_state_params(self._h, dptr(out))
        return out

    def getKappa(self) -> int:
        return int(self._params()[0])

    def getG(self) -> float:
        return float(self._params()[1])

    def getMu(self) -> float:
        return float(self._params()[7])

    def getPsi(self, p: Sequence[float]) -> complex:
        out = np.zeros(2)
        lib.gpe_vortex_state_psi(self._h, float(p[0]), float(p[1]), float(p[2]), dptr(out))
        return complex(out[0], out[1])

    def getPhase(self, time: float) -> complex:
        out = np.zeros(2)
        lib.gpe_vortex_state_phase(self._h, float(time), dptr(out))
        return complex(out[0], out[1])

    def searchFunctionMax(self) -> float:
        return lib.gpe_vortex_state_search_max(self._h)

    def searchMaxR(self, eps: float) -> float:
        return lib.gpe_vortex_state_search_max_r(self._h, float(eps))

    def searchMaxZ(self, eps: float) -> float:
        return lib.gpe_vortex_state_search_max_z(self._h, float(eps))

    def integrateCurvature(self) -> float:
        return lib.gpe_vortex_state_curvature(self._h)

    # -- generating a state (VortexState.cpp:12-47, 74-196, 250-257, 282-311, 342-361) --
    def setKappa(self, kappa: int):
        lib.gpe_vortex_state_set_kappa(self._h, int(kappa))

    def setG(self, g: float):
        lib.gpe_vortex_state_set_g(self._h, float(g))

    def setThreads(self, threads: int):
        """Threads of iterateSolution: 0 automatic, 1 serial, n threads (same result for any n)."""
        lib.gpe_vortex_state_set_threads(self._h, int(threads))

    def setDimensions(self, rsize: int, rstep: float, rplug: float = 0.0, zsize: int = 1, zstep: float = 1.0):
        check(lib.gpe_vortex_state_set_dimensions(self._h, int(rsize), float(rstep), float(rplug), int(zsize),
                                                  float(zstep)), "gpe_vortex_state_set_dimensions")

    def setRange(self, rmin: float, rmax: float, zmax: float, rstep: float, zstep: float):
        check(lib.gpe_vortex_state_set_range(self._h, float(rmin), float(rmax), float(zmax), float(rstep),
                                             float(zstep)), "gpe_vortex_state_set_range")

    def initialize(self, state: "VortexState"):
        check(lib.gpe_vortex_state_initialize(self._h, state._h), "gpe_vortex_state_initialize")

    def iterateSolution(self, potential, maxiters: int = 100000, eps: float = 1e-28, progress=None) -> dict:
        """Relaxes f to a stationary state of the potential (a TrapRZ, evaluated natively, or any
        callable V(r, z)).  progress(iteration, diff, prob) is called every 1000 iterations, where
        the reference prints.  Returns the run's report (iterations, converged, diff, prob, mu, error)."""
        fn, ratio = _rz_potential(potential)
        cb = _lib.RELAX_PROGRESS(lambda it, d, p, _u: progress(it, d, p)) if progress else _lib.RELAX_PROGRESS()
        rep = _lib.GpeRelaxReport()
        check(lib.gpe_vortex_state_iterate(self._h, fn, ratio, None, int(maxiters), float(eps), cb, None,
                                           C.byref(rep)), "gpe_vortex_state_iterate")
        return {k: getattr(rep, k) for k, _ in rep._fields_}

    def integrateFunctionError(self, potential) -> float:
        fn, ratio = _rz_potential(potential)
        return lib.gpe_vortex_state_function_error(self._h, fn, ratio, None)

    def save(self, path: str) -> bool:
        return lib.gpe_vortex_state_save(self._h, str(path).encode()) == _lib.OK

    def getFunction(self, r: float, z: float) -> float:
        return lib.gpe_vortex_state_function(self._h, float(r), float(z))

    def getMinR(self) -> float:
        return float(self._params()[4])

    def getMaxR(self) -> float:
        p = self._params()
        return float(p[4] + int(p[2]) * p[5])

    def getMaxZ(self) -> float:
        p = self._params()
        return float(int(p[3]) * p[6])

    def searchMinR(self, eps: float) -> float:
        return lib.gpe_vortex_state_search_min_r(self._h, float(eps))

    def integrateZ(self, vol: float) -> float:
        return lib.gpe_vortex_state_integrate_z(self._h, float(vol))


class TrapRZ:
    """The trap 0.5(r^2 + ratio^2 z^2) of potentialRZ (GPUGrossPitaevskii_BCC.cu:28-31); passed to
    VortexState.iterateSolution it is evaluated inside the library."""

    def __init__(self, ratio: float):
        self.ratio = float(ratio)

    def __call__(self, r: float, z: float) -> float:
        return 0.5 * (r * r + self.ratio * self.ratio * z * z)


def _rz_potential(potential):
    if isinstance(potential, TrapRZ):
        return _lib.RZ_POTENTIAL(), potential.ratio
    if not callable(potential):
        raise TypeError("potential must be a TrapRZ or a callable V(r, z)")
    return _lib.RZ_POTENTIAL(lambda r, z, _u: float(potential(r, z))), 0.0


def reference_state_recipe(kappa: int, g: float, trap_ratio: float, is_3d: bool = True, iters: int = 10000,
                           eps: float = 1e-29, progress=None):
    """The state recipe of GPUGrossPitaevskii_BCC.cu:516-535 (LOAD_STATE_FROM_DISK 0).

    Returns (state, (minr, maxr, maxz)) with the extents main passes to integrateInTime."""
    cb = _lib.RELAX_PROGRESS(lambda it, d, p, _u: progress(it, d, p)) if progress else _lib.RELAX_PROGRESS()
    h = C.c_void_p()
    ext = np.zeros(3)
    check(lib.gpe_vortex_state_reference_recipe(int(kappa), float(g), float(trap_ratio), 1 if is_3d else 0,
                                                int(iters), float(eps), cb, None, C.byref(h), dptr(ext)),
          "gpe_vortex_state_reference_recipe")
    vs = VortexState.__new__(VortexState)
    vs._h = h
    return vs, (float(ext[0]), float(ext[1]), float(ext[2]))


def reference_dims(block_scale, minp, maxp):
    dims = (C.c_uint32 * 3)()
    origin = (C.c_double * 3)()
    lib.gpe_reference_dims(float(block_scale), _lib.d3(minp), _lib.d3(maxp), dims, origin)
    return tuple(dims), tuple(origin)


def reference_setup(state: VortexState, block_scale, minp, maxp, iteration_period=1.0, trap_ratio=0.1,
                    seed=54363, fields=True):
    """Setup of integrateInTime (GPUGrossPitaevskii_BCC.cu:245-348).

    Returns (run dict, even, odd, V); V is unscaled."""
    run = _lib.GpeReferenceRun()
    check(lib.gpe_reference_setup(state._h, float(block_scale), _lib.d3(minp), _lib.d3(maxp),
                                  float(iteration_period), float(trap_ratio), int(seed), C.byref(run),
                                  None, None, None), "gpe_reference_setup")
    info = _run_dict(run)
    if not fields:
        return info, None, None, None
    n = nodes_of(info["dims"])
    even, odd, V = np.empty(2 * n), np.empty(2 * n), np.empty(n)
    check(lib.gpe_reference_setup(state._h, float(block_scale), _lib.d3(minp), _lib.d3(maxp),
                                  float(iteration_period), float(trap_ratio), int(seed), C.byref(run),
                                  dptr(even), dptr(odd), dptr(V)), "gpe_reference_setup")
    return info, even, odd, V


def _run_dict(run) -> dict:
    return {"dims": tuple(run.dims), "steps_per_iteration": run.steps_per_iteration, "origin": tuple(run.origin),
            "dt": run.dt, "g": run.g, "maxpot": run.maxpot, "lapfac_dt": run.lapfac_dt,
            "lapfac0_dt": run.lapfac0_dt}


def integrateInTime(state: VortexState, block_scale: float, minp, maxp, iteration_period: float,
                    number_of_iterations: int, trap_ratio: float = 0.1, precision: str = "fp64",
                    return_fields: bool = False):
    """Drop-in for the reference's integrateInTime (GPUGrossPitaevskii_BCC.cu:218-499).

    Returns 0 like the reference; with return_fields=True returns
    (0, run, even, odd)."""
    run = _lib.GpeReferenceRun()
    even = odd = None
    if return_fields:
        info, _, _, _ = reference_setup(state, block_scale, minp, maxp, iteration_period, trap_ratio, fields=False)
        n = nodes_of(info["dims"])
        even, odd = np.empty(2 * n), np.empty(2 * n)
    check(lib.gpe_integrate_in_time(state._h, float(block_scale), _lib.d3(minp), _lib.d3(maxp),
                                    float(iteration_period), int(number_of_iterations), float(trap_ratio),
                                    PRECISIONS[precision], C.byref(run), dptr(even), dptr(odd)),
          "gpe_integrate_in_time")
    if return_fields:
        return 0, _run_dict(run), even, odd
    return 0


integrate_in_time = integrateInTime


def integrateInTimeWithOutput(state: VortexState, block_scale: float, minp, maxp, iteration_period: float,
                              number_of_iterations: int, trap_ratio: float = 0.1, precision: str = "fp64",
                              save_picture: bool = True, save_volume: bool = True, prefix: str = "",
                              reference_class0: bool = False) -> dict:
    """integrateInTime with the reference's SAVE_PICTURE / SAVE_VOLUME output
    (GPUGrossPitaevskii_BCC.cu:17-18, 424-468): <prefix>kuva<iter>.bmp and <prefix>volume<iter>.mhd
    (+ .raw) before each iteration and after the last.  Returns the run's setup scalars."""
    run = _lib.GpeReferenceRun()
    opt = _lib.GpeOutputOptions()
    opt.save_picture = 1 if save_picture else 0
    opt.save_volume = 1 if save_volume else 0
    opt.volume_flags = 1 if reference_class0 else 0
    opt.prefix = str(prefix).encode()
    check(lib.gpe_integrate_in_time_output(state._h, float(block_scale), _lib.d3(minp), _lib.d3(maxp),
                                           float(iteration_period), int(number_of_iterations), float(trap_ratio),
                                           PRECISIONS[precision], C.byref(opt), C.byref(run)),
          "gpe_integrate_in_time_output")
    return _run_dict(run)


def write_volume_map(path: str, vol: np.ndarray, spacing) -> None:
    """saveVolumeMap (GPUGrossPitaevskii_BCC.cu:38-67) of a uint16 [z][y][x] volume."""
    vol = np.ascontiguousarray(vol, np.uint16)
    dims = (C.c_uint32 * 3)(vol.shape[2], vol.shape[1], vol.shape[0])
    check(lib.gpe_write_volume_map(str(path).encode(), vol.ctypes.data_as(C.POINTER(C.c_uint16)), dims,
                                   _lib.d3(spacing)), "gpe_write_volume_map")


def write_bmp(path: str, bgr: np.ndarray) -> None:
    """Picture::save (Output/Picture.cpp:81-137) of a uint8 [height][width][3] (blue, green, red) image."""
    bgr = np.ascontiguousarray(bgr, np.uint8)
    check(lib.gpe_write_bmp(str(path).encode(), bgr.shape[1], bgr.shape[0], bgr.ctypes.data_as(C.POINTER(C.c_uint8))),
          "gpe_write_bmp")


# ---------------------------------------------------------------------------
# recipes
# ---------------------------------------------------------------------------
def centered_origin(dims, block_scale):
    out = (C.c_double * 3)()
    lib.gpe_centered_origin(_lib.u32(dims), float(block_scale), out)
    return tuple(out)


def node_positions(dims, origin, block_scale) -> np.ndarray:
    out = np.empty(3 * nodes_of(dims))
    lib.gpe_node_positions(_lib.u32(dims), _lib.d3(origin), float(block_scale), dptr(out))
    return out.reshape(-1, 3)


def trap_potential(dims, origin, block_scale, ratio) -> np.ndarray:
    out = np.empty(nodes_of(dims))
    lib.gpe_trap_potential(_lib.u32(dims), _lib.d3(origin), float(block_scale), float(ratio), dptr(out))
    return out


def field_norm(dims, block_scale, psi) -> float:
    return lib.gpe_field_norm(_lib.u32(dims), float(block_scale), dptr(np.ascontiguousarray(psi, np.float64)))


def gaussian(dims, origin, block_scale, sigma, noise, seed) -> np.ndarray:
    out = np.empty(2 * nodes_of(dims))
    check(lib.gpe_recipe_gaussian(_lib.u32(dims), _lib.d3(origin), float(block_scale), float(sigma),
                                  float(noise), int(seed), dptr(out)), "gpe_recipe_gaussian")
    return out


def thomas_fermi(dims, origin, block_scale, g, ratio):
    out = np.empty(2 * nodes_of(dims))
    mu = C.c_double()
    check(lib.gpe_recipe_thomas_fermi(_lib.u32(dims), _lib.d3(origin), float(block_scale), float(g), float(ratio),
                                      dptr(out), C.byref(mu)), "gpe_recipe_thomas_fermi")
    return out, mu.value


def vortex_line(dims, origin, block_scale, x0, y0, psi) -> np.ndarray:
    check(lib.gpe_recipe_vortex_line(_lib.u32(dims), _lib.d3(origin), float(block_scale), float(x0), float(y0),
                                     dptr(psi)), "gpe_recipe_vortex_line")
    return psi


def vortex_ring(dims, origin, block_scale, center, axis, radius, psi) -> np.ndarray:
    check(lib.gpe_recipe_vortex_ring(_lib.u32(dims), _lib.d3(origin), float(block_scale), _lib.d3(center),
                                     _lib.d3(axis), float(radius), dptr(psi)), "gpe_recipe_vortex_ring")
    return psi


def phase_shift(dims, even, mu, dt) -> np.ndarray:
    odd = np.empty_like(even)
    check(lib.gpe_recipe_phase_shift(_lib.u32(dims), dptr(even), float(mu), float(dt), dptr(odd)),
          "gpe_recipe_phase_shift")
    return odd


def leapfrog_coefs(block_scale: float, dt: float, g: float):
    """(lapfac*dt, lapfac0*dt, g*dt) as GPUGrossPitaevskii_BCC.cu:291-306 form them."""
    lapfac = -0.5 * HODGE / (block_scale * block_scale)
    lapfac0 = 4.0 * (-lapfac)
    return lapfac * dt, lapfac0 * dt, g * dt
