"""Gross-Pitaevskii stepper with the 14-point primal BCC operator (gpe_primal_*).

psi lives on the 2 N^3 sites of a body-centred cubic lattice -- the corners
and centres of N^3 cubic cells of edge ``a`` -- and the Laplacian is the DEC
Laplacian of 0-forms the reference's mesh library builds on that lattice
(DelaunayMesh::createBccGrid, jukan_koodit/Mesh/DelaunayMesh.cpp:178; Hodge
weights Mesh/Mesh.cpp:1479-1489):

    (L psi)_i = (1/a^2) [ 3/4 sum_8 (psi_j - psi_i) + 1/4 sum_6 (psi_j - psi_i) ]

over the 8 diagonal and 6 axial neighbours.  The integrators are those of
the dual stepper (leapfrog in real time, Euler + renormalisation in
imaginary time, GPUGrossPitaevskii_BCC.cu:475-481), with the harmonic trap
evaluated in the kernel.  Fields are flat float64 arrays of 4 nx ny nz values
in the host order [k][j][i][s] (s = 0 corner, s = 1 centre), complex
interleaved.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import check, dptr, lib
from .solver import MODES, PRECISIONS


def sites_of(dims: Sequence[int]) -> int:
    return 2 * int(dims[0]) * int(dims[1]) * int(dims[2])


def site_measure(a: float) -> float:
    """Dual volume of one site, a^3/2."""
    return 0.5 * a ** 3


def primal_coefficients(a: float, dt: float, g: float) -> np.ndarray:
    """{cD dt, cA dt, c0 dt, g dt} as the kernels use them."""
    out = np.zeros(4)
    lib.gpe_primal_coefficients(float(a), float(dt), float(g), dptr(out))
    return out


def primal_centered_origin(dims: Sequence[int], a: float) -> tuple:
    out = np.zeros(3)
    lib.gpe_primal_centered_origin(_lib.u32(dims), float(a), dptr(out))
    return tuple(float(x) for x in out)


def primal_positions(dims: Sequence[int], origin: Sequence[float], a: float) -> np.ndarray:
    """[sites, 3] positions in host order."""
    out = np.zeros((sites_of(dims), 3))
    lib.gpe_primal_positions(_lib.u32(dims), dptr(np.ascontiguousarray(origin, np.float64)), float(a), dptr(out))
    return out


def primal_trap_potential(dims: Sequence[int], origin: Sequence[float], a: float, ratio: float) -> np.ndarray:
    out = np.zeros(sites_of(dims))
    lib.gpe_primal_trap_potential(_lib.u32(dims), dptr(np.ascontiguousarray(origin, np.float64)), float(a),
                                  float(ratio), dptr(out))
    return out


def primal_tiles() -> int:
    return int(lib.gpe_primal_tiles())


class PrimalSolver:
    """One condensate on the BCC sites, resident on the current GPU (gpe_primal)."""

    def __init__(self, dims: Sequence[int], a: float, dt: float, g: float = 0.0,
                 origin: Optional[Sequence[float]] = None, trap_ratio: float = 1.0, precision: str = "fp64",
                 mode: str = "real", tile: int = 0):
        self.dims = tuple(int(d) for d in dims)
        self.a = float(a)
        if origin is None:
            origin = primal_centered_origin(self.dims, a)
        cfg = _lib.GpePrimalConfig()
        cfg.nx, cfg.ny, cfg.nz = self.dims
        cfg.precision = PRECISIONS[precision]
        cfg.a = float(a)
        cfg.dt = float(dt)
        cfg.g = float(g)
        cfg.origin = _lib.d3(origin)
        cfg.trap_ratio = float(trap_ratio)
        cfg.mode = MODES[mode]
        cfg.tile = int(tile)
        self.cfg = cfg
        self.origin = tuple(float(o) for o in origin)
        self.sites = sites_of(self.dims)
        self._h = C.c_void_p()
        check(lib.gpe_primal_create(C.byref(cfg), C.byref(self._h)), "gpe_primal_create")

    def close(self) -> None:
        if self._h:
            lib.gpe_primal_destroy(self._h)
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

    def set_state(self, even: np.ndarray, odd: Optional[np.ndarray] = None) -> None:
        even = self._field(even)
        odd = None if odd is None else self._field(odd)
        check(lib.gpe_primal_set_state(self._h, dptr(even), dptr(odd)), "gpe_primal_set_state")

    def get_state(self, odd: bool = True):
        e = np.empty(2 * self.sites)
        o = np.empty(2 * self.sites) if odd else None
        check(lib.gpe_primal_get_state(self._h, dptr(e), dptr(o)), "gpe_primal_get_state")
        return (e, o) if odd else e

    def set_mode(self, mode: str, dt: float) -> None:
        check(lib.gpe_primal_set_mode(self._h, MODES[mode], float(dt)), "gpe_primal_set_mode")
        self.cfg.mode = MODES[mode]
        self.cfg.dt = float(dt)

    def init_odd(self, mu: float) -> None:
        check(lib.gpe_primal_init_odd(self._h, float(mu)), "gpe_primal_init_odd")

    def step(self, nsteps: int) -> None:
        check(lib.gpe_primal_step(self._h, int(nsteps)), "gpe_primal_step")

    def time_steps(self, nsteps: int):
        """Enqueue nsteps and return (device ms, hot-kernel launches)."""
        ms = C.c_double()
        nl = C.c_uint64()
        check(lib.gpe_primal_time_steps(self._h, int(nsteps), C.byref(ms), C.byref(nl)), "gpe_primal_time_steps")
        return ms.value, nl.value

    def synchronize(self) -> None:
        check(lib.gpe_primal_synchronize(self._h), "gpe_primal_synchronize")

    def observables(self, field: int = 0) -> dict:
        out = np.zeros(4)
        check(lib.gpe_primal_observables(self._h, int(field), dptr(out)), "gpe_primal_observables")
        N, K, P, I = (float(x) for x in out)
        return {"N": N, "K": K, "P": P, "I": I, "E": K + P + 0.5 * I, "mu": (K + P + I) / N if N else float("nan")}

    def _field(self, a: np.ndarray) -> np.ndarray:
        a = np.ascontiguousarray(a)
        if a.dtype == np.complex128:
            a = a.view(np.float64)
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        if a.size != 2 * self.sites:
            raise ValueError(f"field has {a.size} doubles, expected {2 * self.sites}")
        return a
