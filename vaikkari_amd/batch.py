"""Batched stepping of many independent condensates (include/vaikkari_gpe_batch.h).

The reference steps one field per process; a Batch keeps many instances
resident and advances them with shared persistent launches captured in a
hipGraph per chunk of steps.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import check, dptr, lib
from .solver import MODES, PRECISIONS, nodes_of

TILE_CACHED = 0x100  # GPE_BATCH_TILE_CACHED: cached read-modify-write of the real-time pass
TILE_SLABS = 0x200  # GPE_BATCH_TILE_SLABS: each XCD takes a contiguous eighth of an instance's tiles
TILE_GROUP_SLABS = 0x400  # GPE_BATCH_TILE_GROUP_SLABS: ... of a whole group's tiles (whole instances)
TILE_QUAD = 0x800  # GPE_BATCH_TILE_QUAD: one thread per cell and group of 4 classes (3, 6 or 12 waves)


def instance_config(dims: Sequence[int], h: float, g: float, dt: float, steps: int, mode: str = "real",
                    trap_ratio: float = 1.0, init: str = "gaussian", sigma: float = 1.5, noise: float = 1e-3,
                    mu0: float = float("nan"), seed: int = 0, zlayers: int = 0,
                    origin: Optional[Sequence[float]] = None) -> _lib.GpeInstanceConfig:
    """Instance of cube edge h (block_scale h/sqrt(8)), harmonic trap 0.5(x^2+y^2+(ratio z)^2).
    With init "gaussian", mu0 NaN (the default) starts the leapfrog with the state's own chemical
    potential, odd = even exp(i mu dt/2), mu = <H>/N computed on the device."""
    c = _lib.GpeInstanceConfig()
    c.nx, c.ny, c.nz = (int(d) for d in dims)
    c.mode = MODES[mode]
    c.block_scale = float(h) / math.sqrt(8.0)
    c.dt = float(dt)
    c.g = float(g)
    c.trap_ratio = float(trap_ratio)
    c.origin = _lib.d3(origin if origin is not None else (math.nan, math.nan, math.nan))
    c.steps = int(steps)
    c.init = _lib.INIT_GAUSSIAN if init == "gaussian" else _lib.INIT_HOST
    c.zlayers = int(zlayers)
    c.sigma, c.noise, c.mu0 = float(sigma), float(noise), float(mu0)
    c.seed = int(seed)
    return c


def tile_code(waves, cached: bool = False, slabs=False) -> int:
    """gpe_batch_set_tile's argument for (waves, cached, slabs); slabs False/0, True/1 or 2 (see
    Batch.set_tile).  A (waves, cached) or (waves, cached, slabs) tuple is accepted as waves.
    3, 6 or 12 waves select one thread per cell and group of 4 classes (GPE_BATCH_TILE_QUAD)."""
    if isinstance(waves, tuple):
        waves, cached, slabs = (tuple(waves) + (False, False))[:3]
    order = int(slabs)
    quad = TILE_QUAD if int(waves) in (3, 6, 12) else 0
    return int(waves) | quad | (TILE_CACHED if cached else 0) | {0: 0, 1: TILE_SLABS, 2: TILE_GROUP_SLABS}[order]


class Batch:
    def __init__(self, precision: str = "fp32", hbm_budget_bytes: int = 0):
        self._h = C.c_void_p()
        check(lib.gpe_batch_create(PRECISIONS[precision], int(hbm_budget_bytes), C.byref(self._h)),
              "gpe_batch_create")
        self.dims = {}

    def close(self):
        if self._h:
            lib.gpe_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tile(self, waves, cached: bool = False, slabs=False) -> None:
        """Real-time tile: waves per block (1, 2, 4, 8; or 3, 6, 12 for one thread per cell and
        group of 4 classes), whether the updated field's
        read-modify-write is cached (default: streamed) and how the tiles go to the XCDs: slabs
        True (or 1) gives each XCD a contiguous slab of every instance's tiles, 2 a contiguous
        eighth of a whole group's tiles (whole instances in layer order), False y-bands or plain
        order.  A (waves, cached) or (waves, cached, slabs) tuple is accepted."""
        check(lib.gpe_batch_set_tile(self._h, tile_code(waves, cached, slabs)), "gpe_batch_set_tile")

    def set_group_bytes(self, nbytes: int) -> None:
        """Real-time instances advance a chunk group by group, groups of at most nbytes of fields
        (0: one group); results do not depend on it."""
        check(lib.gpe_batch_set_group_bytes(self._h, int(nbytes)), "gpe_batch_set_group_bytes")

    @property
    def group_bytes(self) -> int:
        return int(lib.gpe_batch_group_bytes(self._h))

    def set_launch_mode(self, mode: str) -> None:
        """"graph": replay a captured hipGraph per chunk (default); "stream": enqueue directly."""
        check(lib.gpe_batch_set_launch_mode(self._h, {"graph": 0, "stream": 1}[mode]), "gpe_batch_set_launch_mode")

    def instance_bytes(self, cfg: _lib.GpeInstanceConfig) -> int:
        return int(lib.gpe_batch_instance_bytes(self._h, C.byref(cfg)))

    def add(self, cfg: _lib.GpeInstanceConfig) -> int:
        i = C.c_int32()
        check(lib.gpe_batch_add(self._h, C.byref(cfg), C.byref(i)), "gpe_batch_add")
        self.dims[i.value] = (cfg.nx, cfg.ny, cfg.nz, cfg.mode)
        return i.value

    def try_add(self, cfg: _lib.GpeInstanceConfig) -> Optional[int]:
        """add(), or None when the HBM budget does not fit the instance."""
        i = C.c_int32()
        st = lib.gpe_batch_add(self._h, C.byref(cfg), C.byref(i))
        if st == _lib.EBUSY:
            return None
        check(st, "gpe_batch_add")
        self.dims[i.value] = (cfg.nx, cfg.ny, cfg.nz, cfg.mode)
        return i.value

    def remove(self, i: int) -> None:
        check(lib.gpe_batch_remove(self._h, int(i)), "gpe_batch_remove")
        self.dims.pop(i, None)

    def set_state(self, i: int, even: np.ndarray, odd: Optional[np.ndarray] = None) -> None:
        check(lib.gpe_batch_set_state(self._h, int(i), dptr(np.ascontiguousarray(even, np.float64)),
                                      dptr(None if odd is None else np.ascontiguousarray(odd, np.float64))),
              "gpe_batch_set_state")

    def get_state(self, i: int):
        nx, ny, nz, mode = self.dims[i]
        n = nodes_of((nx, ny, nz))
        e = np.empty(2 * n)
        o = np.empty(2 * n) if mode == _lib.REAL_TIME else None
        check(lib.gpe_batch_get_state(self._h, int(i), dptr(e), dptr(o)), "gpe_batch_get_state")
        return e, o

    def observables(self, i: int) -> dict:
        out = np.zeros(4)
        check(lib.gpe_batch_observables(self._h, int(i), dptr(out)), "gpe_batch_observables")
        N, K, P, I = (float(x) for x in out)
        return {"N": N, "K": K, "P": P, "I": I, "E": K + P + 0.5 * I, "mu": (K + P + I) / N if N else float("nan")}

    def info(self, i: int) -> dict:
        inf = _lib.GpeInstanceInfo()
        check(lib.gpe_batch_instance(self._h, int(i), C.byref(inf)), "gpe_batch_instance")
        return {"steps_done": inf.steps_done, "steps_left": inf.steps_left, "nodes": inf.nodes,
                "device_bytes": inf.device_bytes, "mode": inf.mode, "active": bool(inf.active)}

    def run(self, chunk_steps: int, chunks: int = 1) -> float:
        """Advance every active instance by chunks x min(chunk_steps, steps left); returns device ms."""
        ms = C.c_double()
        check(lib.gpe_batch_run(self._h, int(chunk_steps), int(chunks), C.byref(ms)), "gpe_batch_run")
        return ms.value

    def chunk_node_steps(self, chunk_steps: int) -> int:
        return int(lib.gpe_batch_chunk_node_steps(self._h, int(chunk_steps)))

    @property
    def active(self) -> int:
        return int(lib.gpe_batch_active(self._h))

    @property
    def resident_bytes(self) -> int:
        return int(lib.gpe_batch_resident_bytes(self._h))
