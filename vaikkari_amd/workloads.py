"""The synthetic workloads the stepper is measured and tested on, one function per form, so that
bench.py and the tests build exactly the same inputs.

Conventions: a cell is a cube of edge h (block_scale b = h/sqrt(8)) holding 12 dual nodes; the
box of n^3 cells is centred on 0; the trap is V = 0.5(x^2 + y^2 + (ratio z)^2).

* drop-in form: BCC 128^3 cells, h = 0.1, g = 2000, ratio 2 (trap 0.5(x^2+y^2+4z^2));
  a Gaussian (sigma 1.5, 1e-3 noise, seed 1) relaxed in imaginary time, then a straight vortex
  line along z at an offset drawn from seed 2 in [-0.5, 0.5]^2, real time at dt = 2e-4.
* precision form: 256^3 cells, h = 0.06, g = 5000, isotropic trap; Thomas-Fermi profile with a
  vortex ring (seed 3: radius uniform in [2, 4], axis uniform on the sphere) through the centre,
  dt = 5e-5.
* batched form: 64 instances (seed 4), n uniform in {24, 32, 48, 64, 96}, h = 0.25, g
  log-uniform in [100, 5000], ratio uniform in [0.5, 4], dt = 1e-3 * 32/n capped by the
  reference's stability rule (serve.stable_dt), Gaussian + noise initial states made on the
  device.
* serving form: the Poisson stream of serve.poisson_workload (seed 5, 20 requests/s).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import solver as _s


@dataclass
class DropIn:
    dims: tuple
    h: float = 0.1
    g: float = 2000.0
    ratio: float = 2.0
    dt: float = 2e-4
    sigma: float = 1.5
    noise: float = 1e-3
    seed: int = 1
    vortex_seed: int = 2

    @property
    def b(self) -> float:
        return self.h / math.sqrt(8.0)

    @property
    def origin(self):
        return _s.centered_origin(self.dims, self.b)

    def initial(self) -> np.ndarray:
        """Phase-1 input: the seeded Gaussian."""
        return _s.gaussian(self.dims, self.origin, self.b, self.sigma, self.noise, self.seed)

    def dtau(self, psi0: np.ndarray) -> float:
        """Imaginary-time step: 1/(6/b^2 + max V + g max|psi0|^2), half the linear bound."""
        V = _s.trap_potential(self.dims, self.origin, self.b, self.ratio)
        return 1.0 / (6.0 / self.b ** 2 + float(V.max()) + self.g * float(np.max(psi0[0::2] ** 2 + psi0[1::2] ** 2)))

    def vortex_offset(self):
        rng = np.random.default_rng(self.vortex_seed)
        x0, y0 = rng.uniform(-0.5, 0.5, 2)
        return float(x0), float(y0)

    def imprint(self, ground: np.ndarray) -> np.ndarray:
        """Phase-2 input: the ground state with the vortex line imprinted (returns a new array)."""
        x0, y0 = self.vortex_offset()
        return _s.vortex_line(self.dims, self.origin, self.b, x0, y0, ground.copy())


def drop_in_form(n: int = 128) -> DropIn:
    return DropIn(dims=(n, n, n))


@dataclass
class PrecisionForm:
    dims: tuple
    h: float
    g: float
    dt: float
    psi: np.ndarray
    mu: float
    ring_radius: float
    ring_axis: tuple

    @property
    def b(self) -> float:
        return self.h / math.sqrt(8.0)

    @property
    def origin(self):
        return _s.centered_origin(self.dims, self.b)


def precision_form(n: int = 256, seed: int = 3) -> PrecisionForm:
    """Thomas-Fermi profile (normalised on the grid, with its chemical potential) and a vortex
    ring through the centre: radius uniform in [2, 4], axis uniform on the sphere (seed 3)."""
    dims = (n, n, n)
    h, g, dt = 0.06, 5000.0, 5e-5
    b = h / math.sqrt(8.0)
    origin = _s.centered_origin(dims, b)
    rng = np.random.default_rng(seed)
    psi, mu = _s.thomas_fermi(dims, origin, b, g, 1.0)
    R = float(rng.uniform(2.0, 4.0))
    ax = rng.normal(size=3)
    ax /= np.linalg.norm(ax)
    psi = _s.vortex_ring(dims, origin, b, (0.0, 0.0, 0.0), ax, R, psi)
    return PrecisionForm(dims, h, g, dt, psi, mu, R, tuple(float(x) for x in ax))


BATCH_SIZES = (24, 32, 48, 64, 96)


def batched_form(n_inst: int = 64, steps: int = 2000, seed: int = 4, h: float = 0.25, sizes=BATCH_SIZES):
    """Instance configurations of the batched form: list of dicts (dims, h, g, dt, steps, mode,
    trap_ratio, seed) for vaikkari_amd.batch.instance_config.  `sizes` replaces the set the cube
    edges are drawn from (one size: a batch of equal grids)."""
    from .serve import stable_dt

    rng = np.random.default_rng(seed)
    out = []
    for q in range(n_inst):
        n = int(rng.choice(tuple(sizes)))
        g = float(np.exp(rng.uniform(np.log(100.0), np.log(5000.0))))
        ratio = float(rng.uniform(0.5, 4.0))
        out.append({"dims": (n, n, n), "h": h, "g": g, "dt": stable_dt(n, h, g, ratio), "steps": int(steps),
                    "mode": "real", "trap_ratio": ratio, "seed": seed * 100003 + q})
    return out


def instance_configs(forms, **kw):
    """vaikkari_amd.batch.instance_config of each batched_form entry (Gaussian + 1e-3 noise,
    sigma 1.5, the state's own chemical potential)."""
    from .batch import instance_config

    return [instance_config(f["dims"], f["h"], f["g"], f["dt"], f["steps"], f["mode"], f["trap_ratio"],
                            sigma=1.5, noise=1e-3, seed=f["seed"], **kw) for f in forms]


def serving_form(seconds: float, seed: int = 5, rate: float = 20.0):
    from .serve import poisson_workload

    return poisson_workload(seed=seed, rate=rate, duration=seconds)
