"""GPU parity of the 14-point primal BCC stepper (gpe_primal_*, through the C
ABI) against the pinned CPU oracle (oracle/primal_oracle.c, pinned by
tests/test_primal_pins.py) on the same seeded inputs.

Bars: FP64 real-time leapfrog steps are bit-identical to the oracle for every
tile shape, ragged grids included; FP64 imaginary time and observables match
to <= 1e-12 relative (grid-wide reductions in another order); FP32 matches
the FP64 oracle to the tolerance written in the test.  At the 128^3 bench
size the leapfrog is still compared bit for bit (the C oracle takes well
under a second per step there).
"""
import math

import numpy as np
import pytest

import vaikkari_amd as va
from oracle_lib import PrimalOracle

pytestmark = pytest.mark.gpu

SHAPES = [(37, 19, 11), (64, 32, 6), (96, 8, 5), (5, 3, 2)]
# the z-march tiles with buffer-resource addressing (26.. in the tile table)
BUF_TILES = list(range(26, va.primal_tiles() + 1))
# degenerate grids: one cell, single rows and columns
TINY = [(1, 1, 1), (2, 1, 3), (17, 1, 1), (1, 9, 2)]


def announce(msg):
    print(f"[gpu-case] {msg}", flush=True)


@pytest.fixture(scope="module")
def po():
    return PrimalOracle()


def random_fields(dims, seed, amp=0.05):
    """Fields O(amp) on every site, boundary layers included."""
    rng = np.random.default_rng(seed)
    n = va.sites_of(dims)
    even = amp * rng.uniform(-1, 1, 2 * n)
    odd = even + 0.01 * amp * rng.uniform(-1, 1, 2 * n)
    return even, odd


def gaussian(dims, origin, a, sigma=1.5):
    xyz = va.primal_positions(dims, origin, a)
    r2 = (xyz ** 2).sum(axis=1)
    psi = np.exp(-r2 / (2 * sigma ** 2)) * np.exp(0.3j * xyz[:, 0])
    psi /= math.sqrt(va.site_measure(a) * np.sum(np.abs(psi) ** 2))
    return psi.view(np.float64).copy()


def setup(dims, a, dt, g, ratio):
    origin = va.primal_centered_origin(dims, a)
    V = va.primal_trap_potential(dims, origin, a, ratio)
    coef = va.primal_coefficients(a, dt, g)
    return origin, V, coef


@pytest.mark.parametrize("dims", SHAPES)
@pytest.mark.parametrize("tile", range(1, va.primal_tiles() + 1))
def test_fp64_leapfrog_bit_exact(po, dims, tile):
    a, dt, g, ratio = 0.35, 0.004, 1.3, 1.3
    origin, V, coef = setup(dims, a, dt, g, ratio)
    even, odd = random_fields(dims, seed=11 + tile)
    announce(f"primal fp64 leapfrog dims={dims} tile={tile}")
    with va.PrimalSolver(dims, a, dt, g, origin, ratio, "fp64", "real", tile) as s:
        s.set_state(even, odd)
        s.step(3)
        ge, go = s.get_state()
    re, ro = po.leapfrog(even, odd, V * dt, dims, coef, 3)
    assert np.array_equal(ge, re), f"max diff {np.max(np.abs(ge - re))}"
    assert np.array_equal(go, ro), f"max diff {np.max(np.abs(go - ro))}"


@pytest.mark.parametrize("dims", TINY)
@pytest.mark.parametrize("tile", [0, 1, 2, 10, 13, 14, 19, 21, 25])
def test_fp64_degenerate_grids_bit_exact(po, dims, tile):
    """Grids of one cell or one row: every neighbour of the edge lanes is halo."""
    a, dt, g, ratio = 0.35, 0.004, 1.3, 1.3
    origin, V, coef = setup(dims, a, dt, g, ratio)
    even, odd = random_fields(dims, seed=31 + tile)
    announce(f"primal fp64 degenerate dims={dims} tile={tile}")
    with va.PrimalSolver(dims, a, dt, g, origin, ratio, "fp64", "real", tile) as s:
        s.set_state(even, odd)
        s.step(4)
        ge, go = s.get_state()
    re, ro = po.leapfrog(even, odd, V * dt, dims, coef, 4)
    assert np.array_equal(ge, re) and np.array_equal(go, ro)


@pytest.mark.parametrize("dims", SHAPES[:3])
@pytest.mark.parametrize("tile", [1, 2, 13, 21] + BUF_TILES[::5])
def test_fp64_imaginary_time(po, dims, tile):
    a, dt, g, ratio = 0.35, 0.004, 1.3, 1.3
    origin, V, coef = setup(dims, a, dt, g, ratio)
    psi = gaussian(dims, origin, a)
    announce(f"primal fp64 imag dims={dims} tile={tile}")
    with va.PrimalSolver(dims, a, dt, g, origin, ratio, "fp64", "imag", tile) as s:
        s.set_state(psi)
        s.step(6)
        got = s.get_state(odd=False)
        obs = s.observables(0)
    want, N = po.imag_steps(psi, V * dt, dims, coef, va.site_measure(a), 6)
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))
    assert abs(obs["N"] - 1.0) < 1e-12


@pytest.mark.parametrize("dims", SHAPES[:3])
def test_fp64_observables(po, dims):
    a, dt, g, ratio = 0.35, 0.004, 1.3, 1.3
    origin, V, _ = setup(dims, a, dt, g, ratio)
    psi = gaussian(dims, origin, a)
    even, _ = random_fields(dims, seed=5)
    announce(f"primal fp64 observables dims={dims}")
    with va.PrimalSolver(dims, a, dt, g, origin, ratio, "fp64", "real") as s:
        s.set_state(psi, even)
        for field, f in [(0, psi), (1, even)]:
            got = s.observables(field)
            want = po.observables(f, V, dims, a, g)
            for q, key in enumerate("NKPI"):
                assert abs(got[key] - want[q]) <= 1e-12 * abs(want[q]) + 1e-300, (key, got[key], want[q])


@pytest.mark.parametrize("tile", [0] + BUF_TILES[::5])
def test_fp32_observables_and_imaginary_time(po, tile):
    """FP32 fields: the observables reduce in FP64, so on FP32-rounded input they
    match the oracle to <= 1e-12 relative; 20 imaginary-time steps stay within
    1e-5 relative L2 of the FP64 oracle and end normalised."""
    dims = (37, 19, 11)
    a, dt, g, ratio = 0.35, 0.004, 1.3, 1.3
    origin, V, coef = setup(dims, a, dt, g, ratio)
    psi = gaussian(dims, origin, a).astype(np.float32).astype(np.float64)
    announce(f"primal fp32 observables + imag tile={tile}")
    with va.PrimalSolver(dims, a, dt, g, origin, ratio, "fp32", "imag", tile) as s:
        s.set_state(psi)
        got = s.observables(0)
        want = po.observables(psi, V, dims, a, g)
        for q, key in enumerate("NKPI"):
            assert abs(got[key] - want[q]) <= 1e-12 * abs(want[q]), (key, got[key], want[q])
        s.step(20)
        out = s.get_state(odd=False)
        assert abs(s.observables(0)["N"] - 1.0) < 1e-6
    ref, _ = po.imag_steps(psi, V * dt, dims, coef, va.site_measure(a), 20)
    assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < 1e-5


def test_init_odd_phase():
    dims, a, dt, mu = (37, 19, 11), 0.35, 0.004, 7.25
    origin = va.primal_centered_origin(dims, a)
    psi = gaussian(dims, origin, a)
    announce("primal init_odd")
    with va.PrimalSolver(dims, a, dt, 0.0, origin, 1.0, "fp64", "real") as s:
        s.set_state(psi)
        s.init_odd(mu)
        _, odd = s.get_state()
    ph = (-0.5 * dt) * mu
    want = psi.view(np.complex128) * complex(math.cos(ph), -math.sin(ph))
    np.testing.assert_allclose(odd.view(np.complex128), want, rtol=1e-15, atol=1e-18)


@pytest.mark.parametrize("tile", [1, 2, 13, 21] + BUF_TILES)
def test_fp32_leapfrog_tolerance(po, tile):
    """FP32 storage and stencil (difference form) against the FP64 oracle:
    relative L2 error <= 1e-5 after 3 leapfrog steps."""
    dims = (37, 19, 11)
    a, dt, g, ratio = 0.35, 0.004, 1.3, 1.3
    origin, V, coef = setup(dims, a, dt, g, ratio)
    even, odd = random_fields(dims, seed=23)
    even32 = even.astype(np.float32).astype(np.float64)
    odd32 = odd.astype(np.float32).astype(np.float64)
    announce(f"primal fp32 leapfrog tile={tile}")
    with va.PrimalSolver(dims, a, dt, g, origin, ratio, "fp32", "real", tile) as s:
        s.set_state(even32, odd32)
        s.step(3)
        ge, go = s.get_state()
    re, ro = po.leapfrog(even32, odd32, V * dt, dims, coef, 3)
    assert np.linalg.norm(ge - re) / np.linalg.norm(re) < 1e-5
    assert np.linalg.norm(go - ro) / np.linalg.norm(ro) < 1e-5


def test_fp64_bench_size_bit_exact(po):
    """128^3 cells (4.19M sites), the bench workload's shape: two leapfrog steps
    equal the oracle bit for bit under the default tile, the plain-order
    tile and a z-march tile, and the imaginary-time norm stays 1."""
    dims = (128, 128, 128)
    a, dt, g, ratio = 0.25, 0.002, 1.0, 1.0
    origin, V, coef = setup(dims, a, dt, g, ratio)
    psi = gaussian(dims, origin, a, sigma=3.0)
    odd = (psi.view(np.complex128) * complex(math.cos(-0.5 * dt * 5.0), -math.sin(-0.5 * dt * 5.0))).view(
        np.float64).copy()
    re, ro = po.leapfrog(psi, odd, V * dt, dims, coef, 2)
    for tile in (1, 7, 13, 21):
        announce(f"primal fp64 128^3 leapfrog tile={tile}")
        with va.PrimalSolver(dims, a, dt, g, origin, ratio, "fp64", "real", tile) as s:
            s.set_state(psi, odd)
            s.step(2)
            ge, go = s.get_state()
        assert np.array_equal(ge, re) and np.array_equal(go, ro)
    announce("primal fp64 128^3 imag")
    with va.PrimalSolver(dims, a, dt, g, origin, ratio, "fp64", "imag") as s:
        s.set_state(psi)
        s.step(10)
        assert abs(s.observables(0)["N"] - 1.0) < 1e-12
