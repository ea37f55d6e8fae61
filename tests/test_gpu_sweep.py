"""Seeded random sweep of FP64 parity: grid shapes drawn at random (every axis 1..70, ragged
against the 64-lane rows, the 4/8-wave blocks and the 8 XCD bands), every gather and z-march tile
shape and both potential kinds, random batches of real-time instances under random tiles and group
sizes, and the 14-point primal stepper's 25 tile shapes on random shapes.  Each case is stepped on
the GPU through the C ABI and compared bit for bit with the CPU oracles (oracle/gpe_oracle.c and
oracle/primal_oracle.c, pinned by tests/test_oracle_pins.py and tests/test_primal_pins.py) on the
same inputs."""
import math
import os
import sys

import numpy as np
import pytest

import vaikkari_amd as va
from oracle_lib import Oracle, PrimalOracle
from vaikkari_amd.batch import Batch, instance_config

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import output_oracle as oo  # noqa: E402

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"[gpu-case] {msg}", flush=True)


@pytest.fixture(scope="module")
def orc():
    return Oracle()


def _shapes(seed, count, max_cells=60000):
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < count:
        d = tuple(int(x) for x in rng.integers(1, 71, 3))
        if d[0] * d[1] * d[2] <= max_cells:
            out.append(d)
    return out


def _fields(dims, h, ratio, seed, amp=0.05):
    rng = np.random.default_rng(seed)
    b = h / math.sqrt(8)
    origin = va.centered_origin(dims, b)
    n = va.nodes_of(dims)
    even = amp * rng.uniform(-1, 1, 2 * n)
    odd = even + 0.01 * amp * rng.uniform(-1, 1, 2 * n)
    return b, origin, even, odd, va.trap_potential(dims, origin, b, ratio)


SHAPES = _shapes(2024, 24)
TILES = [(v["kernel"], v["tile"]) for v in va.variants() if v["kernel"] in ("gather", "zmarch")]
# the fused odd+even pass; its 16-row FP64 tile exceeds the LDS budget and is rejected by the library
FUSED_FP64 = [v["tile"] for v in va.variants() if v["kernel"] == "fused" and v["waves"] != 16]


@pytest.mark.parametrize("case", range(len(SHAPES)))
def test_random_shapes_every_tile_bit_exact(orc, case):
    dims = SHAPES[case]
    h, g, dt, ratio = 0.45, 400.0, 8e-4, 1.7
    b, origin, even, odd, V = _fields(dims, h, ratio, seed=100 + case)
    lf, lf0, gdt = va.leapfrog_coefs(b, dt, g)
    oe, oo = orc.leapfrog(even, odd, V * dt, dims, lf, lf0, gdt, 5)
    for q, (kernel, tile) in enumerate(TILES):
        potential = "array" if (case + q) % 2 else "trap"
        announce(f"sweep {dims} {kernel} tile={tile} {potential}")
        s = va.Solver(dims, b, dt, g, origin, "fp64", "real", potential, ratio, kernel, tile)
        if potential == "array":
            s.set_potential(V)
        s.set_state(even, odd)
        s.step(5)
        ge, go = s.get_state()
        s.close()
        assert np.array_equal(ge, oe) and np.array_equal(go, oo), (dims, kernel, tile, potential)


@pytest.mark.parametrize("case", range(2, len(SHAPES), 3))
def test_random_shapes_fused_every_tile_bit_exact(orc, case):
    dims = SHAPES[case]
    h, g, dt, ratio = 0.45, 400.0, 8e-4, 1.7
    b, origin, even, odd, V = _fields(dims, h, ratio, seed=400 + case)
    oe, oo = orc.leapfrog(even, odd, V * dt, dims, *va.leapfrog_coefs(b, dt, g), 6)
    for tile in FUSED_FP64:
        announce(f"sweep fused {dims} tile={tile}")
        s = va.Solver(dims, b, dt, g, origin, "fp64", "real", "trap", ratio, "fused", tile)
        s.set_state(even, odd)
        s.step(6)
        ge, go = s.get_state()
        s.close()
        assert np.array_equal(ge, oe) and np.array_equal(go, oo), (dims, tile)


@pytest.mark.parametrize("case", range(0, len(SHAPES), 4))
def test_random_shapes_fused_observables(orc, case):
    """gpe_step_observe on the random shapes under every gather tile: N, K, P, I summed inside the
    last step's first half-update equal the oracle's observables of its field after nsteps - 1
    steps to 1e-12 relative (FP64 sums in another order); the steps stay bit-identical."""
    dims = SHAPES[case]
    h, g, dt, ratio = 0.45, 400.0, 8e-4, 1.7
    b, origin, even, odd, V = _fields(dims, h, ratio, seed=600 + case)
    oe, _ = orc.leapfrog(even, odd, V * dt, dims, *va.leapfrog_coefs(b, dt, g), 4)
    want = orc.observables(oe, V, dims, b, g)
    fe, fo = orc.leapfrog(even, odd, V * dt, dims, *va.leapfrog_coefs(b, dt, g), 5)
    for kernel, tile in TILES:
        if kernel != "gather":
            continue
        announce(f"sweep observe {dims} tile={tile}")
        s = va.Solver(dims, b, dt, g, origin, "fp64", "real", "trap", ratio, kernel, tile)
        s.set_state(even, odd)
        got = s.step_observe(5)
        ge, go = s.get_state()
        s.close()
        for k, r in zip("NKPI", want):
            assert got[k] == pytest.approx(r, rel=1e-12, abs=1e-300), (dims, tile, k)
        assert np.array_equal(ge, fe) and np.array_equal(go, fo), (dims, tile)


def _rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("case", range(0, len(SHAPES), 3))
def test_random_shapes_imag_and_fp32_every_tile(orc, case):
    """The same shapes in imaginary time (FP64, 6 Euler steps with renormalisation: <= 1e-12
    relative L2, the norm is a grid-wide sum in another order) and in FP32 real time (10 steps of a
    field of full amplitude everywhere: <= 2e-6 relative L2 of the FP64 oracle)."""
    dims = SHAPES[case]
    h, g, dt, ratio = 0.45, 400.0, 8e-4, 1.7
    b, origin, even, odd, V = _fields(dims, h, ratio, seed=200 + case)
    lf, lf0, gdt = va.leapfrog_coefs(b, dt, g)
    dtau = 0.5 / (3.0 / b ** 2 + V.max())
    ilf, ilf0, igdt = va.leapfrog_coefs(b, dtau, g)
    psi = even / math.sqrt(va.field_norm(dims, b, even))
    oi, _ = orc.imag_steps(psi, V * dtau, dims, ilf, ilf0, igdt, b, 6)
    oe, oo = orc.leapfrog(even, odd, V * dt, dims, lf, lf0, gdt, 10)
    for kernel, tile in TILES:
        announce(f"sweep imag+fp32 {dims} {kernel} tile={tile}")
        s = va.Solver(dims, b, dtau, g, origin, "fp64", "imag", "trap", ratio, kernel, tile)
        s.set_state(psi)
        s.step(6)
        gi = s.get_state(odd=False)
        s.close()
        assert _rel_l2(gi, oi) <= 1e-12, (dims, kernel, tile)
        s = va.Solver(dims, b, dt, g, origin, "fp32", "real", "trap", ratio, kernel, tile)
        s.set_state(even, odd)
        s.step(10)
        ge, go = s.get_state()
        s.close()
        assert _rel_l2(ge, oe) < 2e-6 and _rel_l2(go, oo) < 2e-6, (dims, kernel, tile)


@pytest.mark.parametrize("seed", [7, 8, 9, 10, 11, 12, 13, 14, 15, 16])
def test_random_batch_mix_bit_exact(orc, seed):
    rng = np.random.default_rng(seed)
    # waves 3, 6 and 12: one thread per cell and group of 4 classes
    tile = (int(rng.choice([1, 2, 4, 8, 3, 6, 12])), bool(rng.integers(2)), [False, True, 2][int(rng.integers(3))])
    group_bytes = int(rng.choice([0, 1, 1 << 20, 640 << 20]))
    chunk = int(rng.integers(3, 12))
    announce(f"sweep batch seed={seed} tile={tile} group_bytes={group_bytes} chunk={chunk}")
    B = Batch("fp64")
    B.set_tile(tile)
    B.set_group_bytes(group_bytes)
    ids, refs = [], []
    for q in range(6):
        dims = tuple(int(x) for x in rng.integers(1, 40, 3))
        h, ratio = float(rng.uniform(0.3, 0.6)), float(rng.uniform(0.5, 3.0))
        g, dt, steps = float(rng.uniform(50, 900)), 5e-4, int(rng.integers(5, 30))
        b, origin, even, odd, V = _fields(dims, h, ratio, seed=1000 * seed + q)
        i = B.add(instance_config(dims, h, g, dt, steps, "real", ratio, init="host", origin=origin))
        B.set_state(i, even, odd)
        ids.append((i, steps))
        refs.append(orc.leapfrog(even, odd, V * dt, dims, *va.leapfrog_coefs(b, dt, g), steps))
    B.run(chunk, -(-30 // chunk))
    for (i, steps), (oe, oo) in zip(ids, refs):
        assert B.info(i)["steps_done"] == steps
        ge, go = B.get_state(i)
        assert np.array_equal(ge, oe) and np.array_equal(go, oo), (seed, i)
    B.close()


@pytest.mark.parametrize("case", range(1, len(SHAPES), 4))
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_random_shapes_output_bit_exact(case, precision):
    """Volume map (all classes and the reference's class-0 form) and slice pictures of every kind of
    layer (halo, first, middle, last) on the random shapes, against the output oracle on the
    downloaded field."""
    dims = SHAPES[case]
    announce(f"sweep output {dims} {precision}")
    rng = np.random.default_rng(500 + case)
    b = 0.25 / math.sqrt(8)
    s = va.Solver(dims, b, 1e-4, 100.0, None, precision, "real", "trap", 1.0)
    even = 0.08 * rng.standard_normal(2 * va.nodes_of(dims))
    s.set_state(even, even)
    e = s.get_state(odd=False)
    for class0 in (False, True):
        np.testing.assert_array_equal(s.volume_map(3.0e5, reference_class0=class0),
                                      oo.volume_map(e, dims, 3.0e5, class0))
    for k in sorted({0, 1, dims[2] // 2 + 1, dims[2], dims[2] + 1}):
        for middle in (False, True):
            np.testing.assert_array_equal(s.slice_picture(k, 5.0, middle),
                                          oo.colour_bytes(oo.slice_colours(e, dims, k, 5.0), middle))
    s.close()


# batch ground-state tiles and the single-solver variant with the same logical tiles (4 waves,
# one layer / 8 / 4 layers per tile), hence the same norm summation order
# the single solver's tile whose arithmetic and block partials a batched ground state of
# zlayers layers per tile repeats: gather {4 waves, y-bands} (tile 11; the single solver's own
# default depends on the grid), z-march 8 or 4 layers
IMAG_TILES = {0: ("gather", 11), 8: ("zmarch", 1), 4: ("zmarch", 6)}


@pytest.mark.parametrize("seed", [21, 22, 23])
def test_random_batch_ground_states_match_solver(orc, seed):
    """Ground states (one-layer, 8-layer and 4-layer tiles) mixed with real-time instances under a
    random group size: each ground state equals the single solver's run of the same tiles bit for
    bit, each real-time instance the oracle."""
    rng = np.random.default_rng(seed)
    B = Batch("fp64")
    B.set_group_bytes(int(rng.choice([0, 1, 1 << 20])))
    cases = []
    for q in range(6):
        dims = tuple(int(x) for x in rng.integers(2, 36, 3))
        h, ratio, g = float(rng.uniform(0.3, 0.6)), float(rng.uniform(0.5, 3.0)), float(rng.uniform(50, 900))
        b, origin, even, odd, V = _fields(dims, h, ratio, seed=3000 * seed + q)
        if q % 3 == 2:
            dt, steps = 5e-4, int(rng.integers(5, 25))
            i = B.add(instance_config(dims, h, g, dt, steps, "real", ratio, init="host", origin=origin))
            B.set_state(i, even, odd)
            cases.append(("real", i, dims, b, origin, even, odd, V, g, dt, steps, ratio, None))
        else:
            zl = int(rng.choice([0, 4, 8]))
            dtau = 0.5 / (3.0 / b ** 2 + V.max())
            steps = int(rng.integers(5, 25))
            psi = even / math.sqrt(va.field_norm(dims, b, even))
            i = B.add(instance_config(dims, h, g, dtau, steps, "imag", ratio, init="host", origin=origin,
                                      zlayers=zl))
            B.set_state(i, psi)
            cases.append(("imag", i, dims, b, origin, psi, None, V, g, dtau, steps, ratio, zl))
    announce(f"sweep batch ground states seed={seed}")
    B.run(7, 4)
    for mode, i, dims, b, origin, e, o, V, g, dt, steps, ratio, zl in cases:
        assert B.info(i)["steps_done"] == steps
        ge, go = B.get_state(i)
        if mode == "real":
            oe, oo_ = orc.leapfrog(e, o, V * dt, dims, *va.leapfrog_coefs(b, dt, g), steps)
            assert np.array_equal(ge, oe) and np.array_equal(go, oo_), (seed, i)
        else:
            kernel, tile = IMAG_TILES[zl]
            s = va.Solver(dims, b, dt, g, origin, "fp64", "imag", "trap", ratio, kernel, tile)
            s.set_state(e)
            s.step(steps)
            se = s.get_state(odd=False)
            s.close()
            assert np.array_equal(ge, se), (seed, i, zl)
    B.close()


PRIMAL_SHAPES = _shapes(77, 8, max_cells=40000)


@pytest.mark.parametrize("case", range(len(PRIMAL_SHAPES)))
def test_random_shapes_primal_every_tile_bit_exact(case):
    po = PrimalOracle()
    dims = PRIMAL_SHAPES[case]
    a, dt, g, ratio = 0.35, 0.004, 1.3, 1.3
    origin = va.primal_centered_origin(dims, a)
    V = va.primal_trap_potential(dims, origin, a, ratio)
    coef = va.primal_coefficients(a, dt, g)
    rng = np.random.default_rng(300 + case)
    n = va.sites_of(dims)
    even = 0.05 * rng.uniform(-1, 1, 2 * n)
    odd = even + 0.0005 * rng.uniform(-1, 1, 2 * n)
    re, ro = po.leapfrog(even, odd, V * dt, dims, coef, 3)
    for tile in range(1, va.primal_tiles() + 1):
        announce(f"sweep primal {dims} tile={tile}")
        with va.PrimalSolver(dims, a, dt, g, origin, ratio, "fp64", "real", tile) as s:
            s.set_state(even, odd)
            s.step(3)
            ge, go = s.get_state()
        assert np.array_equal(ge, re) and np.array_equal(go, ro), (dims, tile)

