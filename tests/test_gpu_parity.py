"""GPU parity: the HIP stepper (through the C ABI) against the pinned CPU
oracle on the same seeded inputs.

Bars: FP64 real-time updates are bit-identical to the oracle (same
operation sequence, FMA contraction off); FP64 paths that reduce over the
grid (imaginary-time renormalisation, observables) match to <= 1e-12
relative; FP32-bulk matches the FP64 oracle to the tolerance written in
each test.
"""
import math
import os
import struct

import numpy as np
import pytest

import vaikkari_amd as va
from vaikkari_amd._lib import EINVAL, GpeError
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

KERNELS = ["gather", "zmarch", "fused"]


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def case(dims, h, g, dt, ratio, seed, sigma=1.5, noise=1e-3, mu=10.0):
    b = h / math.sqrt(8)
    origin = va.centered_origin(dims, b)
    even = va.gaussian(dims, origin, b, sigma, noise, seed)
    odd = va.phase_shift(dims, even, mu, dt)
    V = va.trap_potential(dims, origin, b, ratio)
    return b, origin, even, odd, V


def random_case(dims, h, ratio, seed, amp=0.05):
    """A field that is O(amp) at every node, boundary layers included, so that an error
    anywhere shows in a relative norm."""
    rng = np.random.default_rng(seed)
    b = h / math.sqrt(8)
    origin = va.centered_origin(dims, b)
    n = va.nodes_of(dims)
    even = amp * rng.uniform(-1, 1, 2 * n)
    odd = even + 0.01 * amp * rng.uniform(-1, 1, 2 * n)
    V = va.trap_potential(dims, origin, b, ratio)
    return b, origin, even, odd, V


def announce(msg):
    print(f"[gpu-case] {msg}", flush=True)


@pytest.fixture(scope="module")
def orc():
    return Oracle()


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("potential", ["trap", "array"])
@pytest.mark.parametrize("dims", [(7, 5, 9), (65, 3, 2), (1, 1, 1), (3, 66, 11)])
def test_fp64_real_time_bit_exact(orc, kernel, potential, dims):
    announce(f"fp64 real {kernel} {potential} {dims}")
    h, g, dt, ratio = 0.5, 300.0, 1e-3, 1.5
    b, origin, even, odd, V = case(dims, h, g, dt, ratio, seed=11, sigma=1.0)
    s = va.Solver(dims, b, dt, g, origin, "fp64", "real", potential, ratio, kernel)
    if potential == "array":
        s.set_potential(V)
    s.set_state(even, odd)
    s.step(25)
    ge, go = s.get_state()
    lf, lf0, gdt = va.leapfrog_coefs(b, dt, g)
    oe, oo = orc.leapfrog(even, odd, V * dt, dims, lf, lf0, gdt, 25)
    np.testing.assert_array_equal(ge, oe)
    np.testing.assert_array_equal(go, oo)
    s.close()


def _fused_tiles():
    return [v["tile"] for v in va.variants() if v["kernel"] == "fused"]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("tile", _fused_tiles())
def test_fused_every_tile(orc, precision, tile):
    """Every tile shape of the fused odd+even pass on a ragged grid with a field of full
    amplitude at every boundary: FP64 bit-identical to the oracle, FP32 within 2e-6 relative L2
    after 30 steps."""
    dims = (67, 13, 37)
    shape = [v for v in va.variants() if v["kernel"] == "fused" and v["tile"] == tile][0]
    announce(f"fused tile {tile} {shape} {precision}")
    h, g, dt, ratio = 0.4, 500.0, 1e-3, 1.2
    b, origin, even, odd, V = random_case(dims, h, ratio, seed=21)
    if precision == "fp64" and shape["waves"] == 16:
        # 16 rows of FP64 exchange slots exceed the LDS: refused at creation and by
        # set_variant, and a refused set_variant leaves the state and the variant as they were
        with pytest.raises(GpeError) as ei:
            va.Solver(dims, b, dt, g, origin, "fp64", "real", "trap", ratio, "fused", tile)
        assert ei.value.status == EINVAL
        s = va.Solver(dims, b, dt, g, origin, "fp64", "real", "trap", ratio, "fused", 1)
        s.set_state(even, odd)
        s.step(3)
        before = s.get_state()
        with pytest.raises(GpeError) as ei:
            s.set_variant("fused", tile)
        assert ei.value.status == EINVAL
        after = s.get_state()
        np.testing.assert_array_equal(after[0], before[0])
        np.testing.assert_array_equal(after[1], before[1])
        s.step(27)
        ge, go = s.get_state()
        lf, lf0, gdt = va.leapfrog_coefs(b, dt, g)
        oe, oo = orc.leapfrog(even, odd, V * dt, dims, lf, lf0, gdt, 30)
        np.testing.assert_array_equal(ge, oe)
        np.testing.assert_array_equal(go, oo)
        s.close()
        return
    s = va.Solver(dims, b, dt, g, origin, precision, "real", "trap", ratio, "fused", tile)
    s.set_state(even, odd)
    s.step(30)
    ge, go = s.get_state()
    lf, lf0, gdt = va.leapfrog_coefs(b, dt, g)
    oe, oo = orc.leapfrog(even, odd, V * dt, dims, lf, lf0, gdt, 30)
    if precision == "fp64":
        np.testing.assert_array_equal(ge, oe)
        np.testing.assert_array_equal(go, oo)
    else:
        assert rel_l2(ge, oe) < 2e-6 and rel_l2(go, oo) < 2e-6
    s.close()


def _row_tiles():
    return [v["tile"] for v in va.variants() if v["kernel"] == "fused" and v["row"]]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("tile", _row_tiles())
def test_fused_row_tiles_array_potential(orc, precision, tile):
    """The one-wave-per-row fused tiles with a streamed potential, on a grid of three x-strips
    (the last one 1 column wide), rows that do not fill the last block and a last z-chunk of
    one layer: FP64 bit-identical to the oracle, FP32 within 2e-6 relative L2 after 12 steps."""
    dims = (125, 11, 33)
    announce(f"fused row tile {tile} array {precision} {dims}")
    h, g, dt, ratio = 0.4, 400.0, 1e-3, 0.8
    b, origin, even, odd, V = random_case(dims, h, ratio, seed=41)
    s = va.Solver(dims, b, dt, g, origin, precision, "real", "array", ratio, "fused", tile)
    s.set_potential(V)
    s.set_state(even, odd)
    s.step(12)
    ge, go = s.get_state()
    lf, lf0, gdt = va.leapfrog_coefs(b, dt, g)
    oe, oo = orc.leapfrog(even, odd, V * dt, dims, lf, lf0, gdt, 12)
    if precision == "fp64":
        np.testing.assert_array_equal(ge, oe)
        np.testing.assert_array_equal(go, oo)
    else:
        assert rel_l2(ge, oe) < 2e-6 and rel_l2(go, oo) < 2e-6
    s.close()


@pytest.mark.parametrize("variant", [(v["kernel"], v["tile"]) for v in va.variants()
                                     if not (v["kernel"] == "fused" and v["waves"] == 16)])
def test_every_variant_fp64_bit_exact(orc, variant):
    """Every (kernel family, tile shape) the autotuner may pick: FP64 bit-identical to the
    oracle on a ragged grid with full-amplitude boundaries, real and imaginary time."""
    kernel, tile = variant
    dims = (66, 7, 11)
    announce(f"variant {kernel} {tile}")
    b, origin, even, odd, V = random_case(dims, 0.45, 1.3, seed=31)
    s = va.Solver(dims, b, 1e-3, 250.0, origin, "fp64", "real", "trap", 1.3, kernel, tile)
    s.set_state(even, odd)
    s.step(6)
    ge, go = s.get_state()
    lf, lf0, gdt = va.leapfrog_coefs(b, 1e-3, 250.0)
    oe, oo = orc.leapfrog(even, odd, V * 1e-3, dims, lf, lf0, gdt, 6)
    np.testing.assert_array_equal(ge, oe)
    np.testing.assert_array_equal(go, oo)
    s.close()


@pytest.mark.parametrize("kernel", KERNELS)
def test_fp32_random_field_all_kernels(orc, kernel):
    """FP32 on a field of full amplitude everywhere (boundary layers, rows and columns of
    every tile included): within 2e-6 relative L2 of the FP64 oracle after 20 steps."""
    dims = (70, 19, 23)
    b, origin, even, odd, V = random_case(dims, 0.35, 1.1, seed=8)
    s = va.Solver(dims, b, 1e-3, 300.0, origin, "fp32", "real", "trap", 1.1, kernel)
    s.set_state(even, odd)
    s.step(20)
    ge, go = s.get_state()
    lf, lf0, gdt = va.leapfrog_coefs(b, 1e-3, 300.0)
    oe, oo = orc.leapfrog(even, odd, V * 1e-3, dims, lf, lf0, gdt, 20)
    assert rel_l2(ge, oe) < 2e-6 and rel_l2(go, oo) < 2e-6
    s.close()


@pytest.mark.parametrize("kernel", ["auto", "fused"])
def test_fp64_small_config_1000_steps(orc, kernel):
    """32^3 cells (393,216 nodes), h = 0.25, g = 500, dt = 1e-3, 1000 steps, seed 54363:
    relative L2 <= 1e-12 on both leapfrog levels (observed: 0, bit-identical)."""
    dims, h, g, dt = (32, 32, 32), 0.25, 500.0, 1e-3
    announce(f"fp64 32^3 x 1000 {kernel}")
    b, origin, even, odd, V = case(dims, h, g, dt, 1.0, seed=54363)
    s = va.Solver(dims, b, dt, g, origin, "fp64", "real", "trap", 1.0, kernel)
    s.set_state(even, odd)
    s.step(1000)
    ge, go = s.get_state()
    lf, lf0, gdt = va.leapfrog_coefs(b, dt, g)
    oe, oo = orc.leapfrog(even, odd, V * dt, dims, lf, lf0, gdt, 1000)
    assert rel_l2(ge, oe) <= 1e-12 and rel_l2(go, oo) <= 1e-12
    np.testing.assert_array_equal(ge, oe)
    s.close()


@pytest.mark.parametrize("kernel", KERNELS)
def test_fp64_imaginary_time(orc, kernel):
    dims, h, g = (12, 10, 14), 0.4, 200.0
    announce(f"fp64 imag {kernel}")
    b, origin, even, _, V = case(dims, h, g, 1e-3, 2.0, seed=5, sigma=1.2)
    dtau = 0.5 / (3.0 / b ** 2 + V.max() + g * np.max(even[0::2] ** 2 + even[1::2] ** 2))
    s = va.Solver(dims, b, dtau, g, origin, "fp64", "imag", "trap", 2.0, kernel)
    s.set_state(even)
    s.step(60)
    ge = s.get_state(odd=False)
    lf, lf0, gdt = va.leapfrog_coefs(b, dtau, g)
    oe, _ = orc.imag_steps(even, V * dtau, dims, lf, lf0, gdt, b, 60)
    assert rel_l2(ge, oe) <= 1e-12
    assert va.field_norm(dims, b, ge) == pytest.approx(1.0, abs=1e-13)
    s.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("potential", ["trap", "array"])
def test_observables(orc, precision, potential):
    dims, h, g = (9, 8, 10), 0.45, 700.0
    announce(f"observables {precision} {potential}")
    b, origin, even, odd, V = case(dims, h, g, 1e-3, 1.3, seed=9, sigma=1.1)
    s = va.Solver(dims, b, 1e-3, g, origin, precision, "real", potential, 1.3)
    if potential == "array":
        s.set_potential(V)
    s.set_state(even, odd)
    ref = orc.observables(even, V, dims, b, g)
    got = s.observables(0)
    tol = 1e-12 if precision == "fp64" else 2e-6
    for k, r in zip("NKPI", ref):
        assert got[k] == pytest.approx(r, rel=tol), k
    s.close()


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("potential", ["trap", "array"])
def test_step_observe_fused_reduction(orc, kernel, precision, potential):
    """gpe_step_observe: the observables reduced inside the last step's odd
    half-update equal gpe_observables of the even field after nsteps - 1
    steps (FP64 reductions in another block order: <= 1e-12 relative), and
    the steps equal gpe_step bit for bit."""
    dims, h, g, dt = (9, 8, 10), 0.45, 700.0, 1e-3
    announce(f"step_observe {kernel} {precision} {potential}")
    b, origin, even, odd, V = case(dims, h, g, dt, 1.3, seed=13, sigma=1.1)

    def solver():
        s = va.Solver(dims, b, dt, g, origin, precision, "real", potential, 1.3, kernel)
        if potential == "array":
            s.set_potential(V)
        s.set_state(even, odd)
        return s

    a, c = solver(), solver()
    got = a.step_observe(7)
    ea, oa = a.get_state()
    c.step(6)
    ref = c.observables(0)
    c.step(1)
    ec, oc = c.get_state()
    for k in "NKPI":
        assert got[k] == pytest.approx(ref[k], rel=1e-12), k
    assert np.array_equal(ea, ec) and np.array_equal(oa, oc)
    if precision == "fp64" and kernel == "gather" and potential == "trap":
        # and against the oracle: N, K, P, I of the oracle's field after 6 steps
        oe, _ = orc.leapfrog(even, odd, V * dt, dims, *va.leapfrog_coefs(b, dt, g), 6)
        want = orc.observables(oe, V, dims, b, g)
        for k, r in zip("NKPI", want):
            assert got[k] == pytest.approx(r, rel=1e-12), k
        a.set_mode("imag", dt)
        with pytest.raises(va._lib.GpeError) as ei:
            a.step_observe(1)
        assert ei.value.status == va._lib.ENOTSUP
    a.close()
    c.close()


def test_init_odd_matches_host_phase(orc):
    dims = (5, 6, 7)
    b, origin, even, _, V = case(dims, 0.5, 100.0, 1e-3, 1.0, seed=2)
    s = va.Solver(dims, b, 1e-3, 100.0, origin, "fp64", "real", "trap", 1.0)
    s.set_state(even)
    s.init_odd(12.5)
    _, go = s.get_state()
    np.testing.assert_array_equal(go, va.phase_shift(dims, even, 12.5, 1e-3))
    s.close()


def test_fp32_bulk_error_vs_fp64_oracle(orc):
    """FP32 storage + FP32 stencil, 32^3 x 1000 steps: psi relative L2 error vs the FP64
    oracle stated and bounded by 1e-5; the leapfrog invariant drifts by < 1e-5 relative."""
    dims, h, g, dt = (32, 32, 32), 0.25, 500.0, 1e-3
    announce("fp32 32^3 x 1000")
    b, origin, even, odd, V = case(dims, h, g, dt, 1.0, seed=54363)
    s = va.Solver(dims, b, dt, g, origin, "fp32", "real", "trap", 1.0)
    s.set_state(even, odd)
    s.step(1000)
    ge, go = s.get_state()
    lf, lf0, gdt = va.leapfrog_coefs(b, dt, g)
    oe, oo = orc.leapfrog(even, odd, V * dt, dims, lf, lf0, gdt, 1000)
    err = rel_l2(ge, oe)
    print(f"[fp32] rel-L2 psi error after 1000 steps: {err:.3e}")
    assert err <= 1e-5
    q0, q1 = orc.overlap(even, odd), orc.overlap(ge, go)
    assert abs(q1 - q0) <= 1e-5 * abs(q0)
    s.close()


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_deterministic(precision, kernel):
    dims = (33, 17, 20)
    b, origin, even, odd, V = case(dims, 0.4, 900.0, 1e-3, 1.0, seed=4)
    outs = []
    for _ in range(2):
        s = va.Solver(dims, b, 1e-3, 900.0, origin, precision, "real", "trap", 1.0, kernel)
        s.set_state(even, odd)
        s.step(40)
        outs.append(s.get_state())
        s.close()
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    np.testing.assert_array_equal(outs[0][1], outs[1][1])


def test_config2_full_size_bit_exact(orc):
    """The bench workload's grid and physics (128^3 cells, 25,165,824 nodes, h = 0.1,
    g = 2000, trap 0.5(x^2+y^2+4z^2), dt = 2e-4): 100 FP64 leapfrog steps equal the
    oracle bit for bit (the OpenMP oracle takes a few seconds)."""
    dims, h, g, dt = (128, 128, 128), 0.1, 2000.0, 2e-4
    announce("fp64 128^3 x 100 vs oracle")
    b, origin, even, odd, V = case(dims, h, g, dt, 2.0, seed=1, sigma=1.5)
    s = va.Solver(dims, b, dt, g, origin, "fp64", "real", "trap", 2.0)
    s.set_state(even, odd)
    s.step(100)
    ge, go = s.get_state()
    s.close()
    lf, lf0, gdt = va.leapfrog_coefs(b, dt, g)
    oe, oo = orc.leapfrog(even, odd, V * dt, dims, lf, lf0, gdt, 100)
    assert np.array_equal(ge, oe) and np.array_equal(go, oo)


def test_full_size_fp64_invariant():
    """BCC 128^3 cells (25,165,824 nodes), h = 0.1, g = 2000, trap 0.5(x^2+y^2+4z^2),
    dt = 2e-4: the exact leapfrog invariant Re<even,odd> holds to 1e-12 relative over 200 steps
    and the norm stays within the O(dt^2) leapfrog oscillation."""
    dims, h, g, dt = (128, 128, 128), 0.1, 2000.0, 2e-4
    announce("fp64 128^3 invariant")
    b, origin, even, odd, V = case(dims, h, g, dt, 2.0, seed=1, sigma=1.2)
    s = va.Solver(dims, b, dt, g, origin, "fp64", "real", "trap", 2.0)
    s.set_state(even, odd)
    q0 = float(np.dot(even, odd))
    n0 = s.observables(0)["N"]
    s.step(200)
    ge, go = s.get_state()
    q1 = float(np.dot(ge, go))
    assert abs(q1 - q0) <= 1e-12 * abs(q0)
    assert s.observables(0)["N"] == pytest.approx(n0, rel=1e-3)
    assert np.all(np.isfinite(ge))
    s.close()


def write_vortex_state(path, kappa=1, g=400.0, rsize=60, zsize=80, rstep=0.1, zstep=0.1, mu=7.5):
    """A VortexState file in the reference's binary format (VortexState.cpp:74-96) holding
    f(r,z) = r^kappa exp(-(r^2+z^2)/2) (this test's own data)."""
    r = np.arange(rsize) * rstep
    z = np.arange(zsize) * zstep
    f = (r[:, None] ** kappa) * np.exp(-(r[:, None] ** 2 + z[None, :] ** 2) / 2)
    with open(path, "wb") as fh:
        fh.write(struct.pack("<IdIIdddd", kappa, g, rsize, zsize, 0.0, rstep, zstep, mu))
        fh.write(f.astype("<f8").tobytes())


def test_integrate_in_time_drop_in(orc, tmp_path):
    """integrateInTime on a VortexState file: setup + 2 iterations of steps_per_iteration
    leapfrog steps, bit-identical to the oracle's setup and leapfrog."""
    path = str(tmp_path / "vs.dat")
    write_vortex_state(path)
    vs = va.VortexState()
    assert vs.load(path)
    eps = 1e-5 * vs.searchFunctionMax()
    mr, mz = vs.searchMaxR(eps), vs.searchMaxZ(eps)
    bs = 0.15
    minp, maxp = (-mr, -mr, -mz), (mr, mr, mz)
    announce("integrateInTime drop-in")
    rc, run, ge, go = va.integrateInTime(vs, bs, minp, maxp, 0.05, 2, trap_ratio=0.7, return_fields=True)
    assert rc == 0
    info, even, odd, V = va.reference_setup(vs, bs, minp, maxp, 0.05, 0.7)
    assert run == info
    oe, oo = orc.leapfrog(even, odd, V * info["dt"], info["dims"], info["lapfac_dt"], info["lapfac0_dt"],
                          info["g"] * info["dt"], 2 * info["steps_per_iteration"])
    np.testing.assert_array_equal(ge, oe)
    np.testing.assert_array_equal(go, oo)


def test_imag_then_real_workflow(orc):
    """Ground state by imaginary time, vortex imprint, then real time: the drop-in benchmark's
    workflow at 16^3, every stage against the oracle."""
    dims, h, g, ratio = (16, 16, 16), 0.3, 800.0, 2.0
    announce("imag->real workflow")
    b, origin, even, _, V = case(dims, h, g, 1e-3, ratio, seed=3, sigma=1.0)
    dtau = 0.5 / (3.0 / b ** 2 + V.max())
    s = va.Solver(dims, b, dtau, g, origin, "fp64", "imag", "trap", ratio)
    s.set_state(even)
    s.step(50)
    ge = s.get_state(odd=False)
    lf, lf0, gdt = va.leapfrog_coefs(b, dtau, g)
    oe, _ = orc.imag_steps(even, V * dtau, dims, lf, lf0, gdt, b, 50)
    assert rel_l2(ge, oe) <= 1e-12
    obs = s.observables(0)
    ref = orc.observables(oe, V, dims, b, g)
    assert obs["N"] == pytest.approx(ref[0], rel=1e-12)
    mu = (ref[1] + ref[2] + ref[3]) / ref[0]
    assert obs["mu"] == pytest.approx(mu, rel=1e-12)
    # real time from the imaginary-time result with an imprinted vortex line
    dt = 5e-4
    psi = va.vortex_line(dims, origin, b, 0.1, -0.2, ge.copy())
    s.set_mode("real", dt)
    s.set_state(psi)
    s.init_odd(mu)
    s.step(100)
    re, ro = s.get_state()
    lf, lf0, gdt = va.leapfrog_coefs(b, dt, g)
    oe2, oo2 = orc.leapfrog(psi, va.phase_shift(dims, psi, mu, dt), V * dt, dims, lf, lf0, gdt, 100)
    np.testing.assert_array_equal(re, oe2)
    np.testing.assert_array_equal(ro, oo2)
    s.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_device_gaussian_init_equals_batch_recipe(precision):
    """gpe_init_gaussian sets the even field with the batch's device recipe, normalised on the grid:
    bit-identical to a batch instance of the same parameters, unit norm, and odd = even turned by
    the state's own chemical potential exactly as gpe_init_odd turns it."""
    from vaikkari_amd.batch import Batch, instance_config
    from vaikkari_amd.serve import stable_dt

    announce(f"device gaussian init {precision}")
    dims, h, g, ratio = (21, 17, 13), 0.3, 700.0, 1.7
    b = h / math.sqrt(8)
    dt = stable_dt(13, h, g, ratio)
    s = va.Solver(dims, b, dt, g, None, precision, "real", "trap", ratio)
    s.init_gaussian(1.5, 1e-3, 99)
    se, so = s.get_state()
    B = Batch(precision)
    i = B.add(instance_config(dims, h, g, dt, 10, "real", ratio, sigma=1.5, noise=1e-3, seed=99))
    be, _ = B.get_state(i)
    B.close()
    np.testing.assert_array_equal(se, be)
    assert va.field_norm(dims, b, se) == pytest.approx(1.0, abs=1e-13 if precision == "fp64" else 1e-6)
    mu = s.observables(0)["mu"]
    ref = va.phase_shift(dims, se, mu, dt)
    if precision == "fp64":
        np.testing.assert_array_equal(so, ref)
    else:
        assert rel_l2(so, ref) < 1e-6
    s.close()
