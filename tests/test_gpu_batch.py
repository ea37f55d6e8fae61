"""GPU tests of the batched stepper: each instance matches the oracle (or the
single-instance solver) and is bit-identical alone and in a mixed batch."""
import math

import numpy as np
import pytest

import vaikkari_amd as va
from vaikkari_amd.batch import Batch, instance_config
from vaikkari_amd.serve import stable_dt
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"[gpu-case] {msg}", flush=True)


@pytest.fixture(scope="module")
def orc():
    return Oracle()


def host_case(dims, h, g, dt, ratio, seed, sigma=1.2, mu=8.0):
    b = h / math.sqrt(8)
    origin = va.centered_origin(dims, b)
    even = va.gaussian(dims, origin, b, sigma, 1e-3, seed)
    odd = va.phase_shift(dims, even, mu, dt)
    V = va.trap_potential(dims, origin, b, ratio)
    return b, origin, even, odd, V


def random_case(dims, h, ratio, seed, amp=0.05):
    rng = np.random.default_rng(seed)
    b = h / math.sqrt(8)
    origin = va.centered_origin(dims, b)
    n = va.nodes_of(dims)
    even = amp * rng.uniform(-1, 1, 2 * n)
    odd = even + 0.01 * amp * rng.uniform(-1, 1, 2 * n)
    return b, origin, even, odd, va.trap_potential(dims, origin, b, ratio)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_real_time_matches_oracle(orc, precision):
    announce(f"batch real {precision}")
    B = Batch(precision)
    cases = [((13, 9, 21), 0.45, 400.0, 1e-3, 1.5, 1), ((70, 5, 8), 0.4, 900.0, 8e-4, 0.7, 2),
             ((24, 24, 24), 0.25, 2000.0, 1e-3, 2.0, 3)]
    ids, refs = [], []
    lf = {}
    for dims, h, g, dt, ratio, seed in cases:
        b, origin, even, odd, V = random_case(dims, h, ratio, seed)
        cfg = instance_config(dims, h, g, dt, 37, "real", ratio, init="host", origin=origin)
        i = B.add(cfg)
        B.set_state(i, even, odd)
        ids.append(i)
        c = va.leapfrog_coefs(b, dt, g)
        refs.append(orc.leapfrog(even, odd, V * dt, dims, *c, 37))
    B.run(10, 4)  # 40 >= 37: every instance stops at its own step count
    for i, (oe, oo) in zip(ids, refs):
        assert B.info(i)["steps_done"] == 37
        ge, go = B.get_state(i)
        if precision == "fp64":
            np.testing.assert_array_equal(ge, oe)
            np.testing.assert_array_equal(go, oo)
        else:
            assert np.linalg.norm(ge - oe) / np.linalg.norm(oe) < 2e-6
    B.close()


@pytest.mark.parametrize("zlayers,kernel", [(0, "gather"), (8, "zmarch")])
def test_batch_imag_time_matches_solver_and_oracle(orc, zlayers, kernel):
    """One-layer tiles (the default) equal the single solver's 4-wave gather tile bit for bit, 8-layer
    z-march tiles its z-march kernel; both match the oracle to 1e-12."""
    dims, h, g, ratio = (20, 14, 17), 0.4, 600.0, 1.7
    announce(f"batch imag fp64 zlayers={zlayers}")
    b, origin, even, _, V = host_case(dims, h, g, 1e-3, ratio, 5)
    dtau = 0.5 / (3.0 / b ** 2 + V.max())
    B = Batch("fp64")
    i = B.add(instance_config(dims, h, g, dtau, 45, "imag", ratio, init="host", origin=origin, zlayers=zlayers))
    B.set_state(i, even)
    B.run(16, 3)
    ge, _ = B.get_state(i)
    # the batch's one-layer ground-state tiles are the single solver's gather tile 11 (4 waves,
    # y-bands), the 8-layer ones its default z-march tile
    s = va.Solver(dims, b, dtau, g, origin, "fp64", "imag", "trap", ratio, kernel, 11 if kernel == "gather" else 0)
    s.set_state(even)
    s.step(45)
    se = s.get_state(odd=False)
    np.testing.assert_array_equal(ge, se)
    lf, lf0, gdt = va.leapfrog_coefs(b, dtau, g)
    oe, _ = orc.imag_steps(even, V * dtau, dims, lf, lf0, gdt, b, 45)
    assert np.linalg.norm(ge - oe) / np.linalg.norm(oe) < 1e-12
    obs = B.observables(i)
    ref = orc.observables(oe, V, dims, b, g)
    assert obs["N"] == pytest.approx(ref[0], rel=1e-12)
    assert obs["E"] == pytest.approx(ref[1] + ref[2] + 0.5 * ref[3], rel=1e-12)
    s.close()
    B.close()


def _mixed_configs():
    rng = np.random.default_rng(12)
    out = []
    for q in range(7):
        n = int(rng.choice([8, 12, 16, 24, 33]))
        mode = "imag" if q % 3 == 0 else "real"
        g = float(np.exp(rng.uniform(np.log(100), np.log(5000))))
        ratio = float(rng.uniform(0.5, 4.0))
        dt = stable_dt(n, 0.25, g, ratio, mode=mode)
        out.append(dict(dims=(n, n, n), h=0.25, g=g, dt=dt, steps=int(rng.integers(20, 60)), mode=mode,
                        trap_ratio=ratio, seed=q + 100))
    return out


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_alone_equals_in_mix(precision):
    """Each instance's result is bit-identical when stepped alone and inside a mixed batch of
    real- and imaginary-time instances of other sizes (chunks of uneven length)."""
    announce(f"alone vs mix {precision}")
    cfgs = _mixed_configs()
    mix = Batch(precision)
    if precision == "fp64":
        mix.set_launch_mode("stream")  # the mix without graphs against graph-launched instances alone
        mix.set_group_bytes(0)  # every real-time instance in one group
    else:
        mix.set_group_bytes(1)  # every real-time instance a group of its own
    ids = [mix.add(instance_config(c["dims"], c["h"], c["g"], c["dt"], c["steps"], c["mode"], c["trap_ratio"],
                                   seed=c["seed"], mu0=5.0)) for c in cfgs]
    mix.run(7, 10)
    for c, i in zip(cfgs, ids):
        alone = Batch(precision)
        j = alone.add(instance_config(c["dims"], c["h"], c["g"], c["dt"], c["steps"], c["mode"], c["trap_ratio"],
                                      seed=c["seed"], mu0=5.0))
        alone.run(13, 6)
        a_e, a_o = alone.get_state(j)
        m_e, m_o = mix.get_state(i)
        assert alone.info(j)["steps_done"] == mix.info(i)["steps_done"] == c["steps"]
        np.testing.assert_array_equal(a_e, m_e)
        if a_o is not None:
            np.testing.assert_array_equal(a_o, m_o)
        alone.close()
    mix.close()


def test_group_schedule_changes_nothing():
    """Real-time instances advanced group by group (group sizes changed between chunks, from one
    instance per group to all in one) end bit-identical to the whole batch in one group."""
    announce("batch group schedule")
    cfgs = _mixed_configs()
    runs = []
    for groups in ([0, 0, 0, 0], [1, 3 << 20, 0, 1]):
        B = Batch("fp32")
        ids = [B.add(instance_config(c["dims"], c["h"], c["g"], c["dt"], c["steps"], c["mode"], c["trap_ratio"],
                                     seed=c["seed"], mu0=5.0)) for c in cfgs]
        for gb in groups:
            B.set_group_bytes(gb)
            assert B.group_bytes == gb
            B.run(9, 2)
        runs.append([B.get_state(i) for i in ids])
        assert all(B.info(i)["steps_done"] == c["steps"] for c, i in zip(cfgs, ids))
        B.close()
    for (e0, o0), (e1, o1) in zip(*runs):
        np.testing.assert_array_equal(e1, e0)
        if o0 is not None:
            np.testing.assert_array_equal(o1, o0)


def test_long_chunk_split_changes_nothing():
    """A chunk too long for one captured graph (3 groups x 2 launches x 3000 steps > 16384 launches)
    runs as shorter pieces and ends bit-identical to 100-step chunks."""
    announce("batch long chunk split")
    cfgs = [instance_config((8, 8, 8), 0.3, 300.0, 5e-4, 3000, "real", 1.0, seed=1, mu0=2.0),
            instance_config((12, 10, 9), 0.3, 700.0, 5e-4, 2950, "real", 1.6, seed=2, mu0=3.0),
            instance_config((5, 7, 3), 0.3, 100.0, 5e-4, 3000, "real", 2.0, seed=3, mu0=1.0)]
    runs = []
    for group_bytes, chunk, chunks in ((1, 3000, 1), (0, 100, 30)):
        B = Batch("fp32")
        B.set_group_bytes(group_bytes)
        ids = [B.add(c) for c in cfgs]
        ms = B.run(chunk, chunks)
        assert ms > 0
        assert [B.info(i)["steps_left"] for i in ids] == [0, 0, 0]
        runs.append([B.get_state(i) for i in ids])
        B.close()
    for (e0, o0), (e1, o1) in zip(*runs):
        np.testing.assert_array_equal(e1, e0)
        np.testing.assert_array_equal(o1, o0)


def test_device_gaussian_init_and_budget():
    announce("batch init + budget")
    cfg = instance_config((32, 32, 32), 0.25, 500.0, 1e-3, 10, "real", 1.0, sigma=1.5, noise=1e-3, seed=7)
    B = Batch("fp32", hbm_budget_bytes=0)
    per = B.instance_bytes(cfg)
    B.close()
    B = Batch("fp32", hbm_budget_bytes=2 * per + per // 2)
    i0, i1 = B.add(cfg), B.add(cfg)
    assert B.try_add(cfg) is None  # over budget
    assert B.resident_bytes == 2 * per
    e, o = B.get_state(i0)
    dims, b = (32, 32, 32), 0.25 / math.sqrt(8)
    assert va.field_norm(dims, b, e) == pytest.approx(1.0, abs=1e-6)  # normalised on the grid (FP32 storage)
    e1, _ = B.get_state(i1)
    np.testing.assert_array_equal(e, e1)  # same seed, same state
    B.remove(i1)
    assert B.try_add(cfg) is not None
    B.close()


def test_graph_replays_match_eager_chunks():
    """One captured chunk replayed several times == the same steps in other chunkings, and ==
    the chunks enqueued directly on the stream without a graph."""
    announce("graph replay consistency")
    cfg = instance_config((30, 20, 25), 0.3, 800.0, 6e-4, 60, "real", 1.3, seed=3, mu0=4.0)
    outs = []
    for chunk, chunks, mode in ((10, 6, "graph"), (60, 1, "graph"), (7, 9, "graph"), (1, 60, "graph"),
                                (10, 6, "stream"), (7, 9, "stream")):
        B = Batch("fp64")
        B.set_launch_mode(mode)
        i = B.add(cfg)
        for _ in range(chunks):
            B.run(chunk, 1)
            B.get_state(i)  # host work between replays
        outs.append(B.get_state(i))
        B.close()
    for e, o in outs[1:]:
        np.testing.assert_array_equal(e, outs[0][0])
        np.testing.assert_array_equal(o, outs[0][1])


def test_automatic_start_phase():
    """mu0 = NaN: odd = even exp(i mu dt/2) with mu = (K + P + I)/N of the initial state, exactly;
    a strongly interacting Gaussian then keeps its norm better than with mu0 = 0."""
    announce("automatic start phase")
    dims, h, g, ratio = (24, 24, 24), 0.25, 3000.0, 2.0
    dt = stable_dt(24, h, g, ratio)
    B = Batch("fp64")
    i = B.add(instance_config(dims, h, g, dt, 400, "real", ratio, seed=9))
    e, o = B.get_state(i)
    obs = B.observables(i)
    mu = (obs["K"] + obs["P"] + obs["I"]) / obs["N"]
    assert mu > 30.0  # strongly interacting: the phase matters
    cr, ci = math.cos(0.5 * mu * dt), math.sin(0.5 * mu * dt)
    np.testing.assert_array_equal(o[0::2], cr * e[0::2] - ci * e[1::2])
    np.testing.assert_array_equal(o[1::2], ci * e[0::2] + cr * e[1::2])
    drift = {}
    for name, m0 in (("auto", float("nan")), ("zero", 0.0)):
        C = Batch("fp64")
        j = C.add(instance_config(dims, h, g, dt, 400, "real", ratio, seed=9, mu0=m0))
        n0 = C.observables(j)["N"]
        worst = 0.0
        for _ in range(8):
            C.run(50, 1)
            worst = max(worst, abs(C.observables(j)["N"] / n0 - 1.0))
        drift[name] = worst
        C.close()
    assert drift["auto"] < 0.8 * drift["zero"], drift
    B.close()


def test_device_gaussian_init_unit_norm_on_clipped_box():
    """A 16^3 box at h = 0.25 (edge 4) clips a sigma = 1.5 Gaussian: the continuum amplitude
    alone gives a discrete norm of about 0.83.  The device recipe normalises on the grid: FP64
    norm 1 to rounding, real and imaginary time, and the odd level is the even one turned by
    the start phase, exactly."""
    announce("batch init clipped box")
    dims, h = (16, 16, 16), 0.25
    b = h / math.sqrt(8)
    B = Batch("fp64")
    for mode in ("real", "imag"):
        i = B.add(instance_config(dims, h, 800.0, stable_dt(16, h, 800.0, 1.0, mode=mode), 10, mode, 1.0,
                                  sigma=1.5, noise=1e-3, seed=77, mu0=3.0))
        e, o = B.get_state(i)
        assert va.field_norm(dims, b, e) == pytest.approx(1.0, abs=1e-13)
        if mode == "real":
            dt = stable_dt(16, h, 800.0, 1.0)
            cr, ci = math.cos(0.5 * 3.0 * dt), math.sin(0.5 * 3.0 * dt)
            np.testing.assert_array_equal(o[0::2], cr * e[0::2] - ci * e[1::2])
            np.testing.assert_array_equal(o[1::2], ci * e[0::2] + cr * e[1::2])
    B.close()
