"""GPU tests at the workloads' own sizes, against the oracle (FP64: bit for bit) and against the
plain FP64 GPU path (FP32 bulk), through the C ABI.

* the reference's own default run: state.dat regenerated by main's recipe, b = 2 pi/(20 sqrt(curv)),
  87 x 87 x 302 cells, one iteration of 691 steps (GPUGrossPitaevskii_BCC.cu:501-555);
* the drop-in form at full size (128^3): imaginary time, vortex imprint, real time;
* the precision form (256^3): FP64 against the oracle, then FP32 beside FP64 for 10,000 steps;
* the batched form (64 instances of 24^3..96^3): FP64 against the oracle, FP32 alone = in the mix.

Each test holds at most a few GB of host memory and runs in well under two minutes on MI355X with
the box's 16 host threads for the OpenMP oracle.
"""
import hashlib
import json
import math
import os

import numpy as np
import pytest

import vaikkari_amd as va
from vaikkari_amd import workloads as wl
from vaikkari_amd.batch import Batch
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "state_files.json")))


def announce(msg):
    print(f"[gpu-case] {msg}", flush=True)


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def orc():
    return Oracle()


@pytest.mark.timeout(600)
def test_reference_default_workload(orc, tmp_path):
    """The reference's main (GPUGrossPitaevskii_BCC.cu:501-555) through the drop-in: the state is
    regenerated by main's recipe and must hash to the shipped state.dat; the extent comes from
    searchMaxR/Z at 1e-5 of the maximum (:505-508), b = 2 pi/(20 sqrt(integrateCurvature)) (:539),
    trap ratio 0.1, period 1.  The product's setup must give 87 x 87 x 302 cells and 691 steps per
    iteration (the value hard-coded at GPUGrossPitaevskii_BCC_multi.cu:473), and one iteration
    of gpe_integrate_in_time must equal the oracle's setup and leapfrog bit for bit."""
    announce("reference default workload: state recipe")
    ref = GOLDEN["state.dat"]
    vs0, _ = va.reference_state_recipe(20, 5000.0, 0.1, True, 10000, 1e-29)
    path = str(tmp_path / "state.dat")
    assert vs0.save(path)
    assert hashlib.sha256(open(path, "rb").read()).hexdigest() == ref["sha256"]
    vs = va.VortexState()
    assert vs.load(path)  # main's LOAD_STATE_FROM_DISK path from here on
    eps = 1e-5 * vs.searchFunctionMax()
    maxr, maxz = vs.searchMaxR(eps), vs.searchMaxZ(eps)
    bs = 2.0 * math.pi / (20.0 * math.sqrt(vs.integrateCurvature()))
    minp, maxp = (-maxr, -maxr, -maxz), (maxr, maxr, maxz)
    announce(f"reference default workload: b = {bs!r}, integrateInTime on the GPU")
    rc, run, ge, go = va.integrateInTime(vs, bs, minp, maxp, 1.0, 1, trap_ratio=0.1, return_fields=True)
    assert rc == 0
    assert run["dims"] == (87, 87, 302) and run["steps_per_iteration"] == 691
    assert va.nodes_of(run["dims"]) == 27430056

    announce("reference default workload: oracle setup + 691 leapfrog steps")
    import ctypes as C
    from vaikkari_amd import _lib
    h = orc.o.or_vs_load(path.encode())
    assert h
    dims = (C.c_uint32 * 3)()
    p0 = (C.c_double * 3)()
    orc.o.or_ref_dims(bs, _lib.d3(minp), _lib.d3(maxp), dims, p0)
    assert tuple(dims) == (87, 87, 302) and tuple(p0) == run["origin"]
    n = va.nodes_of(tuple(dims))
    ev, od, pot, sc = np.zeros(2 * n), np.zeros(2 * n), np.zeros(n), np.zeros(6)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    assert orc.o.or_ref_setup(h, bs, _lib.d3(minp), _lib.d3(maxp), 1.0, 0.1, 54363, dp(ev), dp(od), dp(pot),
                              dp(sc)) == 0
    orc.o.or_vs_free(h)
    assert int(sc[0]) == 691 and sc[1] == run["dt"] and sc[5] == run["maxpot"]
    assert (sc[2], sc[3], sc[4]) == (run["lapfac_dt"], run["lapfac0_dt"], run["g"] * run["dt"])
    oe, oo = orc.leapfrog(ev, od, pot, tuple(dims), sc[2], sc[3], sc[4], 691)
    np.testing.assert_array_equal(ge, oe)
    np.testing.assert_array_equal(go, oo)


@pytest.mark.timeout(600)
def test_drop_in_form_full_size_workflow(orc):
    """The drop-in form at its own 128^3 size (25,165,824 nodes): 200 FP64 imaginary-time steps
    against the oracle (<= 1e-12 relative), the vortex line imprinted at the seeded offset, then
    100 real-time steps bit-identical to the oracle's leapfrog from the same state."""
    w = wl.drop_in_form(128)
    dims, b, origin = w.dims, w.b, w.origin
    announce("drop-in form 128^3: imaginary time")
    psi0 = w.initial()
    dtau = w.dtau(psi0)
    V = va.trap_potential(dims, origin, b, w.ratio)
    s = va.Solver(dims, b, dtau, w.g, origin, "fp64", "imag", "trap", w.ratio)
    s.set_state(psi0)
    s.step(200)
    ground = s.get_state(odd=False)
    lf, lf0, gdt = va.leapfrog_coefs(b, dtau, w.g)
    og, _ = orc.imag_steps(psi0, V * dtau, dims, lf, lf0, gdt, b, 200)
    assert rel_l2(ground, og) <= 1e-12
    del og, psi0
    mu = s.observables(0)["mu"]
    announce("drop-in form 128^3: vortex line + real time")
    psi = w.imprint(ground)
    del ground
    s.set_mode("real", w.dt)
    s.set_state(psi)
    s.init_odd(mu)
    s.step(100)
    ge, go = s.get_state()
    s.close()
    lf, lf0, gdt = va.leapfrog_coefs(b, w.dt, w.g)
    oe, oo = orc.leapfrog(psi, va.phase_shift(dims, psi, mu, w.dt), V * w.dt, dims, lf, lf0, gdt, 100)
    np.testing.assert_array_equal(ge, oe)
    np.testing.assert_array_equal(go, oo)


@pytest.fixture(scope="module")
def precision_state():
    announce("precision form 256^3: Thomas-Fermi + vortex ring")
    return wl.precision_form(256)


@pytest.mark.timeout(600)
def test_precision_form_fp64_against_oracle(orc, precision_state):
    """The precision form at full size (256^3 cells, 201,326,592 nodes): 5 FP64 leapfrog steps of
    the Thomas-Fermi + vortex ring state equal the oracle bit for bit."""
    w = precision_state
    announce("precision form 256^3: fp64 x 5 vs oracle")
    odd = va.phase_shift(w.dims, w.psi, w.mu, w.dt)
    s = va.Solver(w.dims, w.b, w.dt, w.g, w.origin, "fp64", "real", "trap", 1.0)
    s.set_state(w.psi, odd)
    s.step(5)
    ge, go = s.get_state()
    s.close()
    V = va.trap_potential(w.dims, w.origin, w.b, 1.0)
    lf, lf0, gdt = va.leapfrog_coefs(w.b, w.dt, w.g)
    oe, oo = orc.leapfrog(w.psi, odd, V * w.dt, w.dims, lf, lf0, gdt, 5)
    del V, odd
    assert np.array_equal(ge, oe) and np.array_equal(go, oo)


@pytest.mark.timeout(600)
def test_precision_form_fp32_against_fp64_10000_steps(precision_state):
    """FP32 bulk beside the FP64 path (bit-identical to the oracle, test above) for the form's
    10,000 steps at dt = 5e-5, compared on the device every 1000 steps: psi relative L2 <= 1e-5 at
    every checkpoint (the stated FP32-bulk error bound), and the FP32 norm and energy within 1e-7 relative of the
    FP64 ones.  Every checkpoint is printed."""
    w = precision_state
    announce("precision form 256^3: fp32 vs fp64 x 10000")
    s64 = va.Solver(w.dims, w.b, w.dt, w.g, w.origin, "fp64", "real", "trap", 1.0)
    s32 = va.Solver(w.dims, w.b, w.dt, w.g, w.origin, "fp32", "real", "trap", 1.0)
    for s in (s64, s32):
        s.set_state(w.psi)
        s.init_odd(w.mu)
    o64 = s64.observables(0)
    rows = []
    for step in range(1000, 10001, 1000):
        s64.step(1000)
        s32.step(1000)
        d = s32.compare(s64)
        a64, a32 = s64.observables(0), s32.observables(0)
        rows.append((step, d["rel_l2"], a64["N"] / o64["N"] - 1, a32["N"] / a64["N"] - 1, a64["E"] / o64["E"] - 1,
                     a32["E"] / a64["E"] - 1))
        print(f"[precision] step {step}: psi rel-L2 {d['rel_l2']:.3e}, N drift fp64 {rows[-1][2]:.2e}, "
              f"N fp32/fp64-1 {rows[-1][3]:.2e}, E drift fp64 {rows[-1][4]:.2e}, E fp32/fp64-1 {rows[-1][5]:.2e}",
              flush=True)
    s64.close()
    s32.close()
    for step, err, _, dn, _, de in rows:
        assert err <= 1e-5, (step, err)
        assert abs(dn) <= 1e-7 and abs(de) <= 1e-7, (step, dn, de)


@pytest.mark.timeout(600)
def test_batched_form_fp64_against_oracle(orc):
    """The batched form's 64 instances (seed 4, 24^3..96^3) in one FP64 batch, 40 steps: every
    instance of at most 48^3 cells equals the oracle's leapfrog from its own initial state bit for
    bit, and every instance ends with finite fields."""
    forms = wl.batched_form(64, 40, seed=4)
    cfgs = wl.instance_configs(forms)
    announce(f"batched form fp64: {len(cfgs)} instances, sizes {sorted(set(f['dims'][0] for f in forms))}")
    B = Batch("fp64")
    ids = [B.add(c) for c in cfgs]
    starts = {i: B.get_state(i) for i, f in zip(ids, forms) if f["dims"][0] <= 48}
    assert len(starts) >= 20
    B.run(20, 2)
    for i, f in zip(ids, forms):
        assert B.info(i)["steps_done"] == 40
        ge, go = B.get_state(i)
        assert np.all(np.isfinite(ge)) and np.all(np.isfinite(go))
        if i not in starts:
            continue
        dims, b = f["dims"], f["h"] / math.sqrt(8)
        V = va.trap_potential(dims, va.centered_origin(dims, b), b, f["trap_ratio"])
        e0, o0 = starts[i]
        oe, oo = orc.leapfrog(e0, o0, V * f["dt"], dims, *va.leapfrog_coefs(b, f["dt"], f["g"]), 40)
        np.testing.assert_array_equal(ge, oe)
        np.testing.assert_array_equal(go, oo)
    B.close()


@pytest.mark.timeout(600)
def test_batched_form_fp32_alone_equals_mix_2000_steps():
    """The batched form as benchmarked: 64 instances, FP32 bulk, 2000 steps in one batch.  Every
    instance's final fields equal the same instance stepped alone, bit for bit, and no instance's
    norm drifts by more than 2e-3 (the leapfrog's O(dt^2) oscillation; 7.5e-4 measured)."""
    forms = wl.batched_form(64, 2000, seed=4)
    cfgs = wl.instance_configs(forms)
    announce("batched form fp32: 64 instances x 2000 steps, mix")
    mix = Batch("fp32")
    ids = [mix.add(c) for c in cfgs]
    n0 = {i: mix.observables(i)["N"] for i in ids}
    mix.run(100, 20)
    worst = 0.0
    for q, (i, c) in enumerate(zip(ids, cfgs)):
        assert mix.info(i)["steps_done"] == 2000
        worst = max(worst, abs(mix.observables(i)["N"] / n0[i] - 1.0))
        alone = Batch("fp32")
        j = alone.add(c)
        alone.run(250, 8)
        ae, ao = alone.get_state(j)
        me, mo = mix.get_state(i)
        np.testing.assert_array_equal(ae, me)
        np.testing.assert_array_equal(ao, mo)
        alone.close()
        if q % 16 == 15:
            announce(f"batched form fp32: {q + 1} of 64 instances alone = in mix")
    mix.close()
    print(f"[batched] worst relative norm drift over 2000 steps: {worst:.2e}", flush=True)
    assert worst <= 2e-3
