"""GPU tests of the continuous-batching server: every request completes, a request's result
is bit-identical whether it is served alone or among others arriving while it runs, equals the
single solver and the oracle on the same initial state, and admission respects the HBM budget
and drops a request that cannot run exactly once."""
import math

import numpy as np
import pytest

import vaikkari_amd as va
from vaikkari_amd._lib import ENOMEM, EUNSTABLE, GpeError, OK
from vaikkari_amd.batch import Batch, instance_config
from vaikkari_amd.serve import Request, Server, poisson_workload, run_serving, stable_dt
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"[gpu-case] {msg}", flush=True)


def small_requests():
    w = poisson_workload(seed=3, rate=200.0, duration=0.05, n_min=8, n_max=24, steps_min=20, steps_max=120)
    assert len(w) >= 5
    return w


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_alone_equals_in_mix(precision):
    announce(f"serve alone vs mix {precision}")
    reqs = small_requests()
    target = reqs[len(reqs) // 2]
    # alone
    srv = Server(precision, 0, chunk_steps=16)
    t = srv.submit(target, arrival=0.0, keep_result=True)
    while srv.stats()["completed"] < 1:
        srv.tick()
    alone = srv.fetch(t, target.n, target.mode)
    srv.close()
    # in a mix, arriving while others run, different chunk length
    srv = Server(precision, 0, chunk_steps=7)
    tickets = {}
    for q, r in enumerate(reqs):
        tickets[q] = srv.submit(r, arrival=0.0, keep_result=(r is target))
        srv.tick()
    while srv.stats()["completed"] < len(reqs):
        srv.tick()
    done = srv.poll()
    assert len(done) == len(reqs)
    mix = srv.fetch(tickets[reqs.index(target)], target.n, target.mode)
    np.testing.assert_array_equal(alone[0], mix[0])
    if alone[1] is not None:
        np.testing.assert_array_equal(alone[1], mix[1])
    for d in done:
        assert d["finished"] >= d["admitted"] >= d["arrival"]
        assert np.isfinite(d["energy"]) and d["norm"] > 0
    srv.close()


def test_budget_admission_and_queue():
    announce("serve budget")
    r = Request(arrival=0.0, n=20, steps=30, g=500.0, ratio=1.0, mode="real", seed=1)
    from vaikkari_amd.batch import Batch, instance_config
    from vaikkari_amd.serve import stable_dt
    b = Batch("fp32")
    per = b.instance_bytes(instance_config((20, 20, 20), 0.25, 500.0, stable_dt(20, 0.25, 500.0, 1.0), 30))
    b.close()
    srv = Server("fp32", hbm_budget_bytes=2 * per, chunk_steps=10)
    for q in range(5):
        srv.submit(Request(0.0, 20, 30, 500.0, 1.0, "real", q))
    srv.tick()
    st = srv.stats()
    assert st["active"] == 2 and st["queued"] == 3 and st["resident_bytes"] <= 2 * per
    while srv.stats()["completed"] < 5:
        srv.tick()
    assert srv.stats()["queued"] == 0
    srv.close()


def test_run_serving_short():
    announce("serve 3 s replay")
    w = poisson_workload(seed=11, rate=20.0, duration=3.0, n_min=16, n_max=48, steps_min=100, steps_max=400)
    res = run_serving(w, duration=3.0, precision="fp32", hbm_budget_bytes=8 * 10 ** 9, chunk_steps=25)
    assert res["submitted"] >= 1 and res["served_site_steps_per_s"] > 0
    assert res["completed"] >= 1 and res["p50_s"] is not None and res["p99_s"] >= res["p50_s"]


@pytest.mark.parametrize("policy", ["shortest_first", "fifo"])
def test_policy_order_and_node_cap(policy):
    """Long request A arrives before short B and C; the cap fits B + C but not A beside them.
    shortest_first admits B and C first and A only after both finish; fifo admits A alone first."""
    announce(f"serve policy {policy}")
    a = Request(0.0, 24, 400, 500.0, 1.0, "real", 1)
    b = Request(0.0, 16, 30, 500.0, 1.0, "real", 2)
    c = Request(0.0, 20, 60, 500.0, 1.0, "imag", 3)
    cap = 12 * 24 ** 3
    srv = Server("fp32", 0, chunk_steps=10, policy=policy, max_active_nodes=cap)
    ta, tb, tc = (srv.submit(r, arrival=0.0) for r in (a, b, c))
    assert srv.oldest_queued() == 0.0
    srv.tick()
    st = srv.stats()
    assert st["active"] == (2 if policy == "shortest_first" else 1)
    done = {}
    while len(done) < 3:
        srv.tick()
        for d in srv.poll():
            done[d["ticket"]] = d
        assert srv.stats()["active"] <= 2
    assert srv.oldest_queued() is None
    if policy == "shortest_first":
        assert done[ta]["admitted"] >= max(done[tb]["finished"], done[tc]["finished"])
    else:
        assert done[ta]["finished"] <= min(done[tb]["admitted"], done[tc]["admitted"])
    with pytest.raises(Exception):
        srv.set_policy("lottery")
    srv.close()


def _initial_state(r: Request, precision: str):
    """The initial even/odd fields the server gives request r: the batch's device recipe on
    the same instance parameters (Server.submit: sigma 1.5, noise 1e-3, the state's own mu)."""
    dt = stable_dt(r.n, r.h, r.g, r.ratio, mode=r.mode)
    B = Batch(precision)
    i = B.add(instance_config((r.n,) * 3, r.h, r.g, dt, r.steps, r.mode, r.ratio, sigma=1.5, noise=1e-3,
                              seed=r.seed))
    e, o = B.get_state(i)
    B.close()
    return dt, e, o


@pytest.mark.parametrize("launch", ["graph", "stream"])
def test_served_results_match_solver_and_oracle(launch):
    """A real-time and an imaginary-time request of the serving form, served FP64 among requests
    of other sizes that arrive while they run: the real-time result is bit-identical to the single
    solver and to the oracle's leapfrog (GPUGrossPitaevskii_BCC.cu:205-215,475-481) from the same
    initial state, the ground state bit-identical to the single solver and within 1e-12 of the
    oracle's imaginary-time steps."""
    announce(f"served vs solver + oracle ({launch})")
    orc = Oracle()
    real = Request(0.0, 22, 260, 1300.0, 2.2, "real", 424242)
    imag = Request(0.0, 19, 150, 700.0, 0.8, "imag", 515151)
    others = poisson_workload(seed=8, rate=400.0, duration=0.04, n_min=10, n_max=30, steps_min=30, steps_max=200)
    srv = Server("fp64", 0, chunk_steps=23, policy="fifo", launch=launch)
    tickets = {}
    pending = [real] + others[: len(others) // 2] + [imag] + others[len(others) // 2:]
    for q, r in enumerate(pending):
        tickets[q] = srv.submit(r, arrival=srv.now(), keep_result=r in (real, imag))
        if q % 2 == 0:
            srv.tick()  # the next requests arrive while these run
    while srv.stats()["completed"] < len(pending):
        srv.tick()
    done = srv.poll()
    assert sorted(d["ticket"] for d in done) == sorted(tickets.values())
    assert all(d["status"] == OK for d in done)
    served_real = srv.fetch(tickets[pending.index(real)], real.n, "real")
    served_imag = srv.fetch(tickets[pending.index(imag)], imag.n, "imag")
    srv.close()

    for r, (ge, go) in ((real, served_real), (imag, served_imag)):
        dims = (r.n,) * 3
        b = r.h / math.sqrt(8)
        origin = va.centered_origin(dims, b)
        V = va.trap_potential(dims, origin, b, r.ratio)
        dt, e0, o0 = _initial_state(r, "fp64")
        assert va.field_norm(dims, b, e0) == pytest.approx(1.0, abs=1e-13)  # unit norm on the grid
        lf, lf0, gdt = va.leapfrog_coefs(b, dt, r.g)
        # tile 11 (4 waves, y-bands): the tile of the server's ground-state launches
        s = va.Solver(dims, b, dt, r.g, origin, "fp64", r.mode, "trap", r.ratio, "gather", 11)
        if r.mode == "real":
            s.set_state(e0, o0)
            s.step(r.steps)
            se, so = s.get_state()
            np.testing.assert_array_equal(ge, se)
            np.testing.assert_array_equal(go, so)
            oe, oo = orc.leapfrog(e0, o0, V * dt, dims, lf, lf0, gdt, r.steps)
            np.testing.assert_array_equal(ge, oe)
            np.testing.assert_array_equal(go, oo)
        else:
            s.set_state(e0)
            s.step(r.steps)
            se = s.get_state(odd=False)
            np.testing.assert_array_equal(ge, se)
            oe, _ = orc.imag_steps(e0, V * dt, dims, lf, lf0, gdt, b, r.steps)
            assert np.linalg.norm(ge - oe) / np.linalg.norm(oe) < 1e-12
        s.close()


def test_unstable_and_failed_requests():
    """An unstable request is refused at submission (GPE_EUNSTABLE).  A request that passes
    submission but cannot be allocated while nothing else is resident is dropped once, with a
    failed completion, and the requests around it are admitted once each and complete."""
    announce("serve failed admission")
    srv = Server("fp32", 0, chunk_steps=10)
    bad = Request(0.0, 20, 30, 500.0, 1.0, "real", 9)
    with pytest.raises(GpeError) as ei:
        srv.submit(bad, arrival=0.0, dt=0.05)
    assert ei.value.status == EUNSTABLE
    a = Request(0.0, 16, 40, 500.0, 1.0, "real", 1)
    huge = Request(0.0, 2048, 10, 500.0, 1.0, "real", 2)  # ~1.7 TB of fields: the allocation fails
    c = Request(0.0, 18, 40, 500.0, 1.0, "imag", 3)
    ta = srv.submit(a, arrival=0.0)
    srv.tick()  # a is resident: huge must wait, not fail
    th = srv.submit(huge, arrival=0.0)
    tc = srv.submit(c, arrival=0.0)
    seen = []
    for _ in range(200):
        srv.tick()
        seen.extend(srv.poll())
        st = srv.stats()
        if st["active"] == 0 and st["queued"] == 0:
            break
    tickets = [d["ticket"] for d in seen]
    assert sorted(tickets) == sorted([ta, th, tc])  # each exactly once
    by = {d["ticket"]: d for d in seen}
    assert by[th]["status"] == ENOMEM and math.isnan(by[th]["norm"]) and by[th]["node_steps"] == 0
    assert by[ta]["status"] == OK and by[tc]["status"] == OK
    st = srv.stats()
    assert st["completed"] == 2 and st["failed"] == 1 and st["admitted"] == 2 and st["resident_bytes"] == 0
    srv.close()


def _req(n, steps, mode="real", seed=1, g=300.0):
    return Request(arrival=0.0, n=n, steps=steps, g=g, ratio=1.0, mode=mode, h=0.25, seed=seed)


def test_cancel_queued_and_running_requests():
    """gpe_server_cancel: a queued request leaves the queue and a resident one leaves the shared
    launches (its memory freed at once), each with one GPE_ECANCELED completion; the request that
    was not cancelled gives the same bits as when served alone; unknown or finished tickets are
    refused."""
    announce("serve cancel queued + running")
    from vaikkari_amd._lib import ECANCELED, EINVAL

    a, b, c = _req(20, 400, seed=1), _req(16, 60, seed=2), _req(12, 40, seed=3)
    srv = Server("fp32", 0, chunk_steps=8, policy="fifo", max_active_nodes=12 * 20 ** 3)
    ta = srv.submit(a, arrival=0.0)
    tb = srv.submit(b, arrival=0.0)
    tc = srv.submit(c, arrival=0.0, keep_result=True)
    srv.tick()  # a is admitted; the cap keeps b and c queued
    st = srv.stats()
    assert st["active"] == 1 and st["queued"] == 2 and st["resident_bytes"] > 0
    srv.cancel(tb)
    srv.cancel(ta)
    st = srv.stats()
    assert st["active"] == 0 and st["queued"] == 1 and st["resident_bytes"] == 0 and st["cancelled"] == 2
    done = {d["ticket"]: d for d in srv.poll()}
    assert set(done) == {ta, tb}
    assert all(d["status"] == ECANCELED and d["node_steps"] == 0 and math.isnan(d["norm"]) for d in done.values())
    while srv.stats()["completed"] < 1:
        srv.tick()
    rest = srv.poll()
    assert [d["ticket"] for d in rest] == [tc] and rest[0]["status"] == OK
    with pytest.raises(GpeError) as ei:
        srv.cancel(tc)  # finished
    assert ei.value.status == EINVAL
    with pytest.raises(GpeError) as ei:
        srv.cancel(10 ** 9)  # never issued
    assert ei.value.status == EINVAL
    mix = srv.fetch(tc, c.n, c.mode)
    srv.close()
    alone_srv = Server("fp32", 0, chunk_steps=8)
    t = alone_srv.submit(c, arrival=0.0, keep_result=True)
    while alone_srv.stats()["completed"] < 1:
        alone_srv.tick()
    alone = alone_srv.fetch(t, c.n, c.mode)
    alone_srv.close()
    np.testing.assert_array_equal(alone[0], mix[0])
    np.testing.assert_array_equal(alone[1], mix[1])


def test_timeout_drops_queued_and_resident_requests():
    """gpe_server_set_timeout: at the next tick every queued or resident request older than the
    limit is dropped with one GPE_ETIMEDOUT completion (a resident one frees its memory); a
    negative or non-finite limit is refused."""
    announce("serve timeout")
    import time

    from vaikkari_amd._lib import EINVAL, ETIMEDOUT

    srv = Server("fp32", 0, chunk_steps=8)
    for bad in (-1.0, float("nan"), float("inf")):
        with pytest.raises(GpeError) as ei:
            srv.set_timeout(bad)
        assert ei.value.status == EINVAL
    now = srv.now()
    t_run = srv.submit(_req(16, 10 ** 6, seed=4), arrival=now)  # would run for a long time
    srv.tick()
    assert srv.stats()["active"] == 1
    t_old = srv.submit(_req(12, 40, seed=5), arrival=now - 100.0)  # arrived long ago
    srv.set_timeout(50.0)
    srv.tick()  # drops the old request before admission; the young one keeps running
    d = srv.poll()
    assert [x["ticket"] for x in d] == [t_old] and d[0]["status"] == ETIMEDOUT
    assert srv.stats()["active"] == 1
    srv.set_timeout(1e-3)
    time.sleep(0.01)
    srv.tick()
    d = srv.poll()
    assert [x["ticket"] for x in d] == [t_run] and d[0]["status"] == ETIMEDOUT and math.isnan(d[0]["energy"])
    st = srv.stats()
    assert st["active"] == 0 and st["resident_bytes"] == 0 and st["timed_out"] == 2 and st["completed"] == 0
    srv.close()
