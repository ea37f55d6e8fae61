"""GPU tests of the continuous-batching server: every request completes, a request's result
is bit-identical whether it is served alone or among others arriving while it runs, and
admission respects the HBM budget."""
import numpy as np
import pytest

from vaikkari_amd.serve import Request, Server, poisson_workload, run_serving

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
