"""CPU tests of the batch/serving host logic: workload generator, time-step
rule, stability guard and admission arithmetic (no device compute)."""
import math

import numpy as np
import pytest

import vaikkari_amd as va
from vaikkari_amd import _lib
from vaikkari_amd.batch import instance_config
from vaikkari_amd.serve import poisson_workload, stable_dt


def test_poisson_workload_statistics():
    w = poisson_workload(seed=5, rate=20.0, duration=600.0)
    assert abs(len(w) / 600.0 - 20.0) < 1.0
    assert all(a.arrival < b.arrival for a, b in zip(w, w[1:]))
    ns = np.array([r.n for r in w])
    assert ns.min() >= 16 and ns.max() <= 160
    # log-uniform in n: median near sqrt(16*160)
    assert abs(np.median(ns) - math.sqrt(16 * 160)) < 6
    frac_imag = np.mean([r.mode == "imag" for r in w])
    assert abs(frac_imag - 0.3) < 0.03
    steps = np.array([r.steps for r in w])
    assert steps.min() >= 500 and steps.max() <= 20000
    # seeded: identical replay
    w2 = poisson_workload(seed=5, rate=20.0, duration=600.0)
    assert [(r.arrival, r.n, r.seed) for r in w] == [(r.arrival, r.n, r.seed) for r in w2]


@pytest.mark.parametrize("n", [16, 24, 32, 64, 96, 160])
@pytest.mark.parametrize("g", [100.0, 5000.0])
@pytest.mark.parametrize("ratio", [0.5, 4.0])
def test_stable_dt_respects_linear_bound(n, g, ratio):
    h = 0.25
    b = h / math.sqrt(8)
    dt = stable_dt(n, h, g, ratio)
    assert dt <= 1e-3 * 32 / n
    dims = (n, n, n)
    V = va.trap_potential((2, 2, 2), va.centered_origin((2, 2, 2), b), b, ratio)  # placeholder shape
    half = 0.5 * n * h
    vmax = 0.5 * (2 * half ** 2 + ratio ** 2 * half ** 2)
    assert dt * (6.0 / b ** 2 + vmax) < 2.0


def test_solver_rejects_unstable_step_before_device_work():
    b = 0.1 / math.sqrt(8)
    with pytest.raises(_lib.GpeError) as ei:
        va.Solver((16, 16, 16), b, 1e-3, 100.0)  # dt * 6/b^2 = 4.8 > 2
    assert ei.value.status == _lib.EUNSTABLE


def test_instance_bytes_accounting():
    from vaikkari_amd.batch import Batch  # noqa: F401  (import only: creating a Batch needs a GPU)
    c = instance_config((32, 32, 32), 0.25, 500.0, 1e-3, 10, "real")
    assert c.block_scale == pytest.approx(0.25 / math.sqrt(8))
    # host-side layout: 4 FP32 fields of (nz+2)*(ny+2)*12*xp complex values
    xp = ((16 + 32 + 1 + 15) // 16) * 16
    per_field = 34 * 34 * 12 * xp * 8
    assert per_field > 12 * 32 ** 3 * 8


def test_latency_report_counts_unfinished_requests():
    """Completed-only percentiles against percentiles over every submitted request, with the
    unfinished ones at their age at the end (right-censored), a dropped request in neither,
    and per-size-class offered against served load."""
    from vaikkari_amd.serve import Request, latency_report

    reqs = [Request(float(q), n, 100, 500.0, 1.0, "real", q) for q, n in enumerate((16, 20, 40, 150, 18))]
    arrivals = {q + 1: (r, r.arrival) for q, r in enumerate(reqs)}
    done = [{"ticket": 1, "finished": 2.0, "node_steps": reqs[0].node_steps(), "status": 0},
            {"ticket": 3, "finished": 9.0, "node_steps": reqs[2].node_steps(), "status": 0},
            {"ticket": 5, "finished": 4.0, "node_steps": 0.0, "status": _lib.ENOMEM}]
    rep = latency_report(reqs, arrivals, done, end=20.0, duration=10.0)
    # completed: ticket 1 took 2 s, ticket 3 took 7 s
    assert rep["p50_s"] == pytest.approx(4.5) and rep["failed"] == 1 and rep["censored"] == 2
    # all: 2, 7 and the unfinished tickets 2 (age 19) and 4 (age 17)
    assert rep["p50_all_s"] == pytest.approx(np.percentile([2, 7, 19, 17], 50))
    assert rep["p99_all_s"] > rep["p99_s"]
    cls = {tuple(c["n"]): c for c in rep["size_classes"]}
    assert cls[(16, 31)]["submitted"] == 2 and cls[(16, 31)]["completed"] == 1  # ticket 5 was dropped
    assert cls[(32, 63)]["completed_site_steps_per_s"] == pytest.approx(reqs[2].node_steps() / 10.0)
    assert cls[(128, 160)]["p50_all_s"] == pytest.approx(17.0) and cls[(128, 160)]["p50_s"] is None
    assert cls[(16, 31)]["offered_site_steps_per_s"] == pytest.approx(
        (reqs[0].node_steps() + reqs[1].node_steps() + reqs[4].node_steps()) / 10.0)


@pytest.mark.parametrize("n", [6, 8, 12, 16, 24])
@pytest.mark.parametrize("mode", ["real", "imag"])
def test_stable_dt_holds_for_the_normalised_initial_state(n, mode):
    """Batched and served Gaussians are normalised on the grid, which raises the peak density of a
    clipped one (a 8^3 box of edge 2 keeps about 28 % of a sigma = 1.5 Gaussian): stable_dt must
    keep dt * (6/b^2 + max V + g max|psi0|^2) < 2 with that peak, for the largest g of the forms."""
    h, g, ratio, sigma = 0.25, 5000.0, 4.0, 1.5
    b = h / math.sqrt(8)
    dims = (n, n, n)
    origin = va.centered_origin(dims, b)
    psi = va.gaussian(dims, origin, b, sigma, 0.0, 1)  # normalised on the grid, no noise
    assert va.field_norm(dims, b, psi) == pytest.approx(1.0, rel=1e-12)
    peak = float(np.max(psi[0::2] ** 2 + psi[1::2] ** 2))
    vmax = float(va.trap_potential(dims, origin, b, ratio).max())
    dt = stable_dt(n, h, g, ratio, sigma=sigma, mode=mode)
    assert dt * (6.0 / b ** 2 + vmax + g * peak) < 2.0
