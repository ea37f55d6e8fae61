"""Continuous-batching solve server (include/vaikkari_gpe_serve.h) and the
Poisson request workload it is benchmarked on.

Workload (serving form): requests arrive as a seeded Poisson stream; each
asks for an n^3-cell solve, n log-uniform in [16, 160], steps uniform in
[500, 20000], g log-uniform in [100, 5000], trap anisotropy uniform in
[0.5, 4], 30 % imaginary-time ground states and 70 % real-time evolutions,
cube edge h = 0.25, Gaussian initial state (sigma 1.5, 1e-3 noise).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import _lib
from ._lib import check, dptr, lib
from .solver import MODES, PRECISIONS, nodes_of


def stable_dt(n: int, h: float, g: float, ratio: float, sigma: float = 1.5, dt_nominal: Optional[float] = None,
              mode: str = "real") -> float:
    """Time step for an n^3-cell request: the nominal 1e-3*(32/n) (real time), capped by the
    reference's own rule dt <= 1/(max(V + g|psi0|^2) + 3/b^2) (GPUGrossPitaevskii_BCC.cu:298).
    Imaginary time uses half the cap.  max |psi0|^2 is the peak of the Gaussian initial state
    normalised on the box, (pi sigma^2)^(-3/2) / erf(n h / (2 sigma))^3: a box that clips the
    Gaussian raises the peak by the norm it cuts off."""
    b = h / math.sqrt(8.0)
    half = 0.5 * n * h
    vmax = 0.5 * (2 * half * half + ratio * ratio * half * half)
    rho = (math.pi * sigma * sigma) ** -1.5 / math.erf(half / sigma) ** 3
    cap = 1.0 / (vmax + g * rho + 3.0 / (b * b))
    if mode == "imag":
        return 0.5 * cap
    nominal = 1e-3 * 32.0 / n if dt_nominal is None else dt_nominal
    return min(nominal, cap)


@dataclass
class Request:
    arrival: float
    n: int
    steps: int
    g: float
    ratio: float
    mode: str
    seed: int
    h: float = 0.25

    def node_steps(self) -> float:
        return 12.0 * self.n ** 3 * self.steps


def poisson_workload(seed: int, rate: float, duration: float, n_min: int = 16, n_max: int = 160,
                     steps_min: int = 500, steps_max: int = 20000, imag_frac: float = 0.3) -> List[Request]:
    rng = np.random.default_rng(seed)
    t, out, q = 0.0, [], 0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        steps = int(rng.integers(steps_min, steps_max + 1))
        g = float(math.exp(rng.uniform(math.log(100.0), math.log(5000.0))))
        ratio = float(rng.uniform(0.5, 4.0))
        mode = "imag" if rng.uniform() < imag_frac else "real"
        out.append(Request(t, n, steps, g, ratio, mode, seed * 1000003 + q))
        q += 1
    return out


POLICIES = {"fifo": 0, "shortest_first": 1}


LAUNCH_MODES = {"graph": 0, "stream": 1}


class Server:
    def __init__(self, precision: str = "fp32", hbm_budget_bytes: int = 200 * 10 ** 9, chunk_steps: int = 50,
                 policy: str = "fifo", max_active_nodes: int = 0, launch: str = "graph", tile=None):
        self._h = C.c_void_p()
        check(lib.gpe_server_create(PRECISIONS[precision], int(hbm_budget_bytes), int(chunk_steps),
                                    C.byref(self._h)), "gpe_server_create")
        self.set_policy(policy, max_active_nodes)
        self.set_launch_mode(launch)
        if tile is not None:
            self.set_tile(tile)

    def set_tile(self, waves, cached: bool = False, slabs=False) -> None:
        """Real-time tile of the shared launches, as Batch.set_tile (before the first admission)."""
        from .batch import tile_code

        check(lib.gpe_server_set_tile(self._h, tile_code(waves, cached, slabs)), "gpe_server_set_tile")

    def set_policy(self, policy: str, max_active_nodes: int = 0) -> None:
        check(lib.gpe_server_set_policy(self._h, POLICIES[policy], int(max_active_nodes)), "gpe_server_set_policy")

    def set_launch_mode(self, mode: str) -> None:
        """'graph' (captured chunk, the default) or 'stream' (kernels enqueued one by one)."""
        check(lib.gpe_server_set_launch_mode(self._h, LAUNCH_MODES[mode]), "gpe_server_set_launch_mode")

    def cancel(self, ticket: int) -> None:
        """Withdraws a queued or resident request (a GPE_ECANCELED completion follows)."""
        check(lib.gpe_server_cancel(self._h, int(ticket)), "gpe_server_cancel")

    def set_timeout(self, seconds: float) -> None:
        """Drops requests older than `seconds` since arrival at the next tick (0: no limit)."""
        check(lib.gpe_server_set_timeout(self._h, float(seconds)), "gpe_server_set_timeout")

    def oldest_queued(self) -> Optional[float]:
        t = lib.gpe_server_oldest_queued(self._h)
        return None if t < 0 else t

    def close(self):
        if self._h:
            lib.gpe_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def now(self) -> float:
        return lib.gpe_server_now(self._h)

    def submit(self, r: Request, arrival: Optional[float] = None, keep_result: bool = False,
               dt: Optional[float] = None) -> int:
        """Queues r with the time step stable_dt gives it (or `dt`); returns its ticket."""
        q = _lib.GpeRequest()
        q.n = r.n
        q.mode = MODES[r.mode]
        q.steps = r.steps
        q.h = r.h
        q.dt = stable_dt(r.n, r.h, r.g, r.ratio, mode=r.mode) if dt is None else float(dt)
        q.g = r.g
        q.trap_ratio = r.ratio
        q.sigma, q.noise, q.mu0 = 1.5, 1e-3, float("nan")  # leapfrog started with the state's own mu
        q.seed = r.seed
        q.keep_result = 1 if keep_result else 0
        t = C.c_uint64()
        check(lib.gpe_server_submit(self._h, C.byref(q), float(r.arrival if arrival is None else arrival),
                                    C.byref(t)), "gpe_server_submit")
        return t.value

    def tick(self) -> int:
        n = C.c_uint32()
        check(lib.gpe_server_tick(self._h, C.byref(n)), "gpe_server_tick")
        return n.value

    def poll(self, max_items: int = 4096):
        buf = (_lib.GpeCompletion * max_items)()
        n = C.c_uint32()
        check(lib.gpe_server_poll(self._h, buf, max_items, C.byref(n)), "gpe_server_poll")
        return [{"ticket": c.ticket, "arrival": c.arrival, "admitted": c.admitted, "finished": c.finished,
                 "node_steps": c.node_steps, "norm": c.norm, "energy": c.energy, "status": c.status}
                for c in buf[:n.value]]

    def fetch(self, ticket: int, n: int, mode: str):
        nodes = nodes_of((n, n, n))
        e = np.empty(2 * nodes)
        o = np.empty(2 * nodes) if mode == "real" else None
        check(lib.gpe_server_fetch(self._h, int(ticket), dptr(e), dptr(o)), "gpe_server_fetch")
        return e, o

    def stats(self) -> dict:
        s = _lib.GpeServerStats()
        check(lib.gpe_server_get_stats(self._h, C.byref(s)), "gpe_server_get_stats")
        return {k: getattr(s, k) for k, _ in _lib.GpeServerStats._fields_}


SIZE_CLASSES = ((16, 32), (32, 64), (64, 128), (128, 161))


def _pct(a, q):
    return float(np.percentile(a, q)) if len(a) else None


def latency_report(workload: List[Request], arrivals: dict, done: List[dict], end: float, duration: float) -> dict:
    """Latency figures of a serving run.

    arrivals: ticket -> (request, arrival on the server clock) of every submitted request;
    done: completions; end: server clock at the end.  Time in system is finished - arrival for a
    completed request and end - arrival (a lower bound: the request is still queued or running)
    for the others, which are right-censored.  Percentiles over completed requests alone favour
    the short solves that shortest-first admission finishes first; the 'all' figures count every
    submitted request and are lower bounds on what those callers see.  Per size class (cells per
    axis) the offered load (node-steps/s of the requests arriving in the window) sits beside the
    served load (node-steps/s of the completed ones)."""
    fin = {d["ticket"]: d for d in done if d.get("status", 0) == 0}
    failed = sum(1 for d in done if d.get("status", 0) != 0)
    dropped = {d["ticket"] for d in done if d.get("status", 0) != 0}
    lat_done, lat_all = [], []
    for t, (r, arr) in arrivals.items():
        if t in fin:
            lat_done.append(fin[t]["finished"] - arr)
            lat_all.append(fin[t]["finished"] - arr)
        elif t not in dropped:
            lat_all.append(end - arr)
    out = {
        "p50_s": _pct(lat_done, 50), "p99_s": _pct(lat_done, 99),
        "mean_s": float(np.mean(lat_done)) if lat_done else None,
        "p50_all_s": _pct(lat_all, 50), "p99_all_s": _pct(lat_all, 99),
        "censored": len(lat_all) - len(lat_done), "failed": failed,
        "latency_note": "p50_s/p99_s: completed requests only; p50_all_s/p99_all_s: every submitted request, "
                        "unfinished ones at their age at the end (lower bounds)",
    }
    classes = []
    for lo, hi in SIZE_CLASSES:
        ts = [t for t, (r, _) in arrivals.items() if lo <= r.n < hi and t not in dropped]
        ld = [fin[t]["finished"] - arrivals[t][1] for t in ts if t in fin]
        la = [fin[t]["finished"] - arrivals[t][1] if t in fin else end - arrivals[t][1] for t in ts]
        offered = sum(r.node_steps() for r in workload if r.arrival < duration and lo <= r.n < hi) / duration
        served = sum(fin[t]["node_steps"] for t in ts if t in fin) / duration
        classes.append({"n": [lo, hi - 1], "submitted": len(ts), "completed": len(ld),
                        "p50_s": _pct(ld, 50), "p99_s": _pct(ld, 99), "p50_all_s": _pct(la, 50),
                        "p99_all_s": _pct(la, 99), "offered_site_steps_per_s": offered,
                        "completed_site_steps_per_s": served})
    out["size_classes"] = classes
    return out


def run_serving(workload: List[Request], duration: float, precision: str = "fp32",
                hbm_budget_bytes: int = 200 * 10 ** 9, chunk_steps: int = 50, log=None, policy: str = "fifo",
                max_active_nodes: int = 0, sample_every: float = 10.0, launch: str = "graph", tile=None) -> dict:
    """Replays the workload against the wall clock for `duration` seconds: requests are
    submitted when their arrival time has passed, the server ticks continuously.  Every
    `sample_every` seconds the queue depth, resident solves and bytes, completions and the age
    of the oldest queued request are recorded (and logged)."""
    srv = Server(precision, hbm_budget_bytes, chunk_steps, policy, max_active_nodes, launch, tile)
    t_start = srv.now()
    pending = sorted(workload, key=lambda r: r.arrival)
    k = 0
    done = []
    arrivals = {}
    samples = []
    last_sample = t_start
    while True:
        now = srv.now() - t_start
        if now >= duration:
            break
        while k < len(pending) and pending[k].arrival <= now:
            arr = t_start + pending[k].arrival
            arrivals[srv.submit(pending[k], arrival=arr)] = (pending[k], arr)
            k += 1
        srv.tick()
        done.extend(srv.poll())
        if srv.now() - last_sample > sample_every:
            st = srv.stats()
            oldest = srv.oldest_queued()
            samples.append({"t": round(now, 2), "queued": int(st["queued"]), "active": int(st["active"]),
                            "completed": int(st["completed"]), "resident_GB": st["resident_bytes"] / 1e9,
                            "oldest_queued_age_s": None if oldest is None else srv.now() - oldest})
            if log:
                log(f"serve t={now:.0f}s completed={st['completed']} active={st['active']} queued={st['queued']} "
                    f"resident={st['resident_bytes'] / 1e9:.1f} GB")
            last_sample = srv.now()
    end = srv.now()
    wall = end - t_start
    st = srv.stats()
    ok = [d for d in done if d.get("status", 0) == 0]
    offered = sum(r.node_steps() for r in workload if r.arrival < duration) / duration
    res = {
        "duration_s": wall,
        "submitted": int(st["submitted"]), "completed": int(st["completed"]), "queued_at_end": int(st["queued"]),
        "active_at_end": int(st["active"]),
        "served_site_steps_per_s": st["node_steps_done"] / wall if wall > 0 else 0.0,
        "device_site_steps_per_s": st["node_steps_done"] / (st["device_ms"] * 1e-3) if st["device_ms"] else 0.0,
        "offered_site_steps_per_s": offered,
        "completed_norm_range": [float(min(d["norm"] for d in ok)), float(max(d["norm"] for d in ok))]
        if ok else None,
        "ticks": int(st["ticks"]),
        "policy": policy, "max_active_nodes": int(max_active_nodes), "launch": launch,
        "completed_work_site_steps": float(sum(d["node_steps"] for d in ok)),
        "samples": samples,
    }
    res.update(latency_report(workload, arrivals, done, end, duration))
    srv.close()
    return res
