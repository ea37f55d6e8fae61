"""Throughput benchmark of the BCC Gross-Pitaevskii stepper on one MI355X.

Workload (single GPU): BCC 128^3 cubic cells = 25,165,824 dual nodes
(12 per cell), cube edge h = 0.1 (block_scale b = h/sqrt(8)), g = 2000,
anisotropic trap V = 0.5(x^2 + y^2 + 4 z^2).
  phase 1: Gaussian (sigma 1.5, 1e-3 noise, seed 1) relaxed by imaginary-time
           Euler steps with FP64 renormalisation every step (untimed setup,
           its rate is reported);
  phase 2: a straight vortex line along z imprinted at a seeded offset in
           [-0.5, 0.5]^2 (seed 2), odd level = even*exp(i mu dt/2), then
           real-time leapfrog steps at dt = 2e-4 -- the timed region.
value = FP64 node-steps/s of phase 2 (one node-step = one node advanced by
one full leapfrog step = two node-updates).  The FP32-bulk stepper runs the
same phase 2 from the same state; its rate, its ratio to FP64 and its psi
relative L2 error against the FP64 path (bit-identical to the CPU oracle)
are reported beside it.

Inputs are resident in HBM before the timed region; timing uses HIP events
on the solver's stream around exactly --steps steps.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--imag-steps", type=int, default=2000)
    ap.add_argument("--cells", type=int, default=128)
    ap.add_argument("--kernel", default="auto")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(dims, b, g, dt, ratio, seconds):
    """The CPU oracle (oracle/liboracle.so, C + OpenMP, FP64, reference AoS layout) on the same
    grid and coefficients: rate = nodes*(n2-n1)/(t2-t1) from two runs, so padding copies cancel."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle  # noqa: E402
    import vaikkari_amd as va

    orc = Oracle()
    origin = va.centered_origin(dims, b)
    even = va.gaussian(dims, origin, b, 1.5, 1e-3, 1)
    odd = va.phase_shift(dims, even, 10.0, dt)
    V = va.trap_potential(dims, origin, b, ratio) * dt
    lf, lf0, gdt = va.leapfrog_coefs(b, dt, g)
    n = va.nodes_of(dims)
    t0 = time.perf_counter()
    orc.leapfrog(even, odd, V, dims, lf, lf0, gdt, 1)
    t1 = time.perf_counter()
    per_step = max(t1 - t0, 1e-3)
    n2 = max(2, int(seconds / per_step))
    t2 = time.perf_counter()
    orc.leapfrog(even, odd, V, dims, lf, lf0, gdt, n2)
    t3 = time.perf_counter()
    rate = n * (n2 - 1) / ((t3 - t2) - (t1 - t0))
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": rate, "unit": "site-steps/s", "cores": threads, "kind": "port",
            "sample": f"oracle FP64 leapfrog on the same {dims[0]}^3-cell grid ({n} nodes), "
                      f"{n2} steps timed ({t3 - t2:.1f} s) minus a 1-step run"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the stepper shards by z-slabs across GPUs by design, not built here)", file=sys.stderr)
        sys.exit(2)
    import vaikkari_amd as va

    N = args.cells
    dims = (N, N, N)
    h, g, ratio, dt = 0.1, 2000.0, 2.0, 2e-4
    b = h / math.sqrt(8)
    origin = va.centered_origin(dims, b)
    nodes = va.nodes_of(dims)
    log(f"grid {dims} cells, {nodes} nodes")
    copy_gbs = va.copy_bandwidth(1 << 30, 10)
    log(f"plain streaming copy: {copy_gbs:.0f} GB/s ({copy_gbs / HBM_PEAK_GBS * 100:.1f}% of the 8 TB/s peak)")

    # ---- phase 1: imaginary time, FP64 -----------------------------------
    psi0 = va.gaussian(dims, origin, b, 1.5, 1e-3, 1)
    V = va.trap_potential(dims, origin, b, ratio)
    dtau = 1.0 / (6.0 / b ** 2 + V.max() + g * float(np.max(psi0[0::2] ** 2 + psi0[1::2] ** 2)))
    del V
    s64 = va.Solver(dims, b, dtau, g, origin, "fp64", "imag", "trap", ratio, args.kernel)
    s64.set_state(psi0)
    ms_imag, _ = s64.time_steps(args.imag_steps)
    obs = s64.observables(0)
    log(f"phase 1: {args.imag_steps} imaginary steps in {ms_imag:.1f} ms, E = {obs['E']:.6f}, mu = {obs['mu']:.6f}")
    ground = s64.get_state(odd=False)

    # ---- phase 2 input: vortex line at a seeded offset ----------------------
    rng = np.random.default_rng(2)
    x0, y0 = rng.uniform(-0.5, 0.5, 2)
    psi = va.vortex_line(dims, origin, b, x0, y0, ground.copy())
    del ground
    mu = obs["mu"]

    results = {}
    for prec in ("fp64", "fp32"):
        if prec == "fp64":
            s = s64
            s.set_mode("real", dt)
        else:
            s = va.Solver(dims, b, dt, g, origin, "fp32", "real", "trap", ratio, args.kernel)
        s.set_state(psi)
        s.init_odd(mu)
        if args.warmup:
            s.step(args.warmup)
        s.synchronize()
        t0 = time.perf_counter()
        ms, launches = s.time_steps(args.steps)
        wall = time.perf_counter() - t0
        info = s.info()
        rate = nodes * args.steps / (ms * 1e-3)
        upd_bytes = info["bytes_per_node_step"] / 2.0 * nodes  # algorithmic bytes per launch
        avg_launch_s = ms * 1e-3 / launches
        achieved = upd_bytes / avg_launch_s / 1e9
        results[prec] = {
            "rate": rate, "ms": ms, "wall_s": wall, "launches": launches, "kernel": info["kernel"],
            "ms_per_step": ms / args.steps, "avg_launch_us": avg_launch_s * 1e6,
            "bytes_per_node_step": info["bytes_per_node_step"],
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "traffic": None},
            "state": s.get_state(odd=False),
        }
        log(f"{prec}: {rate / 1e9:.2f} G site-steps/s, {ms / args.steps * 1e3:.1f} us/step, "
            f"{achieved:.0f} GB/s algorithmic ({achieved / HBM_PEAK_GBS * 100:.1f}% of peak), kernel {info['kernel']}")
        if prec == "fp32":
            s.close()
    err = float(np.linalg.norm(results["fp32"]["state"] - results["fp64"]["state"])
                / np.linalg.norm(results["fp64"]["state"]))
    s64.close()

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(dims, b, g, dt, ratio, args.cpu_seconds)
        log(f"cpu oracle: {cpu['value'] / 1e6:.1f} M site-steps/s on {cpu['cores']} threads")

    r64, r32 = results["fp64"], results["fp32"]
    line = {
        "metric": "site-steps/s + % HBM peak, FP32-bulk vs plain FP64; p50/p99 solve latency",
        "value": r64["rate"],
        "unit": "site-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r64["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"plain FP64 GPE leapfrog, BCC {N}^3 cells (12 dual nodes/cell), h=0.1, g=2000, "
                        "trap 0.5(x^2+y^2+4z^2), dt=2e-4, vortex line after imaginary-time ground state",
            "cells": list(dims), "nodes": nodes, "site": "dual node (12 per cubic cell)",
            "imag_steps": args.imag_steps, "dtau": dtau, "kernel": r64["kernel"],
        },
        "roofline": r64["roofline"],
        "cpu_baseline": cpu,
        "fp32_bulk": {
            "value": r32["rate"], "ms_per_step": r32["ms_per_step"], "ratio_vs_fp64": r32["rate"] / r64["rate"],
            "psi_rel_l2_vs_fp64": err, "roofline": r32["roofline"], "kernel": r32["kernel"],
        },
        "phase1_imag": {"steps": args.imag_steps, "ms": ms_imag, "site_steps_per_s": nodes * args.imag_steps / (ms_imag * 1e-3),
                        "E": obs["E"], "mu": obs["mu"], "N": obs["N"]},
        "bytes_per_node_step": {"fp64": r64["bytes_per_node_step"], "fp32": r32["bytes_per_node_step"]},
        "hbm_copy_GBs": copy_gbs,
        "avg_launch_us": {"fp64": r64["avg_launch_us"], "fp32": r32["avg_launch_us"]},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
