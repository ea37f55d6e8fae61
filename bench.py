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
on the solver's stream around exactly --steps steps.  Before them run the
--warmup steps and, when those took under 50 ms of device time on the FP64
solver, more untimed steps up to 50 ms, as many for FP32 (warmup_extra_steps):
the GPU idles during the host-side imprint and upload, and its first
milliseconds back run below the steady rate.

Beside the line of record (all untimed or timed separately): fp32_check, FP32
against FP64 over a fixed --check-steps from the phase-2 state, psi relative L2
(reduced on the device) with norm and energy drift at four checkpoints; dual256,
both precisions at the precision form's 256^3 grid; cpu_form_gpu, both precisions
on the CPU form's 32^3 grid (1000 steps); the batched form (64
instances), a 20 s sample of the serving form, the 14-point primal stepper, and
cpu_baseline, the CPU oracle on the host's cores (warm-up, then the median of
timed runs) at the 128^3 grid and at the 32^3 CPU-form grid.
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
    ap.add_argument("--cpu-warmup", type=int, default=10, help="untimed oracle steps before the timed runs")
    ap.add_argument("--cpu-runs", type=int, default=5, help="timed oracle runs (the median is reported)")
    ap.add_argument("--cpu-steps", type=int, default=100, help="steps per timed oracle run at 128^3")
    ap.add_argument("--check-steps", type=int, default=1000,
                    help="length of the untimed FP32-vs-FP64 check from the benchmark state")
    ap.add_argument("--dual256-steps", type=int, default=100,
                    help="timed steps of the 256^3 precision-form grid, FP64 and FP32 (0 = skip)")
    ap.add_argument("--batch-instances", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=2000)
    ap.add_argument("--serve-seconds", type=float, default=20.0,
                    help="length of the Poisson serving sample (0 = skip)")
    ap.add_argument("--suite", default="main",
                    choices=["main", "precision", "serving", "primal", "batched", "cpu_form"],
                    help="main: the JSON line of record; precision: 256^3 FP32-vs-FP64 error study; "
                         "serving: the Poisson serving stream alone (use --serve-seconds 600); "
                         "primal: the 14-point primal stepper's tile sweep alone (with the 1 GiB copy "
                         "probe that calibrates rocprof traffic counters); batched: the 64-instance batched "
                         "workload alone (with the same probe); cpu_form: the CPU form's 32^3 grid on the "
                         "GPU alone")
    ap.add_argument("--batch-tile", default="",
                    help="batched suite: fixed real-time tile instead of the autotuner's, waves followed by "
                         "c (cached read-modify-write) and/or s (slab order per instance) or g (slab order "
                         "per group), e.g. 2, 1cs or 2g")
    ap.add_argument("--batch-sizes", default="",
                    help="batched suite: cube edges to draw the instances from (comma list; default "
                         "24,32,48,64,96)")
    ap.add_argument("--batch-group-mib", type=int, default=None,
                    help="batched suite: fixed group size of the real-time schedule in MiB (0 = one group)")
    ap.add_argument("--batch-launch", default="graph", choices=["graph", "stream"],
                    help="batched workload: replay a captured hipGraph per chunk, or enqueue the kernels "
                         "directly (for per-dispatch counter runs)")
    ap.add_argument("--serve-launch", default="graph", choices=["graph", "stream"],
                    help="serving: replay a captured hipGraph per chunk (default), or enqueue the kernels "
                         "directly (for kernel traces of the serving loop)")
    ap.add_argument("--serve-tile", default="",
                    help="serving suite: real-time tile of the shared launches (as --batch-tile; default the "
                         "batch's own)")
    ap.add_argument("--serve-policy", default="shortest_first", choices=["shortest_first", "fifo"],
                    help="admission policy of --suite serving (shortest_first is capped at 32 M resident nodes)")
    ap.add_argument("--primal-cells", default="128,256",
                    help="cube sizes of the 14-point primal stepper sweep (comma list; empty = skip)")
    ap.add_argument("--primal-steps", type=int, default=200)
    ap.add_argument("--primal-tiles", default="",
                    help="primal suite: tile numbers to sweep (comma list; default every tile)")
    ap.add_argument("--precision-cells", type=int, default=256)
    ap.add_argument("--precision-steps", type=int, default=10000)
    ap.add_argument("--precision-every", type=int, default=1000)
    return ap.parse_args()


def measured_traffic(nodes, precision, info):
    """HBM bytes per launch of the timed kernel on this workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs (profiles/hbm_traffic_128cubed.md), when the
    workload and kernel variant are the ones profiled; else None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_128cubed.json")
    # the default variants at 128^3 cells and their kernels
    tile, key = {"fp64": (24, "k_quad<double,0,1,0,6,true>"), "fp32": (11, "k_gather<float,1,1,0,4,true>")}[precision]
    if nodes != 25165824 or info["kernel"] != "gather" or info["tile"] != tile or not os.path.exists(path):
        return None
    try:
        k = json.load(open(path))["kernels"][key]
    except (OSError, ValueError, KeyError):
        return None
    return k["read_bytes"] + k["write_bytes"]


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cpu():
    """CPU model and core counts of this host from lscpu (physical cores = sockets x cores per socket)."""
    import subprocess

    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    def num(k):
        try:
            return int(info.get(k, ""))
        except ValueError:
            return None
    sockets, per, tpc = num("Socket(s)"), num("Core(s) per socket"), num("Thread(s) per core")
    return {"model": info.get("Model name", "unknown"), "sockets": sockets, "cores_per_socket": per,
            "threads_per_core": tpc, "physical_cores": sockets * per if sockets and per else None,
            "logical_cpus": num("CPU(s)"), "affinity_cpus": len(os.sched_getaffinity(0))}


def cpu_baseline(warmup, runs, steps128):
    """The oracle (oracle/liboracle.so: C + OpenMP, FP64, the reference's AoS layout of 12 complex
    values per cell, -O2 -ffp-contract=off) on this host's cores, timed as the CPU-baseline plan says:
    `warmup` untimed steps, then the median of `runs` runs of >= 100 steps, each run timed inside the
    oracle with the monotonic clock (the fields are padded once, outside the timing).  Two grids:
    the benchmark's own (128^3 cells, h 0.1, g 2000, ratio 2, dt 2e-4: `value`) and the CPU form's
    (32^3 cells, h 0.25, g 500, isotropic trap, dt 1e-3, 1000 steps per run)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle  # noqa: E402
    import vaikkari_amd as va

    orc = Oracle()
    grids = []
    for label, n, h, g, dt, ratio, nsteps in (("drop_in_128", 128, 0.1, 2000.0, 2e-4, 2.0, steps128),
                                              ("cpu_form_32", 32, 0.25, 500.0, 1e-3, 1.0, 1000)):
        dims = (n, n, n)
        b = h / math.sqrt(8)
        origin = va.centered_origin(dims, b)
        even = va.gaussian(dims, origin, b, 1.5, 1e-3, 54363)
        odd = va.phase_shift(dims, even, 10.0, dt)
        V = va.trap_potential(dims, origin, b, ratio) * dt
        lf, lf0, gdt = va.leapfrog_coefs(b, dt, g)
        secs, threads = orc.leapfrog_timed(even, odd, V, dims, lf, lf0, gdt, warmup, runs, nsteps)
        nodes = va.nodes_of(dims)
        med = float(np.median(secs))
        rate = nodes * nsteps / med
        grids.append({"grid": label, "cells": list(dims), "nodes": nodes, "steps_per_run": nsteps,
                      "runs_s": [round(float(x), 4) for x in secs], "median_s": med, "site_steps_per_s": rate,
                      # the oracle streams prev, next (read + write) and V per node-update: 16 + 32 + 8 bytes
                      "host_GBs": rate * 2 * 56 / 1e9, "threads": threads})
        log(f"cpu oracle {label}: {rate / 1e6:.1f} M site-steps/s (median of {runs} x {nsteps} steps, "
            f"{threads} threads)")
    main = grids[0]
    cpu = host_cpu()
    return {"value": main["site_steps_per_s"], "unit": "site-steps/s", "cores": main["threads"], "kind": "port",
            "cpu": cpu, "grids": grids,
            "sample": f"oracle FP64 leapfrog on the benchmark's 128^3-cell grid ({main['nodes']} nodes): "
                      f"{warmup} warm-up steps, median of {runs} runs of {main['steps_per_run']} steps "
                      f"({main['median_s']:.2f} s each) on {main['threads']} OpenMP threads; "
                      f"the 32^3 CPU-form grid timed the same way in 'grids'"}


def batched_workload(n_inst, steps, seed=4, log=log, launch="graph", tile=None, group_mib=None, sizes=None):
    """64 independent condensates: n^3 cells with n uniform in {24,32,48,64,96}, cube edge 0.25,
    g log-uniform in [100, 5000], trap anisotropy uniform in [0.5, 4], dt = 1e-3*32/n capped by
    the reference's stability rule, FP32 bulk; Gaussian + noise initial states (device recipe).
    tile / group_mib fix the real-time tile (waves, cached, slabs) and the group size instead of
    autotuning them."""
    from vaikkari_amd.autotune import tune_batch_groups, tune_batch_tile
    from vaikkari_amd.batch import Batch

    from vaikkari_amd import workloads as wl

    cfgs = wl.instance_configs(wl.batched_form(n_inst, steps, seed=seed, sizes=sizes or wl.BATCH_SIZES))

    def make(waves, cfg_list=cfgs):
        b = Batch("fp32")
        b.set_tile(waves)
        b.set_launch_mode(launch)
        for c in cfg_list:
            b.add(c)
        return b

    if tile is None:
        tuned = tune_batch_tile(lambda t: make(t, cfgs[: max(8, n_inst // 4)]), chunk_steps=10, chunks=2)
    else:
        tuned = {"waves": tile[0], "cached": tile[1], "slabs": tile[2], "tile": tuple(tile), "ms": "fixed"}
    log(f"batched: tile {tuned['waves']} waves, {'cached' if tuned['cached'] else 'streamed'} "
        f"read-modify-write, {'group-slab' if tuned['slabs'] == 2 else 'slab' if tuned['slabs'] else 'y-band'} "
        f"order ({tuned['ms']})")
    if group_mib is None:
        T = make(tuned["tile"])  # the whole workload, only to tune the group size of the schedule
        grouping = tune_batch_groups(T, chunk_steps=10, chunks=2)
        T.close()
    else:
        grouping = {"group_bytes": int(group_mib) << 20, "ms": "fixed"}
    log(f"batched: groups of at most {grouping['group_bytes'] >> 20} MiB of fields ({grouping['ms']})")
    B = make(tuned["tile"])
    B.set_group_bytes(grouping["group_bytes"])
    ids = list(B.dims.keys())
    n0 = {i: B.observables(i)["N"] for i in ids}
    chunk = 100
    B.run(chunk, 1)  # graph capture + first chunk, untimed
    rest = steps - chunk
    nchunks, tail = divmod(rest, chunk)
    node_steps = sum(B.info(i)["nodes"] for i in ids) * rest
    ms_total = B.run(chunk, nchunks) if nchunks else 0.0
    if tail:
        ms_total += B.run(tail, 1)
    assert all(B.info(i)["steps_left"] == 0 for i in ids)
    drift = max(abs(B.observables(i)["N"] / n0[i] - 1.0) for i in ids)
    nodes = sum(B.info(i)["nodes"] for i in ids)
    B.close()
    rate = node_steps / (ms_total * 1e-3)
    gbs = rate * 48.0 / 1e9
    return {"instances": n_inst, "steps": steps, "nodes_total": nodes, "site_steps_per_s": rate,
            "timed_ms": ms_total, "timed_steps": rest, "algorithmic_GBs": gbs,
            "frac_of_peak": gbs / HBM_PEAK_GBS, "tile_waves": tuned["waves"], "tile_cached_rmw": tuned["cached"],
            "tile_slabs": tuned["slabs"], "group_MiB": grouping["group_bytes"] >> 20,
            "max_rel_norm_drift": drift,
            "dtype": "fp32 bulk", "sizes": sorted(set(int(c.nx) for c in cfgs)), "h": 0.25}


# Resident nodes under shortest-first admission.  On this stream (20 s) 32 M finishes the most work:
# 228-230 requests, p50 2.5 s, p99 7.0-7.7 s, 118-119 G site-steps/s served; 24 M gives p50 1.9 s at
# 116 G, 48 M p50 3.4 s / p99 10.0 s at 117 G, 64 M p50 4.2 s / p99 11.4 s at 116 G.
SERVE_CAP_NODES = 32 * 10 ** 6


def serving_workload(seconds, policy="shortest_first", cap=SERVE_CAP_NODES, log=log, launch="graph", tile=None):
    """Poisson stream (rate 20/s, seed 5; n log-uniform in [16, 160], steps in [500, 20000], 30 % ground
    states), FP32 bulk, 200 GB HBM budget, 50-step chunks, replayed against the wall clock."""
    from vaikkari_amd import workloads as wl
    from vaikkari_amd.serve import run_serving

    w = wl.serving_form(seconds)
    res = run_serving(w, duration=seconds, precision="fp32", hbm_budget_bytes=200 * 10 ** 9, chunk_steps=50,
                      log=log, policy=policy, max_active_nodes=cap, launch=launch, tile=tile)
    res["rate_per_s"] = 20.0
    res["hbm_budget_GB"] = 200
    res["chunk_steps"] = 50
    return res


def serving_suite(args):
    """The full serving configuration: 600 s of the Poisson stream (about 12,000 requests) under one
    admission policy; prints one JSON line with latency percentiles, offered vs served load and the
    queue-depth time series."""
    cap = SERVE_CAP_NODES if args.serve_policy == "shortest_first" else 0
    tile = None
    if args.serve_tile:
        w, flags = int(args.serve_tile.rstrip("csg")), args.serve_tile.lstrip("0123456789")
        tile = (w, "c" in flags, 2 if "g" in flags else "s" in flags)
    res = serving_workload(args.serve_seconds, policy=args.serve_policy, cap=cap, launch=args.serve_launch,
                           tile=tile)
    print(json.dumps({"suite": "serving", **res}), flush=True)


def warm_up(s, steps, min_ms=50.0):
    """Untimed steps on solver s: `steps`, then more until at least min_ms of device time has
    run (a solver's first milliseconds after creation run below its steady rate)."""
    ms, _ = s.time_steps(steps)
    if 0 < ms < min_ms:
        s.time_steps(int(math.ceil((min_ms - ms) / (ms / steps))))
    s.synchronize()


def primal_workload(sizes, steps, warmup=100, log=log, tiles_only=None):
    """The 14-point primal BCC stepper (2 sites per cubic cell): FP64 and FP32
    leapfrog steps of a Gaussian condensate (sigma 2, kick e^{0.3ix}) in the trap
    0.5 r^2 on a box of edge 16, g = 100, dt = 0.4 / (7.5/a^2 + max V), over every
    tile shape.  One launch updates every site once: FP64 reads prev (16 B) and
    reads + writes next (32 B), 48 B per site; FP32 24 B.  Each timed run follows at least
    `warmup` untimed steps and 50 ms of device time on the same solver (warm_up): at 128^3 the
    first 200 steps after creation run 3-5 % below the steady rate.  tiles_only limits the
    sweep to those tile numbers (the default shape is always timed)."""
    import vaikkari_amd as va

    out = []
    for n in sizes:
        dims = (n, n, n)
        a = 16.0 / n
        g = 100.0
        origin = va.primal_centered_origin(dims, a)
        vmax = 0.5 * 3 * (8.0 ** 2)
        dt = 0.4 / (7.5 / a ** 2 + vmax)
        xyz = va.primal_positions(dims, origin, a)
        psi = np.exp(-(xyz ** 2).sum(axis=1) / 8.0 + 0.3j * xyz[:, 0])
        del xyz
        psi /= math.sqrt(va.site_measure(a) * float(np.sum(np.abs(psi) ** 2)))
        psi = psi.view(np.float64)
        sites = va.sites_of(dims)
        entry = {"cells": n, "sites": sites, "a": a, "dt": dt, "g": g}
        mu = None
        for prec, bps in (("fp64", 48), ("fp32", 24)):
            tiles = {}
            for tile in (tiles_only or range(1, va.primal_tiles() + 1)):
                with va.PrimalSolver(dims, a, dt, g, origin, 1.0, prec, "real", tile) as s:
                    s.set_state(psi)
                    if mu is None:
                        o = s.observables(0)
                        mu = o["mu"]
                    s.init_odd(mu)
                    warm_up(s, warmup)
                    ms, launches = s.time_steps(steps)
                    avg_s = ms * 1e-3 / launches
                    gbs = sites * bps / avg_s / 1e9
                    tiles[tile] = {"site_steps_per_s": sites * steps / (ms * 1e-3), "avg_launch_us": avg_s * 1e6,
                                   "achieved_GBs": gbs}
            with va.PrimalSolver(dims, a, dt, g, origin, 1.0, prec, "real", 0) as s:
                s.set_state(psi)
                s.init_odd(mu)
                warm_up(s, warmup)
                ms, launches = s.time_steps(steps)
                default_rate = sites * steps / (ms * 1e-3)
            best = max(tiles, key=lambda t: tiles[t]["site_steps_per_s"])
            b = tiles[best]
            entry[prec] = {"best_tile": best, "site_steps_per_s": b["site_steps_per_s"],
                           "avg_launch_us": b["avg_launch_us"], "bytes_per_site_update": bps,
                           "achieved_GBs": b["achieved_GBs"], "frac_of_peak": b["achieved_GBs"] / HBM_PEAK_GBS,
                           "default_tile_site_steps_per_s": default_rate,
                           "tiles": {str(t): round(v["site_steps_per_s"] / 1e9, 3) for t, v in tiles.items()}}
            log(f"primal14 {n}^3 {prec}: default shape {default_rate / 1e9:.2f} G site-steps/s; "
                f"best tile {best}: {b['site_steps_per_s'] / 1e9:.2f} G site-steps/s, "
                f"{b['avg_launch_us']:.1f} us/launch, {b['achieved_GBs']:.0f} GB/s "
                f"({b['achieved_GBs'] / HBM_PEAK_GBS * 100:.1f}% of peak); per tile (G/s) {entry[prec]['tiles']}")
        out.append(entry)
    return out


def monitor_timing(s, reps=20):
    """Cost of energy monitoring on the FP64 solver: gpe_step_observe (N, K, P, I
    reduced inside the last step's odd half-update) against a plain step, and the
    separate observables pass (k_observables + k_finalize) it replaces."""
    def ms(fn):
        s.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        s.synchronize()
        return (time.perf_counter() - t0) * 1e3 / reps
    s.step(5)
    # wall time of an observed step (one synchronisation: it returns the values)
    # against a plain step followed by a synchronisation
    plain_ms, _ = s.time_steps(reps)
    plain_ms /= reps
    observed = ms(lambda: s.step_observe(1))
    separate = ms(lambda: s.observables(0))
    stepsync = ms(lambda: (s.step(1), s.synchronize()))
    return {"plain_step_ms": plain_ms, "observed_step_ms": observed, "plain_step_with_sync_ms": stepsync,
            "observed_step_extra_us": (observed - stepsync) * 1e3, "separate_pass_us": separate * 1e3}


def output_timing(s, dims, reps=5):
    """Per-iteration output of the reference (picture + volume map, GPUGrossPitaevskii_BCC.cu:424-468) done
    on the device: median wall time of one gpe_volume_map / gpe_slice_picture call (kernel + copy of the
    result), beside one download of the full even field, which is what the reference copies each time."""
    def median_ms(fn):
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append((time.perf_counter() - t0) * 1e3)
        return float(np.median(ts))

    k = dims[2] // 2 + 1
    return {"volume_map_call_ms": median_ms(lambda: s.volume_map(1.0)),
            "slice_picture_call_ms": median_ms(lambda: s.slice_picture(k)),
            "field_download_ms": median_ms(lambda: s.get_state(odd=False)),
            "volume_bytes": 2 * (dims[0] + 2) * (dims[1] + 2) * (dims[2] + 2)}


def precision_study(args):
    """The precision form (256^3 cells, 201,326,592 nodes, h = 0.06, g = 5000, isotropic trap,
    Thomas-Fermi profile with the seed-3 vortex ring, dt = 5e-5): FP32 bulk beside the FP64 path
    (bit-identical to the CPU oracle) every `every` steps: psi relative L2 (reduced on the device),
    norm drift and energy drift."""
    import vaikkari_amd as va
    from vaikkari_amd import workloads as wl

    w = wl.precision_form(args.precision_cells)
    log(f"precision: {w.dims} cells, ring R={w.ring_radius:.3f} axis={np.round(w.ring_axis, 3).tolist()}, "
        f"mu_TF={w.mu:.3f}")
    s64 = va.Solver(w.dims, w.b, w.dt, w.g, w.origin, "fp64", "real", "trap", 1.0)
    s32 = va.Solver(w.dims, w.b, w.dt, w.g, w.origin, "fp32", "real", "trap", 1.0)
    for s in (s64, s32):
        s.set_state(w.psi)
        s.init_odd(w.mu)
    rows, ms64, ms32 = precision_checkpoints(s64, s32, args.precision_steps, args.precision_every)
    nodes = va.nodes_of(w.dims)
    done = rows[-1]["step"]
    out = {"suite": "precision", "cells": list(w.dims), "nodes": nodes, "h": w.h, "g": w.g, "dt": w.dt,
           "steps": args.precision_steps, "ring_radius": w.ring_radius, "ring_axis": list(w.ring_axis),
           "fp64_site_steps_per_s": nodes * done / (ms64 * 1e-3),
           "fp32_site_steps_per_s": nodes * done / (ms32 * 1e-3),
           "checkpoints": rows, "max_psi_rel_l2": max(r["psi_rel_l2"] for r in rows)}
    print(json.dumps(out), flush=True)


def precision_checkpoints(s64, s32, steps, every):
    """Steps an FP64 and an FP32 solver side by side from the same state; every `every` steps the
    psi relative L2 error of FP32 against FP64 (gpe_compare_fields: on the device, FP64 sums),
    and the norm and energy drift of both.  Returns (rows, FP64 device ms, FP32 device ms)."""
    o64, o32 = s64.observables(0), s32.observables(0)
    rows = []
    done = 0
    ms64 = ms32 = 0.0
    while done < steps:
        k = min(every, steps - done)
        ms64 += s64.time_steps(k)[0]
        ms32 += s32.time_steps(k)[0]
        done += k
        err = s32.compare(s64)["rel_l2"]
        a64, a32 = s64.observables(0), s32.observables(0)
        rows.append({"step": done, "psi_rel_l2": err,
                     "norm_drift_fp64": a64["N"] / o64["N"] - 1, "norm_drift_fp32": a32["N"] / o32["N"] - 1,
                     "energy_drift_fp64": a64["E"] / o64["E"] - 1, "energy_drift_fp32": a32["E"] / o32["E"] - 1,
                     "norm_fp32_minus_fp64": a32["N"] - a64["N"], "energy_fp32_minus_fp64": a32["E"] - a64["E"]})
        log(f"precision step {done}: rel-L2 {err:.3e}, N drift fp64 {rows[-1]['norm_drift_fp64']:.2e} "
            f"fp32 {rows[-1]['norm_drift_fp32']:.2e}")
    return rows, ms64, ms32


def dual256_leg(steps, warmup=20):
    """The dual-node stepper at the precision form's grid (256^3 cells, 201,326,592 nodes, h = 0.06,
    g = 5000, isotropic trap, dt = 5e-5), FP64 and FP32 bulk, from the device Gaussian (sigma 1.5,
    1e-3 noise, seed 3, normalised on the grid; no host state): `steps` timed steps each with HIP
    events on the solver's stream, the hot kernel's average launch time against the roofline."""
    import vaikkari_amd as va

    n = 256
    dims, h, g, dt = (n, n, n), 0.06, 5000.0, 5e-5
    b = h / math.sqrt(8)
    nodes = va.nodes_of(dims)
    out = {"cells": list(dims), "nodes": nodes, "h": h, "g": g, "dt": dt, "steps": steps}
    for prec in ("fp64", "fp32"):
        with va.Solver(dims, b, dt, g, None, prec, "real", "trap", 1.0) as s:
            s.init_gaussian(1.5, 1e-3, 3)
            warm_up(s, warmup)
            ms, launches = s.time_steps(steps)
            info = s.info()
        upd_bytes = info["bytes_per_node_step"] / 2.0 * nodes
        avg = ms * 1e-3 / launches
        gbs = upd_bytes / avg / 1e9
        out[prec] = {"site_steps_per_s": nodes * steps / (ms * 1e-3), "ms_per_step": ms / steps,
                     "avg_launch_us": avg * 1e6, "achieved_GBs": gbs, "frac_of_peak": gbs / HBM_PEAK_GBS,
                     "kernel": info["kernel"], "tile": info["tile"]}
        log(f"256^3 {prec}: {out[prec]['site_steps_per_s'] / 1e9:.2f} G site-steps/s, {gbs:.0f} GB/s "
            f"({gbs / HBM_PEAK_GBS * 100:.1f}% of peak)")
    out["fp32_vs_fp64"] = out["fp32"]["site_steps_per_s"] / out["fp64"]["site_steps_per_s"]
    return out


def cpu_form_leg(steps=1000, warmup=20):
    """The CPU form's workload on the GPU (32^3 cells, 393,216 nodes, h 0.25, g 500, isotropic
    trap, dt 1e-3; unit-norm Gaussian sigma 1.5 with 1e-3 seeded noise, odd level at mu 10): `steps`
    real-time steps, FP64 and FP32 bulk, default tiles, timed with HIP events -- the GPU figure
    beside cpu_baseline's 32^3 grid."""
    import vaikkari_amd as va

    dims, h, g, dt = (32, 32, 32), 0.25, 500.0, 1e-3
    b = h / math.sqrt(8)
    origin = va.centered_origin(dims, b)
    nodes = va.nodes_of(dims)
    psi = va.gaussian(dims, origin, b, 1.5, 1e-3, 54363)
    out = {"cells": list(dims), "nodes": nodes, "h": h, "g": g, "dt": dt, "steps": steps}
    for prec in ("fp64", "fp32"):
        with va.Solver(dims, b, dt, g, origin, prec, "real", "trap", 1.0) as s:
            s.set_state(psi)
            s.init_odd(10.0)
            warm_up(s, warmup)
            ms, launches = s.time_steps(steps)
            info = s.info()
        out[prec] = {"site_steps_per_s": nodes * steps / (ms * 1e-3), "ms": ms,
                     "avg_launch_us": ms * 1e3 / launches, "kernel": info["kernel"], "tile": info["tile"]}
        log(f"32^3 {prec}: {out[prec]['site_steps_per_s'] / 1e9:.2f} G site-steps/s, 1000 steps in {ms:.1f} ms "
            f"(tile {info['tile']})")
    return out


def main():
    args = parse()
    if args.suite in ("precision", "serving", "primal", "batched", "cpu_form"):
        if args.gpus != 1:
            sys.exit(2)
        if args.suite == "cpu_form":
            print(json.dumps({"suite": "cpu_form", "cpu_form_gpu": cpu_form_leg()}), flush=True)
        elif args.suite == "precision":
            precision_study(args)
        elif args.suite == "primal":
            import vaikkari_amd as va

            copy_gbs = va.copy_bandwidth(1 << 30, 3)
            only = [int(x) for x in args.primal_tiles.split(",")] if args.primal_tiles else None
            res = primal_workload([int(x) for x in args.primal_cells.split(",")], args.primal_steps, tiles_only=only)
            print(json.dumps({"suite": "primal", "hbm_copy_GBs": copy_gbs, "primal14": res}), flush=True)
        elif args.suite == "batched":
            import vaikkari_amd as va

            copy_gbs = va.copy_bandwidth(1 << 30, 3)
            tile = None
            if args.batch_tile:
                w, flags = int(args.batch_tile.rstrip("csg")), args.batch_tile.lstrip("0123456789")
                tile = (w, "c" in flags, 2 if "g" in flags else "s" in flags)
            sizes = [int(x) for x in args.batch_sizes.split(",")] if args.batch_sizes else None
            res = batched_workload(args.batch_instances, args.batch_steps, launch=args.batch_launch, tile=tile,
                                   group_mib=args.batch_group_mib, sizes=sizes)
            print(json.dumps({"suite": "batched", "hbm_copy_GBs": copy_gbs, "batched": res}), flush=True)
        else:
            serving_suite(args)
        return
    if args.gpus != 1:
        print(f"bench.py: this stepper runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the stepper shards by z-slabs across GPUs by design, not built here)", file=sys.stderr)
        sys.exit(2)
    import vaikkari_amd as va
    from vaikkari_amd import workloads as wl

    N = args.cells
    w = wl.drop_in_form(N)
    dims, h, g, ratio, dt = w.dims, w.h, w.g, w.ratio, w.dt
    b, origin = w.b, w.origin
    nodes = va.nodes_of(dims)
    log(f"grid {dims} cells, {nodes} nodes")
    copy_gbs = va.copy_bandwidth(1 << 30, 10)
    log(f"plain streaming copy: {copy_gbs:.0f} GB/s ({copy_gbs / HBM_PEAK_GBS * 100:.1f}% of the 8 TB/s peak)")
    rmw_gbs = va.stream_bandwidth("rmw", 1 << 30, 10)
    read_gbs = va.stream_bandwidth("read", 1 << 30, 10)
    log(f"stencil-pattern stream (read + non-temporal RMW, no stencil): {rmw_gbs:.0f} GB/s; read-only: {read_gbs:.0f} GB/s")

    # ---- phase 1: imaginary time, FP64 -----------------------------------
    psi0 = w.initial()
    dtau = w.dtau(psi0)
    s64 = va.Solver(dims, b, dtau, g, origin, "fp64", "imag", "trap", ratio, args.kernel)
    s64.set_state(psi0)
    ms_imag, _ = s64.time_steps(args.imag_steps)
    obs = s64.observables(0)
    log(f"phase 1: {args.imag_steps} imaginary steps in {ms_imag:.1f} ms, E = {obs['E']:.6f}, mu = {obs['mu']:.6f}")
    ground = s64.get_state(odd=False)

    # ---- phase 2 input: vortex line at a seeded offset ----------------------
    psi = w.imprint(ground)
    del ground
    mu = obs["mu"]

    results = {}
    extra = 0
    for prec in ("fp64", "fp32"):
        if prec == "fp64":
            s = s64
            s.set_mode("real", dt)
        else:
            s = va.Solver(dims, b, dt, g, origin, "fp32", "real", "trap", ratio, args.kernel)
        s.set_state(psi)
        s.init_odd(mu)
        if args.warmup:
            # the GPU idled during the host-side imprint and upload; after the W warm-up steps,
            # more untimed steps until 50 ms of device time have run on the FP64 solver, so that
            # the timed steps see the steady rate however short W is (reported as
            # warmup_extra_steps).  The FP32 run takes as many, so that both runs end at the same
            # step and the psi comparison below compares the same time.
            ms_w, _ = s.time_steps(args.warmup)
            if prec == "fp64":
                extra = int(math.ceil((50.0 - ms_w) / (ms_w / args.warmup))) if 0 < ms_w < 50.0 else 0
            if extra:
                s.step(extra)
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
            "warmup_extra_steps": extra,
            "ms_per_step": ms / args.steps, "avg_launch_us": avg_launch_s * 1e6,
            "bytes_per_node_step": info["bytes_per_node_step"],
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(nodes, prec, info),
                         "algorithmic_bytes_per_launch": upd_bytes},
        }
        log(f"{prec}: {rate / 1e9:.2f} G site-steps/s, {ms / args.steps * 1e3:.1f} us/step, "
            f"{achieved:.0f} GB/s algorithmic ({achieved / HBM_PEAK_GBS * 100:.1f}% of peak), kernel {info['kernel']}")
        if prec == "fp32":
            s32 = s
    # psi error of the timed FP32 run against the timed FP64 run (same state, same steps)
    err = s32.compare(s64)["rel_l2"]
    s32.close()
    output = output_timing(s64, dims)
    monitor = monitor_timing(s64)
    log(f"monitoring: a step with fused observables costs {monitor['observed_step_extra_us']:.1f} us more than a "
        f"plain step; a separate observables pass {monitor['separate_pass_us']:.1f} us")
    log(f"output: volume map {output['volume_map_call_ms']:.2f} ms, slice {output['slice_picture_call_ms']:.2f} ms "
        f"per call vs {output['field_download_ms']:.1f} ms to download the even field")
    s64.close()

    # FP32 against FP64 over a fixed length, whatever --steps is: both from the phase-2 state,
    # compared on the device every 250 steps (untimed)
    c64 = va.Solver(dims, b, dt, g, origin, "fp64", "real", "trap", ratio, args.kernel)
    c32 = va.Solver(dims, b, dt, g, origin, "fp32", "real", "trap", ratio, args.kernel)
    for s in (c64, c32):
        s.set_state(psi)
        s.init_odd(mu)
    del psi
    check_rows, _, _ = precision_checkpoints(c64, c32, args.check_steps, max(1, args.check_steps // 4))
    c64.close()
    c32.close()
    check = {"steps": args.check_steps, "checkpoints": check_rows,
             "psi_rel_l2": check_rows[-1]["psi_rel_l2"],
             "max_abs_norm_drift_fp32": max(abs(r["norm_drift_fp32"]) for r in check_rows),
             "max_abs_energy_drift_fp32": max(abs(r["energy_drift_fp32"]) for r in check_rows),
             "max_abs_norm_drift_fp64": max(abs(r["norm_drift_fp64"]) for r in check_rows),
             "max_abs_energy_drift_fp64": max(abs(r["energy_drift_fp64"]) for r in check_rows)}

    dual256 = dual256_leg(args.dual256_steps) if args.dual256_steps > 0 else None
    cpu_form_gpu = cpu_form_leg()

    batched = None
    if args.batch_instances > 0:
        batched = batched_workload(args.batch_instances, args.batch_steps)
        log(f"batched: {batched['site_steps_per_s'] / 1e9:.1f} G site-steps/s over {batched['instances']} instances "
            f"({batched['frac_of_peak'] * 100:.1f}% of peak), norm drift {batched['max_rel_norm_drift']:.1e}")
        # the smallest size of the batched form alone: 256 instances of 24^3 cells (1 GiB of fields),
        # tile and group size autotuned as for the mixed form
        small = batched_workload(256, 1000, sizes=(24,))
        batched["uniform_24"] = {k: small[k] for k in ("instances", "steps", "nodes_total", "site_steps_per_s",
                                                        "algorithmic_GBs", "frac_of_peak", "tile_waves",
                                                        "tile_cached_rmw", "tile_slabs", "group_MiB",
                                                        "max_rel_norm_drift")}
        log(f"batched, 256 x 24^3: {small['site_steps_per_s'] / 1e9:.1f} G site-steps/s "
            f"({small['frac_of_peak'] * 100:.1f}% of peak)")
    serving = None
    if args.serve_seconds > 0:
        serving = serving_workload(args.serve_seconds)
        fifo = serving_workload(args.serve_seconds, policy="fifo", cap=0)
        serving["fifo_uncapped"] = {k: fifo[k] for k in ("completed", "p50_s", "p99_s", "p50_all_s", "p99_all_s",
                                                          "censored", "served_site_steps_per_s",
                                                          "active_at_end", "queued_at_end")}
        for name, r in (("shortest-first, capped", serving), ("fifo, uncapped", fifo)):
            log(f"serving ({name}): {r['completed']} done, p50 {r['p50_s']} s, p99 {r['p99_s']} s "
                f"(all {r['submitted']} submitted, unfinished at their age: p50 {r['p50_all_s']} s, "
                f"p99 {r['p99_all_s']} s), served {r['served_site_steps_per_s'] / 1e9:.1f} G site-steps/s")

    primal = None
    if args.primal_cells.strip():
        primal = primal_workload([int(x) for x in args.primal_cells.split(",")], args.primal_steps)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.cpu_warmup, args.cpu_runs, args.cpu_steps)

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
            "psi_rel_l2_vs_fp64": err,
            "psi_rel_l2_note": f"after the {args.warmup + extra + args.steps} steps (warm-up included) of the "
                               "timed runs; fixed-length check in fp32_check",
            "roofline": r32["roofline"], "kernel": r32["kernel"],
        },
        "fp32_check": check,
        "dual256": dual256,
        "cpu_form_gpu": cpu_form_gpu,
        "phase1_imag": {"steps": args.imag_steps, "ms": ms_imag, "site_steps_per_s": nodes * args.imag_steps / (ms_imag * 1e-3),
                        "E": obs["E"], "mu": obs["mu"], "N": obs["N"]},
        "bytes_per_node_step": {"fp64": r64["bytes_per_node_step"], "fp32": r32["bytes_per_node_step"]},
        "hbm_copy_GBs": copy_gbs,
        "hbm_rmw_stream_GBs": rmw_gbs,
        "hbm_read_stream_GBs": read_gbs,
        "batched": batched,
        "serving": serving,
        "primal14": primal,
        "output": output,
        "monitor": monitor,
        "avg_launch_us": {"fp64": r64["avg_launch_us"], "fp32": r32["avg_launch_us"]},
        "warmup_extra_steps": {"fp64": r64["warmup_extra_steps"], "fp32": r32["warmup_extra_steps"]},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
