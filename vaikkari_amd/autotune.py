"""Tile-shape autotuner.

Times every kernel variant the library offers (gpe_variant_tiles /
gpe_set_variant) on the caller's own problem -- same grid, precision, mode
and state -- and keeps the fastest.  Results persist in a JSON file keyed by
(cells, precision, mode, potential), so a later solve of the same shape
skips the search.  Every variant performs identical arithmetic per node, so
the choice changes speed only; the FP64 parity tests check each variant.

The cache file is $VAIKKARI_AUTOTUNE_CACHE, default
~/.cache/vaikkari_amd/autotune.json.
"""
from __future__ import annotations

import json
import os
import time
from typing import Optional

from .solver import Solver, variants

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "vaikkari_amd", "autotune.json")


def cache_path() -> str:
    return os.environ.get("VAIKKARI_AUTOTUNE_CACHE", DEFAULT_CACHE)


def _key(s: Solver, families=None) -> str:
    prec = "fp64" if s.cfg.precision == 0 else "fp32"
    mode = "real" if s.cfg.mode == 0 else "imag"
    pot = "array" if s.cfg.potential == 0 else "trap"
    key = f"{s.dims[0]}x{s.dims[1]}x{s.dims[2]}/{prec}/{mode}/{pot}"
    if families is not None and tuple(families) != DEFAULT_FAMILIES:
        key += "/" + "+".join(sorted(families))  # a tuning over other families is cached apart
    return key


def load_cache(path: Optional[str] = None) -> dict:
    path = path or cache_path()
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def save_cache(data: dict, path: Optional[str] = None) -> None:
    path = path or cache_path()
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


# The fused odd+even family stays selectable but is not tried by default: on
# MI355X its best tiles run at 96 % (FP64) and 69 % (FP32) of the two-pass
# gather rate at 128^3 cells, so it never wins a default tuning.
DEFAULT_FAMILIES = ("gather", "zmarch")


def tune(s: Solver, steps: int = 20, families=DEFAULT_FAMILIES, use_cache: bool = True,
         path: Optional[str] = None) -> dict:
    """Selects the fastest variant of the given kernel families for solver s and leaves it set.

    The solver's state is advanced by the trial steps; callers tune on a
    scratch solver or reset the state afterwards (set_state)."""
    key = _key(s, families)
    cache = load_cache(path) if use_cache else {}
    if key in cache:
        best = cache[key]
        s.set_variant(best["kernel"], best["tile"])
        return dict(best, cached=True)
    results = []
    real = s.cfg.mode == 0
    for v in variants():
        if v["kernel"] not in families:
            continue
        if v["kernel"] == "fused" and not real:
            continue  # imaginary time runs the z-march whatever the family
        try:
            s.set_variant(v["kernel"], v["tile"])
            s.step(2)
            ms, _ = s.time_steps(steps)
        except Exception:
            continue  # e.g. a shape the precision cannot hold in LDS
        results.append({"kernel": v["kernel"], "tile": v["tile"], "ms_per_step": ms / steps})
    if not results:
        raise RuntimeError("autotune: no variant ran")
    best = min(results, key=lambda r: r["ms_per_step"])
    s.set_variant(best["kernel"], best["tile"])
    out = {"kernel": best["kernel"], "tile": best["tile"], "ms_per_step": best["ms_per_step"],
           "site_steps_per_s": s.nodes / (best["ms_per_step"] * 1e-3), "tuned_at": time.time(),
           "candidates": len(results)}
    if use_cache:
        cache = load_cache(path)
        cache[key] = out
        save_cache(cache, path)
    return dict(out, cached=False)


# tile orders: False y-bands (or plain), True a slab of every instance per XCD, 2 a slab of the
# whole group per XCD
BATCH_TILES = tuple((w, c, sl) for sl in (False, True, 2) for c in (False, True) for w in (1, 2, 4, 8)) + \
    tuple((w, c, sl) for sl in (True, 2) for c in (False, True) for w in (3, 6, 12))  # class-group tiles


def _tile_triple(t):
    return (tuple(t) + (False, False))[:3] if isinstance(t, tuple) else (int(t), False, False)


def tune_batch_tile(make_batch, chunk_steps: int = 10, chunks: int = 2, candidates=BATCH_TILES) -> dict:
    """Picks the real-time tile of a batch -- waves per block, streamed or cached
    read-modify-write, and y-band or slab order of the tiles over the XCDs: make_batch(tile) must
    return a populated Batch whose set_tile was given `tile` (an int number of waves, or a
    (waves, cached[, slabs]) tuple); each candidate is timed over chunks x chunk_steps steps."""
    timings = {}
    for t in candidates:
        b = make_batch(t)
        b.run(chunk_steps, 1)  # capture + warm-up
        ms = b.run(chunk_steps, chunks)
        timings[t] = ms
        b.close()
    best = min(timings, key=timings.get)
    waves, cached, slabs = _tile_triple(best)
    return {"waves": waves, "cached": cached, "slabs": slabs, "tile": best,
            "ms": {"%d%s%s" % (w, "c" if c else "", "g" if sl == 2 else "s" if sl else ""): v
                   for (w, c, sl), v in ((_tile_triple(k), v) for k, v in timings.items())}}


BATCH_GROUP_BYTES = (0, 192 << 20, 256 << 20, 384 << 20, 512 << 20, 640 << 20, 1 << 30, 2 << 30)


def tune_batch_groups(batch, chunk_steps: int = 10, chunks: int = 2, candidates=BATCH_GROUP_BYTES) -> dict:
    """Picks the group size of a batch's real-time schedule (Batch.set_group_bytes) by timing
    chunks x chunk_steps steps of `batch` itself under each candidate (every candidate advances
    the instances: give them steps to spare), and leaves the best one set."""
    timings = {}
    for gb in candidates:
        batch.set_group_bytes(gb)
        batch.run(chunk_steps, 1)  # re-grouping, capture + warm-up
        ms = batch.run(chunk_steps, chunks)
        # a run in which no instance had steps left took no time and measures nothing
        timings[gb] = ms if ms > 0 else float("inf")
    if all(v == float("inf") for v in timings.values()):
        raise ValueError("tune_batch_groups: the instances ran out of steps; give them more")
    best = min(timings, key=timings.get)
    batch.set_group_bytes(best)
    return {"group_bytes": best, "ms": {str(k >> 20) + "MiB": v for k, v in timings.items()}}
