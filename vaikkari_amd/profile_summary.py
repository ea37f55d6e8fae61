"""Summaries of rocprofv3 SQLite output (``<dir>/<name>_results.db``).

    python -m vaikkari_amd.profile_summary stats <kernel-trace db>
    python -m vaikkari_amd.profile_summary traffic <FETCH_SIZE db> <WRITE_SIZE db>

``stats`` prints calls / total / average duration per kernel.  ``traffic``
prints HBM bytes per dispatch for every kernel from the FETCH_SIZE and
WRITE_SIZE counters (KiB per dispatch), scaled by what the same counters read
for the plain copy kernel k_copy_stream, whose true traffic is known (it reads
and writes exactly its buffer size once; the size is taken from its grid,
16 bytes per work-item).
"""
from __future__ import annotations

import sqlite3
import sys
from collections import defaultdict


def short(name: str) -> str:
    n = name.split("(")[0]
    return n.replace("void ", "").replace("gpe::", "").replace("(anonymous namespace)::", "")


def stats(db: str):
    con = sqlite3.connect(db)
    rows = con.execute("select name, total_calls, total_duration, average, percentage from top_kernels").fetchall()
    print(f"{'kernel':70s} {'calls':>6s} {'total_us':>12s} {'avg_us':>10s} {'%':>6s}")
    for name, calls, tot, avg, pct in rows:
        print(f"{short(name):70s} {calls:6d} {tot / 1e3:12.1f} {avg / 1e3:10.2f} {pct:6.2f}")
    return rows


def _counter(db: str, counter: str):
    con = sqlite3.connect(db)
    rows = con.execute("select kernel_name, grid_size, value from counters_collection where counter_name = ?",
                       (counter,)).fetchall()
    per = defaultdict(list)
    for name, grid, val in rows:
        per[(name, grid)].append(val)
    return per


def traffic(db_fetch: str, db_write: str):
    f = _counter(db_fetch, "FETCH_SIZE")
    w = _counter(db_write, "WRITE_SIZE")
    # calibration from the copy kernel: bytes = 16 * work-items, read once and written once
    cf = cw = None
    for (name, grid), vals in f.items():
        if "k_copy_stream" in name:
            vals = sorted(vals)[1:] or vals  # first dispatch is a warm-up
            known = 16.0 * 0  # filled below
            cf = (name, grid, sum(vals) / len(vals))
    for (name, grid), vals in w.items():
        if "k_copy_stream" in name:
            vals = sorted(vals)[1:] or vals
            cw = (name, grid, sum(vals) / len(vals))
    if cf is None or cw is None:
        raise SystemExit("no k_copy_stream dispatch in the counter databases: cannot calibrate")
    out = {}
    print("copy-kernel calibration: FETCH_SIZE %.0f KiB, WRITE_SIZE %.0f KiB per dispatch" % (cf[2], cw[2]))
    return f, w, cf, cw, out


def main(argv):
    if len(argv) >= 2 and argv[0] == "stats":
        stats(argv[1])
    elif len(argv) >= 3 and argv[0] == "traffic":
        f, w, cf, cw, _ = traffic(argv[1], argv[2])
        copy_bytes = None
        for (name, grid), vals in f.items():
            pass
        print(f"{'kernel':70s} {'grid':>9s} {'fetch_KiB':>12s} {'write_KiB':>12s}")
        for (name, grid), vals in sorted(f.items(), key=lambda kv: -max(kv[1])):
            wv = w.get((name, grid), [float("nan")])
            print(f"{short(name):70s} {grid:9d} {sum(vals) / len(vals):12.0f} {sum(wv) / len(wv):12.0f}")
    else:
        print(__doc__)
        return 2
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
