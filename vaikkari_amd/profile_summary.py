"""Summaries of rocprofv3 SQLite output (``<dir>/<name>_results.db``).

    python -m vaikkari_amd.profile_summary stats <kernel-trace db>
    python -m vaikkari_amd.profile_summary traffic <FETCH_SIZE db> <WRITE_SIZE db>
    python -m vaikkari_amd.profile_summary traffic-json <FETCH_SIZE db> <WRITE_SIZE db> <out.json> [workload]
    python -m vaikkari_amd.profile_summary counters <pmc db> [kernel-name substring]

``stats`` prints calls / total / average duration per kernel; ``counters``
prints the mean per dispatch of every counter of a --pmc run, per kernel.  ``traffic``
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
    n = name.replace("(anonymous namespace)::", "").replace("void ", "").replace("gpe::", "")
    return n.split("(")[0]


def stats(db: str):
    con = sqlite3.connect(db)
    rows = con.execute("select name, total_calls, total_duration, average, percentage from top_kernels").fetchall()
    print(f"{'kernel':60s} {'calls':>6s} {'total_us':>12s} {'avg_us':>10s} {'%':>6s}")
    for name, calls, tot, avg, pct in rows:
        print(f"{short(name):60s} {calls:6d} {tot:12.1f} {avg:10.2f} {pct:6.2f}")
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
    """Per-kernel HBM bytes per dispatch, each counter scaled by its reading on the copy kernel."""
    f = _counter(db_fetch, "FETCH_SIZE")
    w = _counter(db_write, "WRITE_SIZE")
    cal = {}
    for name, per in (("fetch", f), ("write", w)):
        for (kname, grid), vals in per.items():
            if "k_copy_stream" in kname:
                cal[name] = (grid, sum(vals) / len(vals))
    if "fetch" not in cal or "write" not in cal:
        raise SystemExit("no k_copy_stream dispatch in the counter databases: cannot calibrate")
    return f, w, cal


def traffic_table(db_fetch: str, db_write: str, copy_bytes: float = float(1 << 30)):
    """copy_bytes: what k_copy_stream reads and writes per dispatch (bench.py copies 1 GiB)."""
    f, w, cal = traffic(db_fetch, db_write)
    kf = copy_bytes / (cal["fetch"][1] * 1024.0)
    kw = copy_bytes / (cal["write"][1] * 1024.0)
    rows = []
    for (name, grid), vals in f.items():
        wv = w.get((name, grid))
        if not wv:
            continue
        fb = sum(vals) / len(vals) * 1024.0 * kf
        wb = sum(wv) / len(wv) * 1024.0 * kw
        rows.append((short(name), grid, fb, wb, len(vals)))
    rows.sort(key=lambda r: -(r[2] + r[3]))
    return kf, kw, rows


def key(name: str) -> str:
    """Kernel name without spaces, e.g. k_gather<double,0,1,0,4,true>."""
    return short(name).replace(" ", "")


def traffic_json(db_fetch: str, db_write: str, out: str, workload: str = ""):
    """The traffic table as JSON (bytes per dispatch per kernel), the form bench.py reads."""
    import json

    kf, kw, rows = traffic_table(db_fetch, db_write)
    doc = {"workload": workload,
           "calibration": {"kernel": "k_copy_stream 1 GiB", "fetch_scale": kf, "write_scale": kw},
           "kernels": {}}
    for name, grid, fb, wb, n in rows:
        doc["kernels"].setdefault(key(name), {"grid": grid, "read_bytes": fb, "write_bytes": wb, "dispatches": n})
    with open(out, "w") as fh:
        json.dump(doc, fh, indent=1)
    return doc


def counters(db: str, match: str = ""):
    """Mean value per dispatch of every collected counter, per kernel whose name contains `match`."""
    con = sqlite3.connect(db)
    rows = con.execute("select kernel_name, counter_name, value from counters_collection").fetchall()
    per = defaultdict(lambda: defaultdict(list))
    for name, cname, val in rows:
        if match in name:
            per[short(name)][cname].append(val)
    names = sorted({c for k in per.values() for c in k})
    print("kernel," + ",".join(names) + ",dispatches")
    for k, cs in sorted(per.items()):
        n = max(len(v) for v in cs.values())
        print(k.replace(",", ";") + "," + ",".join(f"{sum(cs[c]) / len(cs[c]):.6g}" if cs.get(c) else "" for c in names)
              + f",{n}")
    return per


def main(argv):
    if len(argv) >= 2 and argv[0] == "stats":
        stats(argv[1])
    elif len(argv) >= 2 and argv[0] == "counters":
        counters(argv[1], argv[2] if len(argv) > 2 else "")
    elif len(argv) >= 4 and argv[0] == "traffic-json":
        traffic_json(argv[1], argv[2], argv[3], " ".join(argv[4:]))
    elif len(argv) >= 3 and argv[0] == "traffic":
        kf, kw, rows = traffic_table(argv[1], argv[2])
        print(f"calibration on the 1 GiB copy kernel: FETCH_SIZE x {kf:.4f}, WRITE_SIZE x {kw:.4f}")
        print(f"{'kernel':60s} {'grid':>9s} {'read_MB':>10s} {'write_MB':>10s} {'n':>4s}")
        for name, grid, fb, wb, n in rows:
            print(f"{name:60s} {grid:9d} {fb / 1e6:10.1f} {wb / 1e6:10.1f} {n:4d}")
    else:
        print(__doc__)
        return 2
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
