"""Writes the golden fixtures of tests/golden/ from the reference tree.

Run in a container that has the reference mounted at /root/reference and
oracle/_ref/libvsref.so built (``make -C oracle``):

    python tests/golden/make_golden.py

Fixtures written (data only: inputs and the reference's outputs):
  bcc_geometry_ref.json  class positions and the 48-entry neighbour table as
                         the reference source states them
                         (GPUGrossPitaevskii_BCC.cu:77-88, 93-140), byte-offset
                         expressions rewritten as cell offsets, plus the
                         hard-coded steps_per_iteration of
                         GPUGrossPitaevskii_BCC_multi.cu:473.
  vortex_state_ref.json  outputs of the reference's own VortexState.cpp and
                         Random.cpp (compiled from the reference tree into
                         oracle/_ref) on state.dat at seeded points.
  state_files.json       size, SHA-256 and header of the shipped state.dat and
                         state2.dat, and what main's state recipe
                         (GPUGrossPitaevskii_BCC.cu:516-535), run with the
                         reference's own VortexState.cpp, produces for each:
                         the SHA-256 of the saved state and the extent
                         (minr, maxr, maxz) main passes on.
"""
from __future__ import annotations

import hashlib
import json
import os
import struct
import tempfile
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))

REF = "/root/reference"
BCC = os.path.join(REF, "koodit/GrossPitaevskii/GPUGrossPitaevskii_BCC.cu")
MULTI = os.path.join(REF, "koodit/GrossPitaevskii/GPUGrossPitaevskii_BCC_multi.cu")
STATE = os.path.join(REF, "koodit/GrossPitaevskii/state.dat")
STATE2 = os.path.join(REF, "koodit/GrossPitaevskii/state2.dat")


def parse_geometry():
    text = open(BCC).read().splitlines()
    pos = {}
    for ln in text[76:88]:
        m = re.search(r"pos\[(\d+)\]\s*=\s*SQ1_8\s*\*\s*Vector3\((\d+),\s*(\d+),\s*(\d+)\)", ln)
        if m:
            pos[int(m.group(1))] = [int(m.group(2)), int(m.group(3)), int(m.group(4))]
    table = {}
    for ln in text[92:140]:
        m = re.search(r"ind\[(\d+)\]\s*=\s*make_int2\(([^,]+),\s*(\d+)\)", ln)
        if not m:
            continue
        expr = m.group(2).replace(" ", "")
        off = {"nx": 0, "ny": 0, "nz": 0}
        if expr != "0":
            for sign, name in re.findall(r"([+-]?)(n[xyz])", expr):
                off[name] += -1 if sign == "-" else 1
        table[int(m.group(1))] = [off["nx"], off["ny"], off["nz"], int(m.group(3))]
    hodge = None
    for ln in text[140:142]:
        m = re.search(r"return\s+([0-9.]+);", ln)
        if m:
            hodge = float(m.group(1))
    steps = None
    for ln in open(MULTI).read().splitlines()[470:476]:
        m = re.search(r"steps_per_iteration\s*=\s*(\d+);", ln)
        if m:
            steps = int(m.group(1))
    assert len(pos) == 12 and len(table) == 48 and hodge is not None and steps is not None
    return {
        "source": "koodit/GrossPitaevskii/GPUGrossPitaevskii_BCC.cu:77-88,93-141; "
                  "GPUGrossPitaevskii_BCC_multi.cu:473",
        "positions": [pos[i] for i in range(12)],
        "neighbours": [table[i] for i in range(48)],
        "hodge": hodge,
        "state_dat_steps_per_iteration": steps,
    }


def vortex_state_values():
    from oracle_lib import load_ref
    import ctypes as C

    ref = load_ref()
    assert ref is not None, "build oracle/_ref first: make -C oracle"
    h = ref.ref_vs_load(STATE.encode())
    assert h
    rng = np.random.default_rng(20240611)
    pts = np.concatenate([rng.uniform(-9, 9, size=(150, 2)), rng.uniform(-31, 31, size=(150, 1))], axis=1)
    pts[:5] = [[0.0, 0.0, 0.0], [1e-31, 0.0, 1.0], [3.0, 4.0, 0.0], [8.0, 0.5, -29.9], [-2.0, 2.0, 40.0]]
    out = np.zeros(2)
    psi = []
    for p in pts:
        ref.ref_vs_psi(h, float(p[0]), float(p[1]), float(p[2]), out.ctypes.data_as(C.POINTER(C.c_double)))
        psi.append([float(out[0]), float(out[1])])
    rz = np.stack([rng.uniform(0, 9, 100), rng.uniform(-31, 31, 100)], axis=1)
    func = [float(ref.ref_vs_function(h, float(r), float(z))) for r, z in rz]
    times = [-0.5 * 1.447178e-3, 0.0, 0.1, 1.0, -3.5]
    phase = []
    for t in times:
        ref.ref_vs_phase(h, t, out.ctypes.data_as(C.POINTER(C.c_double)))
        phase.append([float(out[0]), float(out[1])])
    fmax = ref.ref_vs_search_max(h)
    eps = 1e-5 * fmax
    circles = np.zeros(2 * 256)
    ref.ref_random_circles(54363, 256, circles.ctypes.data_as(C.POINTER(C.c_double)))
    uints = (C.c_uint32 * 1000)()
    ref.ref_random_uints(5489, 1000, uints)
    res = {
        "source": "reference VortexState.cpp / Random.cpp compiled from the reference tree (oracle/_ref) on state.dat",
        "kappa": ref.ref_vs_kappa(h), "g": ref.ref_vs_g(h), "mu": ref.ref_vs_mu(h),
        "search_max": fmax, "eps": eps,
        "search_max_r": ref.ref_vs_search_max_r(h, eps), "search_max_z": ref.ref_vs_search_max_z(h, eps),
        "curvature": ref.ref_vs_curvature(h),
        "psi_points": pts.tolist(), "psi": psi,
        "function_points": rz.tolist(), "function": func,
        "phase_times": times, "phase": phase,
        "circles_seed54363": circles.tolist(),
        "uints_seed5489_first1000": list(uints),
    }
    ref.ref_vs_free(h)
    return res


def ref_recipe(ref, kappa, g, ratio, path):
    """main's state recipe (GPUGrossPitaevskii_BCC.cu:516-535, 3-D branch) driven through the
    reference's own VortexState.cpp; saves the fine state to path, returns (minr, maxr, maxz)."""
    s0 = ref.ref_vs_create()
    ref.ref_vs_set_physics(s0, kappa, g)
    ref.ref_vs_set_range(s0, 0.0, 15.0, 35.0, 0.2, 0.2)
    ref.ref_vs_iterate(s0, ratio, 10000, 1e-29)
    eps = 1e-5 * ref.ref_vs_search_max(s0)
    minr, maxr, maxz = ref.ref_vs_search_min_r(s0, eps), ref.ref_vs_search_max_r(s0, eps), ref.ref_vs_search_max_z(s0, eps)
    s = ref.ref_vs_create()
    ref.ref_vs_set_physics(s, kappa, g)
    ref.ref_vs_set_range(s, minr, maxr, maxz, 0.03, 0.03)
    ref.ref_vs_initialize(s, s0)
    ref.ref_vs_iterate(s, ratio, 10000, 1e-29)
    assert ref.ref_vs_save(s, path.encode()) == 1
    ext = (minr, ref.ref_vs_search_max_r(s, eps), ref.ref_vs_search_max_z(s, eps))
    ref.ref_vs_free(s0)
    ref.ref_vs_free(s)
    return ext


def state_files():
    from oracle_lib import load_ref

    ref = load_ref()
    assert ref is not None, "build oracle/_ref first: make -C oracle"
    out = {"source": "koodit/GrossPitaevskii/state.dat, state2.dat; recipe GPUGrossPitaevskii_BCC.cu:516-535 "
                     "run with the reference's VortexState.cpp (oracle/_ref)"}
    # (file, kappa, g, RATIO): state.dat is main's configuration (BCC.cu:13-15); state2.dat's header
    # gives kappa 10 and g 5000, RATIO 0.1 as in every GrossPitaevskii source.
    for path, kappa, g, ratio in ((STATE, 20, 5000.0, 0.1), (STATE2, 10, 5000.0, 0.1)):
        raw = open(path, "rb").read()
        hdr = struct.unpack("<IdIIdddd", raw[:52])
        with tempfile.TemporaryDirectory() as d:
            regen = os.path.join(d, "state.dat")
            ext = ref_recipe(ref, kappa, g, ratio, regen)
            rraw = open(regen, "rb").read()
        out[os.path.basename(path)] = {
            "bytes": len(raw), "sha256": hashlib.sha256(raw).hexdigest(),
            "header": dict(zip(["kappa", "g", "rsize", "zsize", "rplug", "rstep", "zstep", "mu"], hdr)),
            "recipe": {"kappa": kappa, "g": g, "ratio": ratio},
            "recipe_sha256": hashlib.sha256(rraw).hexdigest(), "recipe_bytes": len(rraw),
            "recipe_extent": list(ext),
            "recipe_reproduces_file": rraw == raw,
        }
    return out


def main():
    geo = parse_geometry()
    with open(os.path.join(HERE, "bcc_geometry_ref.json"), "w") as f:
        json.dump(geo, f, indent=1)
    vs = vortex_state_values()
    with open(os.path.join(HERE, "vortex_state_ref.json"), "w") as f:
        json.dump(vs, f)
    sf = state_files()
    with open(os.path.join(HERE, "state_files.json"), "w") as f:
        json.dump(sf, f, indent=1)
    print("wrote bcc_geometry_ref.json, vortex_state_ref.json, state_files.json")


if __name__ == "__main__":
    main()
