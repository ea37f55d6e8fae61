"""CPU tests of the product library: it loads, exports every symbol the
headers declare, and its host-side setup (geometry, VortexState, the
reference setup, recipes) matches the pinned oracle bit for bit.  No
device compute runs here."""
import ctypes as C
import json
import math
import os
import re

import numpy as np
import pytest

import vaikkari_amd as va
from vaikkari_amd import _lib
from oracle_lib import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HERE = os.path.dirname(os.path.abspath(__file__))
GEO = json.load(open(os.path.join(HERE, "golden", "bcc_geometry_ref.json")))
VSREF = json.load(open(os.path.join(HERE, "golden", "vortex_state_ref.json")))


def declared_functions():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(gpe_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) > 30
    lib = C.CDLL(_lib.LIB_PATH)
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    bound = {s[0] for s in _lib.SIGNATURES}
    assert names <= bound, sorted(names - bound)


def test_geometry_tables():
    np.testing.assert_array_equal(va.bcc_positions(), np.array(GEO["positions"]))
    np.testing.assert_array_equal(va.bcc_stencil().reshape(48, 4), np.array(GEO["neighbours"]))
    n, derived = va.bcc_stencil_derived()
    assert n == 48
    table = va.bcc_stencil()
    for l in range(12):
        assert {tuple(e) for e in derived[l].tolist()} == {tuple(e) for e in table[l].tolist()}


def test_create_rejects_bad_config_before_touching_the_device():
    for kw in ({"dims": (0, 4, 4)}, {"dt": -1.0}, {"block_scale": 0.0}):
        args = dict(dims=(4, 4, 4), block_scale=0.1, dt=1e-3, g=1.0)
        args.update(kw)
        with pytest.raises(_lib.GpeError) as ei:
            va.Solver(**args)
        assert ei.value.status == _lib.EINVAL


def test_vortex_state_mirror(state_dat):
    vs = va.VortexState()
    assert vs.load(state_dat)
    assert vs.getKappa() == VSREF["kappa"] and vs.getG() == VSREF["g"] and vs.getMu() == VSREF["mu"]
    for p, ref in zip(VSREF["psi_points"], VSREF["psi"]):
        c = vs.getPsi(p)
        assert [c.real, c.imag] == ref
    for t, ref in zip(VSREF["phase_times"], VSREF["phase"]):
        c = vs.getPhase(t)
        assert [c.real, c.imag] == ref
    assert vs.searchFunctionMax() == VSREF["search_max"]
    assert vs.searchMaxR(VSREF["eps"]) == VSREF["search_max_r"]
    assert vs.searchMaxZ(VSREF["eps"]) == VSREF["search_max_z"]
    assert vs.integrateCurvature() == VSREF["curvature"]
    assert not va.VortexState().load("/nonexistent/state.dat")


def test_reference_setup_matches_oracle_bitwise(state_dat):
    """Product setup (psi0, V, dt rule, disk noise, odd phase) == oracle restatement, on a
    coarser mesh than the reference default so it runs in a second."""
    vs = va.VortexState()
    vs.load(state_dat)
    eps = 1e-5 * vs.searchFunctionMax()
    mr, mz = vs.searchMaxR(eps), vs.searchMaxZ(eps)
    bs = 0.35
    minp, maxp = (-mr, -mr, -mz), (mr, mr, mz)
    info, even, odd, V = va.reference_setup(vs, bs, minp, maxp, 1.0, 0.1, 54363)
    orc = Oracle()
    h = orc.o.or_vs_load(state_dat.encode())
    dims = (C.c_uint32 * 3)()
    p0 = (C.c_double * 3)()
    orc.o.or_ref_dims(bs, _lib.d3(minp), _lib.d3(maxp), dims, p0)
    assert tuple(dims) == info["dims"] and tuple(p0) == info["origin"]
    n = va.nodes_of(info["dims"])
    ev, od, pot, sc = np.zeros(2 * n), np.zeros(2 * n), np.zeros(n), np.zeros(6)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    assert orc.o.or_ref_setup(h, bs, _lib.d3(minp), _lib.d3(maxp), 1.0, 0.1, 54363, dp(ev), dp(od), dp(pot),
                              dp(sc)) == 0
    np.testing.assert_array_equal(even, ev)
    np.testing.assert_array_equal(odd, od)
    np.testing.assert_array_equal(V * info["dt"], pot)
    assert info["steps_per_iteration"] == int(sc[0]) and info["dt"] == sc[1]
    assert (info["lapfac_dt"], info["lapfac0_dt"], info["maxpot"]) == (sc[2], sc[3], sc[5])
    assert info["g"] * info["dt"] == sc[4]
    orc.o.or_vs_free(h)


def test_node_positions_match_oracle():
    dims, b = (3, 4, 5), 0.0883883476483184
    origin = va.centered_origin(dims, b)
    pos = va.node_positions(dims, origin, b)
    orc = Oracle()
    out = np.zeros(pos.size)
    p0 = np.array(origin)
    orc.o.or_node_positions(p0.ctypes.data_as(C.POINTER(C.c_double)), b, *dims,
                            out.ctypes.data_as(C.POINTER(C.c_double)))
    np.testing.assert_array_equal(pos.reshape(-1), out)
    # nodes are 12 per cell, strictly inside, at q/8 of the cube edge
    a = math.sqrt(8) * b
    assert np.all(np.abs(pos) < 0.5 * a * max(dims))


def test_recipes():
    dims, b = (16, 16, 16), 0.25 / math.sqrt(8)
    origin = va.centered_origin(dims, b)
    g = va.gaussian(dims, origin, b, 1.5, 1e-3, 54363)
    assert va.field_norm(dims, b, g) == pytest.approx(1.0, abs=1e-4)
    g0 = va.gaussian(dims, origin, b, 1.5, 0.0, 1)
    assert va.field_norm(dims, b, g0) == pytest.approx(1.0, abs=1e-13)
    # deterministic per seed, different across seeds
    np.testing.assert_array_equal(g, va.gaussian(dims, origin, b, 1.5, 1e-3, 54363))
    assert not np.array_equal(g, va.gaussian(dims, origin, b, 1.5, 1e-3, 54364))
    # noise statistics: (psi/psi0 - 1)/1e-3 ~ N(0,1) in both parts
    r = g.view(np.complex128) / g0.view(np.complex128)
    xi = np.concatenate([(r.real - 1) / 1e-3, r.imag / 1e-3])
    assert abs(xi.mean()) < 0.02 and abs(xi.std() - 1) < 0.02
    # vortex line keeps |psi| and winds once around (x0, y0)
    line = va.vortex_line(dims, origin, b, 0.2, -0.1, g0.copy())
    np.testing.assert_allclose(np.abs(line.view(np.complex128)), np.abs(g0.view(np.complex128)), rtol=1e-14)
    pos = va.node_positions(dims, origin, b)
    ang = np.angle(line.view(np.complex128))
    np.testing.assert_allclose(np.exp(1j * ang), np.exp(1j * np.arctan2(pos[:, 1] + 0.1, pos[:, 0] - 0.2)),
                               atol=1e-12)
    tf, mu = va.thomas_fermi(dims, origin, b, 500.0, 1.0)
    assert va.field_norm(dims, b, tf) == pytest.approx(1.0, abs=1e-12) and mu > 0
    V = va.trap_potential(dims, origin, b, 1.0)
    dens = tf[0::2] ** 2
    inside = V < mu - 1e-9
    ratio = dens[inside] * 500.0 / (mu - V[inside])  # g|psi|^2 = mu - V inside the cloud
    np.testing.assert_allclose(ratio, 1.0, rtol=1e-6)
    assert np.all(dens[V > mu + 1e-9] == 0)
    ring = va.vortex_ring(dims, origin, b, (0, 0, 0), (0, 0, 1), 1.0, tf.copy())
    np.testing.assert_allclose(np.abs(ring.view(np.complex128)), np.abs(tf.view(np.complex128)), rtol=1e-14)
    odd = va.phase_shift(dims, g, 2.0, 1e-3)
    np.testing.assert_allclose(odd.view(np.complex128), g.view(np.complex128) * np.exp(1j * 2.0 * 1e-3 / 2),
                               rtol=1e-15, atol=1e-18)


def test_vortex_state_from_table_equals_file(tmp_path):
    """gpe_vortex_state_from_table (the in-memory handoff of a VortexState) gives the same state as
    saving those members in the reference's file format and loading it: same parameters, table,
    psi, curvature, and the same integrateInTime setup bit for bit."""
    rng = np.random.default_rng(3)
    f = np.abs(rng.normal(size=(40, 55))) * 0.1
    vs = va.VortexState.from_table(3, 420.0, 0.2, 0.07, 0.09, 6.25, f)
    np.testing.assert_array_equal(vs.table(), f)
    path = str(tmp_path / "t.dat")
    assert vs.save(path)
    back = va.VortexState()
    assert back.load(path)
    np.testing.assert_array_equal(vs._params(), back._params())
    np.testing.assert_array_equal(back.table(), f)
    for p in ((0.5, -0.3, 0.2), (1.1, 0.4, -1.9), (0.0, 0.0, 0.3)):
        assert vs.getPsi(p) == back.getPsi(p)
    assert vs.integrateCurvature() == back.integrateCurvature()
    ext = (-2.5, -2.5, -4.0), (2.5, 2.5, 4.0)
    a = va.reference_setup(vs, 0.3, *ext, 0.05, 0.7)
    b = va.reference_setup(back, 0.3, *ext, 0.05, 0.7)
    assert a[0] == b[0]
    for x, y in zip(a[1:], b[1:]):
        np.testing.assert_array_equal(x, y)
    with pytest.raises(_lib.GpeError):
        va.VortexState.from_table(1, 1.0, 0.0, -0.1, 0.1, 1.0, f)  # non-positive step


@pytest.mark.parametrize("precision,eb", [("fp64", 16), ("fp32", 8)])
def test_field_layout(precision, eb):
    """gpe_get_field_layout: rows start on 128-byte lines and hold nx + 1 elements or more (the
    x halo at nx, and x = -1 as the last padding element of the row before), 12 class rows per
    cell row, a one-cell halo all round, one line in front of the first row; pack_field puts
    node (i, j, k, l) at origin + k sz + j sy + l xp + i and leaves everything else zero."""
    dims = (37, 5, 4)
    L = va.field_layout(dims, precision)
    line = 128 // eb
    assert L["elem_bytes"] == eb and L["xp"] % line == 0 and L["xp"] >= dims[0] + 1
    assert L["xp"] - line < dims[0] + 1  # no line of a row is padding alone
    assert L["sy"] == 12 * L["xp"] and L["sz"] == (dims[1] + 2) * L["sy"]
    assert L["origin"] == L["sz"] + L["sy"] + line and L["elems"] == (dims[2] + 2) * L["sz"] + line
    # x = -1 of the first halo row and x = nx of the last one lie inside the buffer
    assert L["origin"] - L["sz"] - L["sy"] - 1 >= 0
    top = L["origin"] + dims[2] * L["sz"] + dims[1] * L["sy"] + 11 * L["xp"] + dims[0]
    assert top < L["elems"]
    rng = np.random.default_rng(1)
    psi = rng.uniform(-1, 1, 2 * va.nodes_of(dims))
    packed = va.pack_field(dims, psi, precision)
    assert packed.size == 2 * L["elems"] and np.count_nonzero(packed) == psi.size
    i, j, k, l = 36, 4, 3, 11
    n = ((k * dims[1] + j) * dims[0] + i) * 12 + l
    e = L["origin"] + k * L["sz"] + j * L["sy"] + l * L["xp"] + i
    assert packed[2 * e] == np.asarray(psi[2 * n], packed.dtype)
    assert packed[2 * e + 1] == np.asarray(psi[2 * n + 1], packed.dtype)


def test_observables_noalloc_argument_checks():
    """gpe_observables_noalloc refuses, before touching the device, a grid whose block partials
    exceed its 8 MiB scratch (more than 2^18 blocks of 64 x 4 cells per layer), a missing output,
    field or streamed potential, and a bad configuration."""
    cfg = _lib.GpeConfig()
    cfg.nx, cfg.ny, cfg.nz = 512, 512, 512
    cfg.precision, cfg.block_scale, cfg.dt, cfg.g = _lib.FP64, 0.05, 1e-4, 10.0
    cfg.potential, cfg.trap_ratio = _lib.POT_TRAP, 1.0
    dummy = 0x1000  # never dereferenced: every call below is refused first
    with pytest.raises(_lib.GpeError) as ei:
        va.observables_noalloc(cfg, dummy, dummy)
    assert ei.value.status == _lib.EINVAL and "scratch" in str(ei.value)
    cfg.nx, cfg.ny, cfg.nz = 8, 8, 8
    for args in ((0, dummy), (dummy, 0)):
        with pytest.raises(_lib.GpeError):
            va.observables_noalloc(cfg, *args)
    cfg.potential = _lib.POT_ARRAY
    with pytest.raises(_lib.GpeError):
        va.observables_noalloc(cfg, dummy, dummy)
    cfg.potential, cfg.block_scale = _lib.POT_TRAP, -1.0
    with pytest.raises(_lib.GpeError):
        va.observables_noalloc(cfg, dummy, dummy)


def _checkpoint_bytes(dims=(1, 1, 1), precision=0, mode=0, has_odd=True):
    """A checkpoint file in the layout gpe_checkpoint_save writes (header, fields, FNV-1a 64)."""
    import struct

    nodes = 12 * dims[0] * dims[1] * dims[2]
    body = b"VGPECKP1" + struct.pack("<I", 1)
    body += struct.pack("<8I", *dims, precision, 1, mode, 1, 1)
    body += struct.pack("<7d", 0.1, 1e-4, 10.0, -0.1, -0.1, -0.1, 1.0)
    body += struct.pack("<QII", 7, 1 if has_odd else 0, 0)
    rng = np.random.default_rng(5)
    body += rng.uniform(-1, 1, 2 * nodes).tobytes()
    if has_odd:
        body += rng.uniform(-1, 1, 2 * nodes).tobytes()
    h = 1469598103934665603
    for b in body:
        h = ((h ^ b) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return body + struct.pack("<Q", h)


def test_checkpoint_load_rejects_bad_files_before_device_work(tmp_path):
    """gpe_checkpoint_load refuses a missing file, a wrong magic, a short file and a file whose
    checksum does not match with GPE_EIO, before any device work (so on a host without a GPU)."""
    good = _checkpoint_bytes()
    cases = {"missing": None, "magic": b"XXXXXXXX" + good[8:], "short": good[:-9],
             "checksum": good[:200] + bytes([good[200] ^ 1]) + good[201:]}
    for name, data in cases.items():
        path = tmp_path / f"{name}.ckpt"
        if data is not None:
            path.write_bytes(data)
        with pytest.raises(_lib.GpeError) as ei:
            va.Solver.load_checkpoint(str(path))
        assert ei.value.status == _lib.EIO, name
