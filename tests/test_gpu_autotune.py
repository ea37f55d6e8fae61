"""GPU tests of the autotuner: it picks a variant that runs, persists it, reuses it, and the
tuned solver still matches the oracle bit for bit (FP64)."""
import math

import numpy as np
import pytest

import vaikkari_amd as va
from vaikkari_amd import autotune
from vaikkari_amd.batch import Batch, instance_config
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu


def test_tune_persist_and_parity(tmp_path):
    path = str(tmp_path / "tune.json")
    dims, h, g, dt = (40, 24, 30), 0.3, 400.0, 8e-4
    b = h / math.sqrt(8)
    origin = va.centered_origin(dims, b)
    even = va.gaussian(dims, origin, b, 1.2, 1e-3, 9)
    odd = va.phase_shift(dims, even, 6.0, dt)
    s = va.Solver(dims, b, dt, g, origin, "fp64", "real", "trap", 1.0)
    s.set_state(even, odd)
    r1 = autotune.tune(s, steps=4, path=path)
    assert not r1["cached"] and r1["candidates"] >= 10
    r2 = autotune.tune(s, steps=4, path=path)
    assert r2["cached"] and (r2["kernel"], r2["tile"]) == (r1["kernel"], r1["tile"])
    assert r1["kernel"] in autotune.DEFAULT_FAMILIES
    # the fused family is tuned on request only, and cached apart from the default tuning
    rf = autotune.tune(s, steps=2, path=path, families=("fused",))
    assert not rf["cached"] and rf["kernel"] == "fused"
    assert autotune.tune(s, steps=2, path=path)["cached"]
    assert s.info()["kernel"] == r1["kernel"]
    s.set_state(even, odd)
    s.step(10)
    ge, go = s.get_state()
    V = va.trap_potential(dims, origin, b, 1.0)
    lf, lf0, gdt = va.leapfrog_coefs(b, dt, g)
    oe, oo = Oracle().leapfrog(even, odd, V * dt, dims, lf, lf0, gdt, 10)
    np.testing.assert_array_equal(ge, oe)
    np.testing.assert_array_equal(go, oo)
    s.close()


def test_batch_tile_choice_does_not_change_results():
    cfg = instance_config((33, 21, 27), 0.3, 700.0, 6e-4, 25, "real", 1.6, seed=4, mu0=3.0)
    outs = []
    for w in (1, 2, 4, 8, (1, True), (2, True), (8, True), (1, False, True), (4, False, True), (8, True, True),
              3, (6, True, True), (12, False, 2), (3, True, 2)):
        B = Batch("fp32")
        B.set_tile(w)
        i = B.add(cfg)
        B.run(25, 1)
        outs.append(B.get_state(i))
        B.close()
    for e, o in outs[1:]:
        np.testing.assert_array_equal(e, outs[0][0])
        np.testing.assert_array_equal(o, outs[0][1])
    res = autotune.tune_batch_tile(lambda w: _mk(w, cfg), chunk_steps=5, chunks=1)
    assert res["waves"] in (1, 2, 4, 8, 3, 6, 12) and res["cached"] in (False, True) and res["slabs"] in (False, True, 2)


def test_batch_slab_order_mixed_instances():
    """Slab order per instance and per group over instances whose tile counts are and are not
    multiples of 8, one of them a y-band instance (16 block rows) under the default order, with one
    thread per cell or per group of 4 classes: states equal the default order's."""
    cfgs = [instance_config((16, 128, 6), 0.3, 300.0, 6e-4, 12, "real", 1.0, seed=1, mu0=2.0),
            instance_config((33, 21, 27), 0.3, 700.0, 6e-4, 9, "real", 1.6, seed=2, mu0=3.0),
            instance_config((5, 7, 3), 0.3, 100.0, 6e-4, 12, "real", 2.0, seed=3, mu0=1.0)]
    runs = []
    for waves, slabs in ((4, False), (4, True), (4, 2), (6, False), (6, True), (3, 2)):
        B = Batch("fp32")
        B.set_tile(waves, False, slabs)
        ids = [B.add(c) for c in cfgs]
        B.run(5, 3)
        runs.append([B.get_state(i) for i in ids])
        B.close()
    for other in runs[1:]:
        for (e0, o0), (e1, o1) in zip(runs[0], other):
            np.testing.assert_array_equal(e1, e0)
            np.testing.assert_array_equal(o1, o0)


def _mk(w, cfg):
    B = Batch("fp32")
    B.set_tile(w)
    B.add(cfg)
    return B
