"""GPU tests of checkpoint and resume (gpe_checkpoint_save / gpe_checkpoint_load): a run stopped,
saved, loaded into a new solver and continued gives the same bits as the run continued without the
interruption -- FP64 and FP32, real and imaginary time, in-kernel trap and streamed potential, the
two-pass and the fused kernel families -- and FP64 also equals the oracle's uninterrupted run."""
import numpy as np
import pytest

import vaikkari_amd as va
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"[gpu-case] {msg}", flush=True)


DIMS, H, G, DT, RATIO = (21, 14, 9), 0.4, 300.0, 6e-4, 1.4


def _start(precision, mode, potential, kernel):
    b = H / np.sqrt(8.0)
    origin = va.centered_origin(DIMS, b)
    s = va.Solver(DIMS, b, DT, G, origin, precision, mode, potential, RATIO, kernel)
    if potential == "array":
        s.set_potential(va.trap_potential(DIMS, origin, b, RATIO))
    even = va.gaussian(DIMS, origin, b, 1.2, 1e-3, 77)
    if mode == "real":
        s.set_state(even, va.phase_shift(DIMS, even, 5.0, DT))
    else:
        s.set_state(even)
    return s, b, origin, even


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("mode", ["real", "imag"])
@pytest.mark.parametrize("potential", ["trap", "array"])
def test_checkpoint_resume_is_bit_identical(tmp_path, precision, mode, potential):
    announce(f"checkpoint {precision} {mode} {potential}")
    a, *_ = _start(precision, mode, potential, "auto")
    a.step(13)
    path = str(tmp_path / "run.ckpt")
    a.save_checkpoint(path)
    a.step(11)
    want = a.get_state()
    cfg_a = a.config()
    a.close()
    b = va.Solver.load_checkpoint(path)
    cfg_b = b.config()
    for k in ("nx", "ny", "nz", "precision", "potential", "mode", "kernel", "tile", "dt", "g", "trap_ratio"):
        assert cfg_b[k] == cfg_a[k], k
    assert b.info()["steps_done"] == 13
    b.step(11)
    got = b.get_state()
    b.close()
    np.testing.assert_array_equal(got[0], want[0])
    if mode == "real":
        np.testing.assert_array_equal(got[1], want[1])


def test_checkpoint_fused_family_and_oracle(tmp_path):
    """A run on the fused (one pass per step) family saved between steps and resumed equals the
    oracle's uninterrupted leapfrog (FP64)."""
    announce("checkpoint fused fp64 vs oracle")
    a, b, origin, even = _start("fp64", "real", "trap", "fused")
    odd = va.phase_shift(DIMS, even, 5.0, DT)
    a.step(5)
    path = str(tmp_path / "fused.ckpt")
    a.save_checkpoint(path)
    a.close()
    r = va.Solver.load_checkpoint(path)
    assert r.config()["kernel"] == 3  # GPE_KERNEL_FUSED
    r.step(9)
    ge, go = r.get_state()
    r.close()
    V = va.trap_potential(DIMS, origin, b, RATIO)
    oe, oo = Oracle().leapfrog(even, odd, V * DT, DIMS, *va.leapfrog_coefs(b, DT, G), 14)
    assert np.array_equal(ge, oe) and np.array_equal(go, oo)
