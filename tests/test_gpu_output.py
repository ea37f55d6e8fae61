"""GPU tests of the output kernels: the device volume map and slice picture of the even field
against the numpy output oracle on the downloaded field (bit-exact), and
integrateInTimeWithOutput's files against the oracle on the same run's fields."""
import math
import os
import sys

import numpy as np
import pytest

import vaikkari_amd as va
from oracle_lib import load_ref

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import output_oracle as oo  # noqa: E402

pytestmark = pytest.mark.gpu
REF = load_ref()


def announce(msg):
    print(f"[gpu-case] {msg}", flush=True)


def random_solver(dims, precision, seed, amp=0.08):
    rng = np.random.default_rng(seed)
    b = 0.25 / math.sqrt(8)
    s = va.Solver(dims, b, 1e-4, 100.0, None, precision, "real", "trap", 1.0)
    n = va.nodes_of(dims)
    even = amp * rng.standard_normal(2 * n)
    s.set_state(even, even)
    return s


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("class0", [False, True])
def test_volume_map_matches_oracle(precision, class0):
    for dims, unit in (((13, 7, 9), 3.0e5), ((70, 5, 3), 2.5e5), ((33, 31, 6), 4.0e5)):
        announce(f"volume map {precision} class0={class0} dims={dims}")
        s = random_solver(dims, precision, sum(dims))
        e = s.get_state(odd=False)
        got = s.volume_map(unit, reference_class0=class0)
        ref = oo.volume_map(e, dims, unit, class0)
        np.testing.assert_array_equal(got, ref)
        assert (ref == 65535).any() and (ref[1:-1, 1:-1, 1:-1] < 65535).any()   # both branches taken
        s.close()


def test_volume_map_cell_placement_and_nan():
    """A field nonzero in one (cell, class) lights exactly that padded cell; a NaN gives 0."""
    dims = (5, 4, 3)
    n = va.nodes_of(dims)
    s = va.Solver(dims, 0.1, 1e-4, 0.0, None, "fp64", "real", "trap", 1.0)
    for (i, j, k, l) in [(0, 0, 0, 0), (4, 3, 2, 11), (2, 1, 1, 6)]:
        announce(f"volume placement {(i, j, k, l)}")
        f = np.zeros(2 * n)
        node = ((k * dims[1] + j) * dims[0] + i) * 12 + l
        f[2 * node], f[2 * node + 1] = 0.75, -0.5
        s.set_state(f, f)
        v = s.volume_map(1000.0)
        assert v[k + 1, j + 1, i + 1] == 812 and int(v.sum()) == 812
        assert int(s.volume_map(1000.0, reference_class0=True).sum()) == (9750 if l == 0 else 0)
    f = np.zeros(2 * n)
    f[0] = np.nan
    s.set_state(f, f)
    assert int(s.volume_map(1.0).sum()) == 0
    with pytest.raises(Exception):
        s.volume_map(-1.0)
    with pytest.raises(Exception):
        s.volume_map(float("inf"))
    s.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_slice_picture_matches_oracle(tmp_path, precision):
    dims = (21, 11, 8)
    announce(f"slice picture {precision}")
    s = random_solver(dims, precision, 4, amp=0.2)
    e = s.get_state(odd=False)
    for k in (0, 1, dims[2] // 2 + 1, dims[2], dims[2] + 1):
        for middle in (False, True):
            got = s.slice_picture(k, 5.0, middle)
            col = oo.slice_colours(e, dims, k, 5.0)
            np.testing.assert_array_equal(got, oo.colour_bytes(col, middle))
            if REF is not None:
                ours, ref = str(tmp_path / "o.bmp"), str(tmp_path / "r.bmp")
                va.write_bmp(ours, got)
                c = np.ascontiguousarray(col)
                assert REF.ref_picture_save(dims[0] + 2, dims[1] + 2, c.ctypes.data_as(REF.ref_picture_save.argtypes[2]),
                                            ref.encode(), int(middle))
                assert open(ours, "rb").read() == open(ref, "rb").read()
    with pytest.raises(Exception):
        s.slice_picture(dims[2] + 2)
    s.close()


def test_integrate_in_time_with_output(tmp_path):
    """generate a VortexState -> integrateInTime with SAVE_PICTURE and SAVE_VOLUME: files for
    iterations 0..N, the first from the setup's even field, the last from the final field."""
    vs = va.VortexState()
    vs.setKappa(2)
    vs.setG(300.0)
    vs.setRange(0.0, 4.0, 4.0, 0.1, 0.1)
    vs.iterateSolution(va.TrapRZ(0.7), 400, 1e-29)
    eps = 1e-5 * vs.searchFunctionMax()
    mr, mz = vs.searchMaxR(eps), vs.searchMaxZ(eps)
    bs, period, iters = 0.15, 0.05, 2
    minp, maxp = (-mr, -mr, -mz), (mr, mr, mz)
    prefix = str(tmp_path) + "/"
    announce("integrateInTime with output")
    run = va.integrateInTimeWithOutput(vs, bs, minp, maxp, period, iters, trap_ratio=0.7, prefix=prefix)
    info, even, _, _ = va.reference_setup(vs, bs, minp, maxp, period, 0.7)
    assert run == info
    _, _, ge, _ = va.integrateInTime(vs, bs, minp, maxp, period, iters, trap_ratio=0.7, return_fields=True)
    dims = info["dims"]
    unit = 60000.0 / (12 * vs.searchFunctionMax() * vs.searchFunctionMax())
    k = dims[2] // 2 + 1
    for it, field in ((0, even), (iters, ge)):
        bmp = str(tmp_path / "expect.bmp")
        va.write_bmp(bmp, oo.colour_bytes(oo.slice_colours(field, dims, k, 5.0), False))
        assert open(prefix + f"kuva{it}.bmp", "rb").read() == open(bmp, "rb").read()
        raw = np.fromfile(prefix + f"volume{it}.mhd.raw", dtype="<u2")
        np.testing.assert_array_equal(raw, oo.volume_map(field, dims, unit, False).reshape(-1))
        assert f"DimSize                 = {dims[0] + 2} {dims[1] + 2} {dims[2] + 2}" in open(
            prefix + f"volume{it}.mhd").read()
    assert os.path.exists(prefix + "kuva1.bmp") and os.path.exists(prefix + "volume1.mhd")
    assert not os.path.exists(prefix + f"kuva{iters + 1}.bmp")
