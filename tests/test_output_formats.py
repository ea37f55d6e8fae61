"""Output files (CPU): the BMP and MetaImage writers against the reference's own Picture::save and
a restatement of saveVolumeMap over the reference's own Text class (oracle/_ref), and the numpy
output oracle pinned by Picture::save and closed forms.

The GPU kernels that produce the pixels and the volume are checked against this oracle in
test_gpu_output.py."""
import os
import struct
import sys

import numpy as np
import pytest

import vaikkari_amd as va
from oracle_lib import load_ref

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import output_oracle as oo  # noqa: E402

REF = load_ref()
needs_ref = pytest.mark.skipif(REF is None, reason="oracle/_ref not built (reference tree absent)")


def ref_bmp(tmp_path, colours, middle):
    h, w = colours.shape[:2]
    c = np.ascontiguousarray(colours, np.float64)
    p = str(tmp_path / "ref.bmp")
    assert REF.ref_picture_save(w, h, c.ctypes.data_as(REF.ref_picture_save.argtypes[2]), p.encode(), int(middle))
    return open(p, "rb").read()


def our_bmp(tmp_path, bgr):
    p = str(tmp_path / "ours.bmp")
    va.write_bmp(p, bgr)
    return open(p, "rb").read()


@needs_ref
@pytest.mark.parametrize("w,h", [(1, 1), (2, 3), (3, 2), (4, 4), (5, 7), (130, 90)])
@pytest.mark.parametrize("middle", [False, True])
def test_bmp_matches_picture_save(tmp_path, w, h, middle):
    """Colours spanning every branch of the byte conversion (below 1, above 255, in range,
    negative, NaN): oracle bytes written by gpe_write_bmp == the reference's Picture::save."""
    rng = np.random.default_rng(w * 1000 + h + (7 if middle else 0))
    col = rng.uniform(-1.5, 1.5, size=(h, w, 4))
    col.reshape(-1, 4)[::7, 0] = np.nan
    col.reshape(-1, 4)[::5, 1] = 1.0 / 256.0          # c == 1 exactly (middle False)
    col.reshape(-1, 4)[::11, 2] = 255.0 / 256.0       # c == 255 exactly
    col.reshape(-1, 4)[::13, 2] = 0.9999999 / 256.0   # just below 1
    assert our_bmp(tmp_path, oo.colour_bytes(col, middle)) == ref_bmp(tmp_path, col, middle)


def test_bmp_layout_closed_form(tmp_path):
    """Header fields and row padding of a 24-bit BMP (the reference's size fields count 4 B/pixel)."""
    w, h = 5, 3
    bgr = np.arange(h * w * 3, dtype=np.uint8).reshape(h, w, 3)
    raw = our_bmp(tmp_path, bgr)
    assert raw[:2] == b"BM"
    size, zero, offs, hdr, W, H, planes, bits, comp, data = struct.unpack("<IIIIiiHHII", raw[2:38])
    assert (offs, hdr, W, H, planes, bits, comp) == (54, 40, w, h, 1, 24, 0)
    assert data == 4 * w * h and size == 54 + 4 * w * h and zero == 0
    row = 3 * w + (-3 * w) % 4
    assert len(raw) == 54 + h * row
    for y in range(h):
        assert raw[54 + y * row: 54 + y * row + 3 * w] == bgr[y].tobytes()
        assert raw[54 + y * row + 3 * w: 54 + (y + 1) * row] == b"\0" * (row - 3 * w)


@needs_ref
@pytest.mark.parametrize("dims,h", [((3, 4, 5), (0.197, 0.197, 0.197)), ((90, 89, 304), (0.19790, 0.19790, 0.19790)),
                                    ((1, 1, 1), (1.0, 2.5, 1e-7)), ((130, 130, 130), (0.28284271247461906,) * 3)])
def test_volume_map_file_matches_reference(tmp_path, dims, h):
    rng = np.random.default_rng(sum(dims))
    vol = rng.integers(0, 65536, size=(dims[2], dims[1], dims[0]), dtype=np.uint16)
    ours, ref = str(tmp_path / "ours.mhd"), str(tmp_path / "ref.mhd")
    va.write_volume_map(ours, vol, h)
    assert REF.ref_save_volume_map(ref.encode(), vol.ctypes.data_as(REF.ref_save_volume_map.argtypes[1]),
                                   dims[0], dims[1], dims[2], *h)
    a, b = open(ours).read(), open(ref).read()
    assert a.replace(ours, "<raw>") == b.replace(ref, "<raw>")
    assert open(ours + ".raw", "rb").read() == open(ref + ".raw", "rb").read() == vol.tobytes()


def test_volume_map_header_closed_form(tmp_path):
    vol = np.zeros((4, 3, 2), np.uint16)
    p = str(tmp_path / "v.mhd")
    va.write_volume_map(p, vol, (0.5, 0.25, 2.0))
    kv = {}
    for line in open(p).read().splitlines():
        k, v = line.split("=", 1)
        assert len(k) == 24 and k.rstrip() == k.strip()
        kv[k.strip()] = v.strip()
    assert kv["DimSize"] == "2 3 4" and kv["ElementType"] == "MET_USHORT" and kv["NDims"] == "3"
    assert [float(x) for x in kv["Offset"].split()] == [-0.5, -0.375, -4.0]
    assert [float(x) for x in kv["ElementSpacing"].split()] == [0.5, 0.25, 2.0]
    assert kv["ElementDataFile"] == p + ".raw"
    assert os.path.getsize(p + ".raw") == 2 * vol.size


def test_volume_oracle_closed_forms():
    """One nonzero (cell, class) lands in its padded cell; class0 reads class 0 only (x 12)."""
    dims = (3, 4, 2)
    n = 12 * 3 * 4 * 2
    for (i, j, k, l) in [(0, 0, 0, 0), (2, 3, 1, 7), (1, 2, 0, 11)]:
        f = np.zeros(2 * n)
        node = ((k * dims[1] + j) * dims[0] + i) * 12 + l
        f[2 * node], f[2 * node + 1] = 0.75, -0.5            # |psi|^2 = 0.8125 exactly
        v = oo.volume_map(f, dims, 1000.0, False)
        assert v[k + 1, j + 1, i + 1] == 812 and v.sum() == 812
        v0 = oo.volume_map(f, dims, 1000.0, True)
        assert v0.sum() == (9750 if l == 0 else 0)
    f = np.zeros(2 * n)
    f[0::2] = 0.5                                              # uniform: s = 12 * 0.25 = 3
    v = oo.volume_map(f, dims, 30000.0, False)
    assert (v[1:-1, 1:-1, 1:-1] == 65535).all() and v.sum() == 65535 * 24
    v = oo.volume_map(f, dims, 1.5, False)
    assert (v[1:-1, 1:-1, 1:-1] == 4).all()
    f[0] = np.nan
    assert oo.volume_map(f, dims, 1.0, False)[1, 1, 1] == 0


def test_slice_oracle_closed_form():
    dims = (2, 2, 3)
    n = 12 * 12
    f = np.zeros(2 * n)
    node = ((1 * 2 + 1) * 2 + 0) * 12 + 0                     # cell (0, 1, 1), class 0
    f[2 * node], f[2 * node + 1] = 0.375, -0.5                 # |psi| = 0.625, all exact
    col = oo.slice_colours(f, dims, 2, 0.5)
    assert col[2, 1].tolist() == [0.1875, 0.3125, -0.25, 0.5]
    b = oo.colour_bytes(col, False)
    assert b[2, 1].tolist() == [0, 80, 48]                      # blue < 1 -> 0, green 80, red 48
    assert (b[0] == 0).all()                                   # halo row
    assert (oo.colour_bytes(col, True)[0] == 128).all()        # halo at 128 with middle


def test_writers_reject_bad_arguments(tmp_path):
    with pytest.raises(Exception):
        va.write_bmp(str(tmp_path / "no_such_dir" / "x.bmp"), np.zeros((2, 2, 3), np.uint8))
    with pytest.raises(Exception):
        va.write_volume_map(str(tmp_path / "no_such_dir" / "x.mhd"), np.zeros((1, 1, 1), np.uint16), (1, 1, 1))
