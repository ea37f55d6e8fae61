"""GPU test of the gather family's default tile (gpe_config.tile = 0): the tile follows the grid,
the precision and the cache fit, and stepping on it gives the bits of every other tile."""
import numpy as np
import pytest

import vaikkari_amd as va

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"[gpu-case] {msg}", flush=True)


# (dims, precision) -> real-time default tile (1-based)
CASES = [
    ((32, 32, 32), "fp64", 23),   # both fields cached, 512 one-wave blocks: class groups in y-bands
    ((32, 32, 32), "fp32", 23),
    ((24, 24, 24), "fp64", 18),   # cached: one thread per node in FP64 (one-wave rows 12, no y-bands)
    ((24, 24, 24), "fp32", 19),   # cached: one thread per group of 4 classes in FP32
    ((48, 48, 48), "fp32", 19),   # cached
    ((80, 80, 80), "fp64", 18),   # 197 MB of fields: cached
    ((96, 96, 96), "fp32", 19),   # 170 MB: cached
    ((96, 96, 96), "fp64", 24),   # 340 MB, FP64: class groups in y-bands (48 wave rows)
    ((112, 112, 112), "fp32", 7),  # 270 MB; rows 7, 14, 28: no y-bands
    ((128, 128, 128), "fp64", 24),
    ((128, 128, 128), "fp32", 11),
    ((160, 160, 160), "fp64", 24),
    ((160, 160, 160), "fp32", 12),  # 4-wave rows 20, 2-wave rows 40
    ((87, 87, 302), "fp64", 7),     # 44 wave rows: no y-bands; one thread per cell in eighths
    ((112, 112, 112), "fp64", 7),   # 28 wave rows: the same
]


@pytest.mark.parametrize("dims,precision,want", CASES)
def test_default_gather_tile_by_grid(dims, precision, want):
    announce(f"default tile {dims} {precision}")
    h = 0.25
    b = h / np.sqrt(8.0)
    s = va.Solver(dims, b, 1e-3, 500.0, va.centered_origin(dims, b), precision, "real", "trap", 1.0)
    info = s.info()
    s.close()
    assert info["kernel"] == "gather" and info["tile"] == want


# (dims, precision) -> imaginary-time default tile (1-based)
IMAG_CASES = [
    ((32, 32, 32), "fp64", 12),    # 128 four-wave blocks: y-bands of 2 waves
    ((48, 48, 48), "fp64", 19),    # cached, 432 four-wave blocks: cached stores, groups of 4 classes
    ((48, 48, 48), "fp32", 7),     # FP32 needs 700 four-wave blocks for cached stores; no y-bands
    ((56, 56, 56), "fp32", 6),     # cached, 784 four-wave blocks: cached stores, 8 waves
    ((64, 64, 64), "fp64", 19),    # cached, 1024 four-wave blocks
    ((64, 64, 64), "fp32", 6),
    ((96, 96, 96), "fp32", 6),     # 170 MB: cached
    ((96, 96, 96), "fp64", 12),    # 340 MB: y-bands, 8- and 4-wave rows (6, 12) do not divide by 8
    ((128, 128, 128), "fp64", 13),
    ((128, 128, 128), "fp32", 11),
]


@pytest.mark.parametrize("dims,precision,want", IMAG_CASES)
@pytest.mark.parametrize("kernel", ["gather", "fused"])
def test_default_imaginary_time_tile_by_grid(dims, precision, want, kernel):
    """gpe_get_info reports what the current mode runs: in imaginary time the gather family's
    imaginary-time default, for the fused family too."""
    announce(f"default imag tile {dims} {precision} {kernel}")
    h = 0.25
    b = h / np.sqrt(8.0)
    s = va.Solver(dims, b, 1e-3, 500.0, va.centered_origin(dims, b), precision, "imag", "trap", 1.0, kernel)
    info = s.info()
    s.close()
    assert info["kernel"] == "gather" and info["tile"] == want


@pytest.mark.parametrize("n,precision", [(32, "fp64"), (40, "fp64"), (40, "fp32"), (96, "fp64")])
def test_default_tile_steps_equal_other_tiles(n, precision):
    """The defaults (the class-group y-band tile 23 at 32^3, the node tile 18 at 40^3 FP64, the
    class-group tile 19 at 40^3 FP32, the class-group y-band tile 24 at 96^3 FP64) step the same
    bits as tile 11."""
    announce(f"default tile {n}^3 {precision} bits")
    dims, h = (n, n, n), 0.25
    b = h / np.sqrt(8.0)
    origin = va.centered_origin(dims, b)
    psi = va.gaussian(dims, origin, b, 1.5, 1e-3, 5)
    out = []
    for tile in (0, 11):
        s = va.Solver(dims, b, 1e-3, 500.0, origin, precision, "real", "trap", 1.0, "gather", tile)
        s.set_state(psi)
        s.init_odd(10.0)
        s.step(40)
        out.append(s.get_state())
        s.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
