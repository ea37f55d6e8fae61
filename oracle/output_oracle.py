"""CPU restatement of the reference's picture and volume output (numpy).

TEST INFRASTRUCTURE ONLY.  Only tests/ may import this module, and only as the
checker; the product library never imports anything under oracle/.

Restated (reference = markus-kivioja/vaikkari):
  * volume map   koodit/GrossPitaevskii/GPUGrossPitaevskii_BCC.cu:446-465: per cell of the padded
                 host copy (x fastest), sum += |psi|^2 over the classes, sum *= unit,
                 vol = sum > 65535 ? 65535 : ushort(sum).  The reference reads values[0] inside the
                 class loop (line 458); class0=True restates that, class0=False sums all 12 classes.
  * slice colour GPUGrossPitaevskii_BCC.cu:426-437: colour = 5 * (Re psi_0, |psi_0|, Im psi_0, 1) on
                 padded layer k = zsize/2 + 1 (the gain and the layer are arguments here).
  * colour bytes jukan_koodit/Output/Picture.cpp:123-126: c = 256 colour (middle False) or
                 128 colour + 128; byte = 0 if c < 1, 255 if c > 255, else char(c).  The reference
                 converts through a signed char; for 1 <= c <= 255 that is trunc(c) as gcc builds it
                 for x86-64, and a NaN falls through both tests to 0.

Pinned in tests/ by the reference's own Picture::save (oracle/_ref) on the same colours, and by
closed forms: a field nonzero in one cell and one class lands at that padded cell with the
expected value, and a uniform field gives trunc(unit * 12 |c|^2) everywhere inside.

Host arrays use the library's order [k][j][i][l], complex interleaved.
"""
from __future__ import annotations

import numpy as np


def padded_classes(field: np.ndarray, dims) -> np.ndarray:
    """[nz+2, ny+2, nx+2, 12] complex128 with a zero halo, from a host-order field."""
    nx, ny, nz = dims
    c = field[0::2] + 1j * field[1::2]
    out = np.zeros((nz + 2, ny + 2, nx + 2, 12), np.complex128)
    out[1:-1, 1:-1, 1:-1, :] = c.reshape(nz, ny, nx, 12)
    return out


def volume_map(field: np.ndarray, dims, unit: float, class0: bool) -> np.ndarray:
    """uint16 [nz+2, ny+2, nx+2]; FP64 sums in class order (GPUGrossPitaevskii_BCC.cu:452-461)."""
    p = padded_classes(field, dims)
    s = np.zeros(p.shape[:3])
    for l in range(12):
        v = p[..., 0] if class0 else p[..., l]
        s = s + (v.real * v.real + v.imag * v.imag)
    s = s * unit
    out = np.zeros(s.shape, np.uint16)
    ok = s >= 0.0
    out[ok] = np.where(s[ok] > 65535.0, 65535.0, np.trunc(s[ok])).astype(np.uint16)
    return out


def slice_colours(field: np.ndarray, dims, k: int, gain: float) -> np.ndarray:
    """float64 [ny+2, nx+2, 4] colours gain*(Re, |psi|, Im, 1) of class 0 on padded layer k."""
    p = padded_classes(field, dims)[k, :, :, 0]
    re, im = p.real.copy(), p.imag.copy()
    nrm = np.sqrt(re * re + im * im)
    return np.stack([gain * re, gain * nrm, gain * im, np.full_like(re, gain)], axis=-1)


def colour_bytes(colours: np.ndarray, middle: bool) -> np.ndarray:
    """uint8 [..., 3] in the order blue, green, red (Picture.cpp:123-130)."""
    c = 128.0 * colours[..., :3] + 128.0 if middle else 256.0 * colours[..., :3]
    b = np.zeros(c.shape, np.uint8)
    hi = c > 255.0
    mid = (c >= 1.0) & ~hi
    b[hi] = 255
    b[mid] = np.trunc(c[mid]).astype(np.uint8)
    return b[..., ::-1].copy()
