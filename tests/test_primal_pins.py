"""Pins of the 14-point primal BCC oracle (oracle/primal_oracle.c), on the CPU.

No reference stepper uses this operator, so the oracle is pinned by
  * the reference's own mesh library: the DEC Laplacian rows
    star0^-1 d0^T star1 d0 of DelaunayMesh::createBccGrid
    (jukan_koodit/Mesh/DelaunayMesh.cpp:178, Mesh.cpp:1479-1489), compiled
    from the reference sources into oracle/_ref/libmeshref.so;
  * closed forms of the operator on polynomials and plane waves (which fix
    the diagonal and axial weights separately);
  * physics of the integrators: the harmonic-oscillator ground state
    (E0 -> 3/2 with second-order convergence, virial K = P) and the phase
    e^{-i mu t} of a stationary state under the leapfrog.
The product's host helpers (coefficients, positions, trap) are checked
against independent numpy evaluations; they make no GPU call.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle_lib import PrimalOracle, load_mesh_ref, mesh_stencils


@pytest.fixture(scope="module")
def po():
    return PrimalOracle()


def site_grid(dims, origin, a):
    """x, y, z of every site in [k][j][i][s] order."""
    nx, ny, nz = dims
    k, j, i, s = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), np.arange(2), indexing="ij")
    off = 0.5 * s
    x = origin[0] + a * (i + off)
    y = origin[1] + a * (j + off)
    z = origin[2] + a * (k + off)
    return x.reshape(-1), y.reshape(-1), z.reshape(-1)


def np_trap(dims, origin, a, ratio):
    x, y, z = site_grid(dims, origin, a)
    rr = ratio * ratio
    return 0.5 * (x * x + y * y + rr * z * z)


def interior_mask(dims, margin=1):
    """Sites whose 14 neighbours all lie inside the box (cells margin..n-1-margin)."""
    nx, ny, nz = dims
    k, j, i, s = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), np.arange(2), indexing="ij")
    m = (i >= margin) & (i < nx - margin) & (j >= margin) & (j < ny - margin) & (k >= margin) & (k < nz - margin)
    return m.reshape(-1)


def centered(dims, a):
    return [-0.5 * a * (d - 0.5) for d in dims]


# ---------------------------------------------------------------------------
# the reference's mesh library
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def meshref():
    m = load_mesh_ref()
    if m is None:
        pytest.skip("oracle/_ref/libmeshref.so not built (reference tree absent)")
    return m


@pytest.mark.parametrize("h", [1.0, 0.7])
def test_reference_mesh_rows_are_the_14_point_stencil(meshref, h):
    pos, cnt, nb = mesh_stencils(meshref, h, 2.5 * h)
    inner = np.all(np.abs(pos) <= 1.0 * h + 1e-9, axis=1)
    assert inner.sum() >= 27
    for n in np.where(inner)[0]:
        assert cnt[n] == 14
        d = nb[n, :14, :3]
        w = nb[n, :14, 3]
        diag = np.all(np.isclose(np.abs(d), 0.5 * h, rtol=0, atol=1e-12), axis=1)
        axial = np.isclose(np.linalg.norm(d, axis=1), h, rtol=0, atol=1e-12) & (
            np.sum(np.isclose(np.abs(d), 0.0, atol=1e-12), axis=1) == 2)
        assert diag.sum() == 8 and axial.sum() == 6
        np.testing.assert_allclose(w[diag], 0.75 / h ** 2, rtol=1e-12)
        np.testing.assert_allclose(w[axial], 0.25 / h ** 2, rtol=1e-12)


def test_reference_mesh_laplacian_equals_oracle(meshref, po):
    """sum_e w_e (f_j - f_i) from the reference's mesh equals the oracle's L f
    at every interior node, for a smooth non-polynomial f."""
    h = 0.8
    pos, cnt, nb = mesh_stencils(meshref, h, 2.5 * h)
    # createBccGrid puts the cube corners at p0 + h i with p0 = -size h / 2
    size = int(5.0 - 1e-8) + 1
    p0 = -0.5 * size * h
    dims = (size + 1, size + 1, size + 1)
    origin = (p0, p0, p0)

    def f(x, y, z):
        return np.sin(0.7 * x + 0.3) * np.cos(0.4 * y) + 0.2 * z * z + 0.1 * x * y * z

    x, y, z = site_grid(dims, origin, h)
    Lf = po.laplacian(f(x, y, z), dims, h)
    checked = 0
    for n in np.where(np.all(np.abs(pos) <= 1.2 * h, axis=1))[0]:
        p = pos[n]
        u = (p - p0) / h
        s = 0 if np.allclose(u, np.round(u), atol=1e-9) else 1
        ijk = np.round(u - 0.5 * s).astype(int)
        assert np.allclose(u - 0.5 * s, ijk, atol=1e-9)
        e = ((ijk[2] * dims[1] + ijk[1]) * dims[0] + ijk[0]) * 2 + s
        fp = f(*p)
        ref = sum(nb[n, q, 3] * (f(*(p + nb[n, q, :3])) - fp) for q in range(cnt[n]))
        assert abs(Lf[e] - ref) <= 1e-12 * (1 + abs(ref))
        checked += 1
    assert checked >= 27


# ---------------------------------------------------------------------------
# closed forms
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("a", [1.0, 0.3])
def test_laplacian_of_polynomials(po, a):
    dims = (7, 6, 8)
    origin = (-1.3, 0.4, -2.1)
    x, y, z = site_grid(dims, origin, a)
    m = interior_mask(dims)
    tol = 1e-9 * (1 + 1 / a ** 2)
    for f, Lf in [
        (np.ones_like(x), 0.0 * x),
        (x, 0.0 * x),
        (x * y - 2 * z, 0.0 * x),
        (x * x, 2.0 + 0 * x),
        (x * x + y * y + z * z, 6.0 + 0 * x),
        # fourth-order terms separate the weights: the axial moment enters only x^4,
        # the diagonal one also x^2 y^2
        (x ** 4, 12 * x * x + 7 * a * a / 8),
        (x * x * y * y, 2 * (x * x + y * y) + 3 * a * a / 8),
    ]:
        got = po.laplacian(f, dims, a)
        np.testing.assert_allclose(got[m], Lf[m], rtol=1e-10, atol=tol)


def test_plane_wave_dispersion(po):
    """L cos(k.x) = -lambda(k) cos(k.x) at interior sites with
    lambda = (6/a^2)(1 - cx cy cz) + (1/(2a^2))(3 - sum cos(k_d a)), c_d = cos(k_d a/2)."""
    a = 0.45
    dims = (9, 8, 7)
    origin = (0.1, -0.2, 0.3)
    x, y, z = site_grid(dims, origin, a)
    m = interior_mask(dims)
    for k in [(0.3, 0.0, 0.0), (1.1, -0.7, 2.3), (3.0, 2.0, 1.0)]:
        f = np.cos(k[0] * x + k[1] * y + k[2] * z + 0.25)
        c = [math.cos(kd * a / 2) for kd in k]
        lam = 6 / a ** 2 * (1 - c[0] * c[1] * c[2]) + 0.5 / a ** 2 * (3 - sum(math.cos(kd * a) for kd in k))
        got = po.laplacian(f, dims, a)
        np.testing.assert_allclose(got[m], -lam * f[m], rtol=0, atol=1e-10 * (1 + lam))


def test_laplacian_symmetric_negative(po):
    rng = np.random.default_rng(3)
    dims = (5, 4, 6)
    f = rng.standard_normal(2 * 5 * 4 * 6)
    g = rng.standard_normal(2 * 5 * 4 * 6)
    Lf, Lg = po.laplacian(f, dims, 0.6), po.laplacian(g, dims, 0.6)
    assert abs(f @ Lg - Lf @ g) <= 1e-12 * abs(f @ Lg)
    assert f @ Lf < 0
    # Gershgorin bound used for the stability check: spectrum of -L in [0, 15/a^2]
    assert -(f @ Lf) / (f @ f) < 15 / 0.36


# ---------------------------------------------------------------------------
# integrators
# ---------------------------------------------------------------------------
def test_half_update_on_constant_field(po):
    """Away from the boundary L c = 0, so s = dt (V + g|c|^2) c and
    the imaginary-time update gives c - s: pins the potential and nonlinear terms."""
    dims, a, dt, g, ratio = (6, 6, 6), 0.5, 0.01, 2.5, 1.7
    origin = centered(dims, a)
    V = np_trap(dims, origin, a, ratio)
    c = 0.3 - 0.4j
    psi = np.full(2 * 6 ** 3, c).view(np.float64).copy()
    coef = [(-0.375 / a ** 2) * dt, (-0.125 / a ** 2) * dt, (3.75 / a ** 2) * dt, g * dt]
    out, _ = po.half_update(np.zeros_like(psi), psi, V * dt, dims, coef, imag=True)
    m = interior_mask(dims)
    got = out.view(np.complex128)[m]
    want = c - dt * (V[m] + g * abs(c) ** 2) * c
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-13)
    # real time: next += -i s
    nxt = np.zeros_like(psi)
    out, _ = po.half_update(nxt, psi, V * dt, dims, coef, imag=False)
    np.testing.assert_allclose(out.view(np.complex128)[m], -1j * dt * (V[m] + g * abs(c) ** 2) * c, rtol=0,
                               atol=1e-13)


def _ground_state(po, n, a, steps, dtau):
    dims = (n, n, n)
    origin = centered(dims, a)
    V = np_trap(dims, origin, a, 1.0)
    x, y, z = site_grid(dims, origin, a)
    psi = np.exp(-0.3 * (x * x + y * y + z * z)).astype(np.complex128).view(np.float64).copy()
    w = 0.5 * a ** 3
    psi /= math.sqrt(w * (psi @ psi))
    coef = [(-0.375 / a ** 2) * dtau, (-0.125 / a ** 2) * dtau, (3.75 / a ** 2) * dtau, 0.0]
    psi, N = po.imag_steps(psi, V * dtau, dims, coef, w, steps)
    obs = po.observables(psi, V, dims, a, 0.0)
    return psi, obs, dims, V


@pytest.mark.slow
def test_harmonic_ground_state_converges(po):
    """g = 0, V = r^2/2: E0 -> 3/2 at second order in a, K = P (virial)."""
    errs = []
    for n, a in [(16, 0.6), (24, 0.4)]:
        _, obs, _, _ = _ground_state(po, n, a, 1500, 0.012 if a == 0.6 else 0.01)
        N, K, P, _ = obs
        E = (K + P) / N
        errs.append(abs(E - 1.5))
        assert abs(N - 1) < 1e-12
        assert abs(K - P) < 0.05 * E
    assert errs[1] < 0.02
    ratio = errs[0] / errs[1]
    assert 1.6 < ratio < 3.2, (errs, ratio)  # (0.6/0.4)^2 = 2.25


def test_leapfrog_phase_of_stationary_state(po):
    """A converged ground state evolves as e^{-i mu t} under the leapfrog
    started with odd = even e^{i mu dt/2}."""
    n, a = 12, 0.6
    psi, obs, dims, V = _ground_state(po, n, a, 1500, 0.012)
    N, K, P, I = obs
    mu = (K + P + I) / N
    dt = 0.01
    coef = [(-0.375 / a ** 2) * dt, (-0.125 / a ** 2) * dt, (3.75 / a ** 2) * dt, 0.0]
    even = psi.copy()
    odd = (psi.view(np.complex128) * np.exp(0.5j * mu * dt)).view(np.float64).copy()
    T = 200
    e, _ = po.leapfrog(even, odd, V * dt, dims, coef, T)
    ov = np.vdot(psi.view(np.complex128), e.view(np.complex128)) * 0.5 * a ** 3
    assert abs(abs(ov) - 1) < 1e-4
    assert abs(ov / abs(ov) - np.exp(-1j * mu * T * dt)) < 2e-3


# ---------------------------------------------------------------------------
# product host helpers (no GPU call)
# ---------------------------------------------------------------------------
def test_product_host_helpers_match_numpy():
    import vaikkari_amd as va

    dims, a, ratio = (5, 3, 4), 0.37, 1.3
    origin = va.primal_centered_origin(dims, a)
    assert origin == tuple(-0.5 * a * (d - 0.5) for d in dims)
    x, y, z = site_grid(dims, origin, a)
    assert np.array_equal(va.primal_positions(dims, origin, a), np.stack([x, y, z], axis=1))
    assert np.array_equal(va.primal_trap_potential(dims, origin, a, ratio), np_trap(dims, origin, a, ratio))
    dt, g = 0.013, 2.2
    assert np.array_equal(va.primal_coefficients(a, dt, g),
                          np.array([(-0.375 / a ** 2) * dt, (-0.125 / a ** 2) * dt, (3.75 / a ** 2) * dt, g * dt]))
    # c0 = -(8 cD + 6 cA): the two neighbour-sum forms are the same operator
    c = va.primal_coefficients(a, 1.0, 0.0)
    assert abs(c[2] + 8 * c[0] + 6 * c[1]) < 1e-12 * c[2]


def test_product_rejects_unstable_and_bad_config():
    import vaikkari_amd as va
    from vaikkari_amd import _lib

    a = 0.25
    # dt (7.5/a^2 + max V) >= 2
    with pytest.raises(_lib.GpeError) as ei:
        va.PrimalSolver((8, 8, 8), a, 2.0 / (7.5 / a ** 2))
    assert ei.value.status == _lib.EUNSTABLE
    with pytest.raises(_lib.GpeError) as ei:
        va.PrimalSolver((0, 8, 8), a, 0.001)
    assert ei.value.status == _lib.EINVAL
    with pytest.raises(_lib.GpeError) as ei:
        va.PrimalSolver((8, 8, 8), -a, 0.001)
    assert ei.value.status == _lib.EINVAL
    with pytest.raises(_lib.GpeError) as ei:
        va.PrimalSolver((8, 8, 8), a, 0.001, tile=va.primal_tiles() + 1)
    assert ei.value.status == _lib.EINVAL
