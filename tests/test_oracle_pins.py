"""Pins the CPU oracle before anything is checked against it.

Each check is chosen so that a plausible mistake in the restatement (a wrong
table entry, sign, index, dropped term or transposed operand) fails:
  * golden fixtures from the reference tree (geometry table, the reference's
    own compiled VortexState/Random outputs, the known answer 691);
  * closed forms (L*1 = 0, L*x = 0, L*r^2 = 6, node measure, Gershgorin);
  * invariants (Hermitian H, the exact leapfrog invariant Re<even,odd>);
  * brute force against a dense-matrix Hamiltonian on tiny grids.
"""
import json
import math
import os

import numpy as np
import pytest

from oracle_lib import Oracle, load_ref

HERE = os.path.dirname(os.path.abspath(__file__))
GEO = json.load(open(os.path.join(HERE, "golden", "bcc_geometry_ref.json")))
VSREF = json.load(open(os.path.join(HERE, "golden", "vortex_state_ref.json")))


@pytest.fixture(scope="module")
def orc():
    return Oracle()


# ---------------------------------------------------------------- geometry
def test_oracle_table_matches_reference_text(orc):
    assert orc.positions().tolist() == GEO["positions"]
    assert orc.stencil().reshape(48, 4).tolist() == GEO["neighbours"]
    assert GEO["hodge"] == 1.5


def test_table_is_the_distance_b_relation(orc):
    q = np.array(GEO["positions"])
    nb = orc.stencil()
    for l in range(12):
        found = set()
        for d in np.ndindex(3, 3, 3):
            off = np.array(d) - 1
            for m in range(12):
                v = q[m] + 8 * off - q[l]
                if int(v @ v) == 8:
                    found.add((int(off[0]), int(off[1]), int(off[2]), m))
        assert found == {tuple(int(x) for x in e) for e in nb[l]}
        # symmetric: l sees m at d  <=>  m sees l at -d
        for dx, dy, dz, m in nb[l]:
            assert (-dx, -dy, -dz, l) in {tuple(int(x) for x in e) for e in nb[m]}


# ------------------------------------------------------------- MT19937
def test_mt19937_known_answers(orc):
    import ctypes as C
    out = (C.c_uint32 * 10000)()
    orc.o.or_mt_uints(5489, 10000, out)
    assert out[9999] == 4123659995  # C++ standard [rand.predef]: 10000th mt19937 output, default seed
    assert list(out[:1000]) == VSREF["uints_seed5489_first1000"]
    circ = np.zeros(2 * 256)
    orc.o.or_mt_circles(54363, 256, circ.ctypes.data_as(C.POINTER(C.c_double)))
    np.testing.assert_array_equal(circ, np.array(VSREF["circles_seed54363"]))


# --------------------------------------------------------- VortexState
def test_vortex_state_against_reference_outputs(orc, state_dat):
    import ctypes as C
    h = orc.o.or_vs_load(state_dat.encode())
    assert h
    info = np.zeros(8)
    orc.o.or_vs_info(h, info.ctypes.data_as(C.POINTER(C.c_double)))
    assert info[0] == VSREF["kappa"] and info[1] == VSREF["g"] and info[7] == VSREF["mu"]
    out = np.zeros(2)
    for p, ref in zip(VSREF["psi_points"], VSREF["psi"]):
        orc.o.or_vs_psi(h, p[0], p[1], p[2], out.ctypes.data_as(C.POINTER(C.c_double)))
        assert out.tolist() == ref
    for (r, z), ref in zip(VSREF["function_points"], VSREF["function"]):
        assert orc.o.or_vs_function(h, r, z) == ref
    for t, ref in zip(VSREF["phase_times"], VSREF["phase"]):
        orc.o.or_vs_phase(h, t, out.ctypes.data_as(C.POINTER(C.c_double)))
        assert out.tolist() == ref
    assert orc.o.or_vs_search_max(h) == VSREF["search_max"]
    assert orc.o.or_vs_search_max_r(h, VSREF["eps"]) == VSREF["search_max_r"]
    assert orc.o.or_vs_search_max_z(h, VSREF["eps"]) == VSREF["search_max_z"]
    assert orc.o.or_vs_curvature(h) == VSREF["curvature"]
    orc.o.or_vs_free(h)


@pytest.mark.slow
def test_reference_setup_known_answer_691(orc, state_dat):
    """state.dat with block_scale = 2pi/(20 sqrt(curvature)) (GPUGrossPitaevskii_BCC.cu:539) gives
    steps_per_iteration = 691, the value hard-coded in GPUGrossPitaevskii_BCC_multi.cu:473."""
    import ctypes as C
    h = orc.o.or_vs_load(state_dat.encode())
    curv = orc.o.or_vs_curvature(h)
    bs = 6.2831853071795864770 / (20.0 * math.sqrt(curv))
    eps = 1e-5 * orc.o.or_vs_search_max(h)
    mr, mz = orc.o.or_vs_search_max_r(h, eps), orc.o.or_vs_search_max_z(h, eps)
    minp = (C.c_double * 3)(-mr, -mr, -mz)
    maxp = (C.c_double * 3)(mr, mr, mz)
    dims = (C.c_uint32 * 3)()
    p0 = (C.c_double * 3)()
    orc.o.or_ref_dims(bs, minp, maxp, dims, p0)
    assert tuple(dims) == (87, 87, 302)
    n = 87 * 87 * 302 * 12
    ev, od, pot, sc = np.zeros(2 * n), np.zeros(2 * n), np.zeros(n), np.zeros(6)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    assert orc.o.or_ref_setup(h, bs, minp, maxp, 1.0, 0.1, 54363, dp(ev), dp(od), dp(pot), dp(sc)) == 0
    assert int(sc[0]) == GEO["state_dat_steps_per_iteration"] == 691
    # the noise-free field is the normalised VortexState: discrete norm ~ 1
    w = math.sqrt(32 / 9) * bs ** 3
    assert abs(w * np.sum(ev ** 2) - 1.0) < 1e-3
    orc.o.or_vs_free(h)


# ------------------------------------------------- operator closed forms
def _grid(dims, b):
    import ctypes as C
    orc = Oracle()
    a = math.sqrt(8) * b
    p0 = np.array([-0.5 * a * d for d in dims])
    pos = np.zeros(3 * 12 * int(np.prod(dims)))
    orc.o.or_node_positions(p0.ctypes.data_as(C.POINTER(C.c_double)), b, *dims,
                            pos.ctypes.data_as(C.POINTER(C.c_double)))
    return pos.reshape(-1, 3)


def _laplacian(orc, f, dims, b):
    """(L f) from one imaginary-time half-update with pot = g = 0:
    next = prev - s, s = -1/2 (1.5/b^2) sum_f (f_n - f)  =>  L f = 2 (next - prev)."""
    lapfac = -0.5 * 1.5 / (b * b)
    psi = np.zeros(2 * f.size)
    psi[0::2] = f
    out = orc.half_update(np.zeros_like(psi), psi, np.zeros(f.size), dims, lapfac, 4 * -lapfac, 0.0, 1)
    return 2.0 * (out[0::2] - f)


def _interior(dims):
    X, Y, Z = dims
    mask = np.zeros((Z, Y, X, 12), dtype=bool)
    mask[1:-1, 1:-1, 1:-1, :] = True
    return mask.reshape(-1)


@pytest.mark.parametrize("b", [0.1, 0.0883883476483184])
def test_laplacian_closed_forms(orc, b):
    dims = (6, 5, 7)
    pos = _grid(dims, b)
    inner = _interior(dims)
    L1 = _laplacian(orc, np.ones(len(pos)), dims, b)
    Lx = _laplacian(orc, pos[:, 0] + 2 * pos[:, 1] - 3 * pos[:, 2], dims, b)
    Lr2 = _laplacian(orc, np.sum(pos ** 2, axis=1), dims, b)
    assert np.max(np.abs(L1[inner])) < 1e-12
    assert np.max(np.abs(Lx[inner])) < 1e-9
    np.testing.assert_allclose(Lr2[inner], 6.0, rtol=0, atol=1e-9)
    # L x^2 along one axis averages to the continuum value 2 over the 12 classes
    Lxx = _laplacian(orc, pos[:, 0] ** 2, dims, b).reshape(-1, 12)[inner.reshape(-1, 12)[:, 0]]
    np.testing.assert_allclose(Lxx.mean(axis=1), 2.0, atol=1e-9)


def _dense_H(orc, dims, b, V, g_psi2):
    """Dense H = -1/2 L + diag(V + g|psi|^2) assembled column by column from the oracle."""
    n = 12 * int(np.prod(dims))
    lapfac = -0.5 * 1.5 / (b * b)
    H = np.zeros((n, n))
    pot = V + g_psi2
    for c in range(n):
        e = np.zeros(2 * n)
        e[2 * c] = 1.0
        out = orc.half_update(np.zeros(2 * n), e, pot, dims, lapfac, 4 * -lapfac, 0.0, 1)
        H[:, c] = (e - out)[0::2]  # s = H e
    return H


def test_hamiltonian_hermitian_and_bounded(orc):
    dims, b = (2, 3, 2), 0.2
    n = 12 * 12
    rng = np.random.default_rng(1)
    V = rng.uniform(0, 1, n)
    H = _dense_H(orc, dims, b, V, np.zeros(n))
    np.testing.assert_allclose(H, H.T, atol=1e-12)
    ev = np.linalg.eigvalsh(H)
    assert ev.min() > 0 and ev.max() <= 6.0 / b ** 2 + V.max() + 1e-9  # Gershgorin bound
    # diagonal = 4 * 0.75/b^2 + V, off-diagonal = -0.75/b^2 on each of the 4 faces
    np.testing.assert_allclose(np.diag(H), 3.0 / b ** 2 + V, rtol=1e-14)
    off = H - np.diag(np.diag(H))
    assert set(np.round(off[off != 0] * b ** 2, 12)) == {-0.75}
    assert np.all(np.count_nonzero(off, axis=1) <= 4)


def test_leapfrog_against_dense_matrix(orc):
    """One odd and one even update against numpy with the dense H(psi_mid), nonlinearity included."""
    dims, b, dt, g = (2, 2, 3, ), 0.25, 1e-3, 300.0
    dims = (2, 2, 3)
    n = 12 * 12
    rng = np.random.default_rng(7)
    even = rng.normal(size=2 * n) * 0.1
    odd = rng.normal(size=2 * n) * 0.1
    V = rng.uniform(0, 3, n)
    lf, lf0, gdt = -0.75 / b ** 2 * dt, 3.0 / b ** 2 * dt, g * dt
    e2, o2 = orc.leapfrog(even, odd, V * dt, dims, lf, lf0, gdt, 1)
    ce, co = even.view(np.complex128), odd.view(np.complex128)
    H0 = _dense_H(orc, dims, b, V, np.zeros(n))
    co_new = co - 1j * dt * ((H0 + np.diag(g * np.abs(ce) ** 2)) @ ce)
    ce_new = ce - 1j * dt * ((H0 + np.diag(g * np.abs(co_new) ** 2)) @ co_new)
    np.testing.assert_allclose(o2.view(np.complex128), co_new, rtol=0, atol=1e-13)
    np.testing.assert_allclose(e2.view(np.complex128), ce_new, rtol=0, atol=1e-13)


def test_leapfrog_invariant_conserved(orc):
    """Re sum conj(even)*odd is exactly invariant under both half-updates (H(psi_mid) Hermitian)."""
    dims, b, dt, g = (5, 4, 6), 0.15, 2e-3, 800.0
    n = 12 * 5 * 4 * 6
    rng = np.random.default_rng(3)
    even = rng.normal(size=2 * n) * 0.05
    odd = even + rng.normal(size=2 * n) * 1e-3
    pos = _grid(dims, b)
    V = 0.5 * np.sum(pos ** 2, axis=1)
    q0 = orc.overlap(even, odd)
    e, o = orc.leapfrog(even, odd, V * dt, dims, -0.75 / b ** 2 * dt, 3.0 / b ** 2 * dt, g * dt, 200)
    q1 = orc.overlap(e, o)
    assert abs(q1 - q0) < 1e-12 * abs(q0)
    assert np.linalg.norm(e - even) > 1e-3 * np.linalg.norm(even)  # it did evolve


def test_observables_against_dense_matrix(orc):
    dims, b, g = (3, 2, 2), 0.3, 50.0
    n = 12 * 12
    rng = np.random.default_rng(11)
    psi = rng.normal(size=2 * n)
    V = rng.uniform(0, 2, n)
    N, K, P, I = orc.observables(psi, V, dims, b, g)
    w = math.sqrt(32 / 9) * b ** 3
    c = psi.view(np.complex128)
    H0 = _dense_H(orc, dims, b, np.zeros(n), np.zeros(n))  # = -1/2 L
    assert N == pytest.approx(w * np.sum(np.abs(c) ** 2), rel=1e-14)
    assert K == pytest.approx(w * np.real(np.vdot(c, H0 @ c)), rel=1e-13)
    assert P == pytest.approx(w * np.sum(V * np.abs(c) ** 2), rel=1e-14)
    assert I == pytest.approx(w * g * np.sum(np.abs(c) ** 4), rel=1e-14)
    # K as the dual-edge sum (a/2) sum |psi_j - psi_i|^2, Dirichlet edges to the halo counted once
    assert K > 0


def test_imaginary_time_harmonic_ground_state(orc):
    """g = 0, isotropic trap: E -> 3/2 (up to O(a^2)), norm 1 every step, E decreasing."""
    dims, b = (10, 10, 10), 0.25 / math.sqrt(8) * 2.4  # cube edge 0.6, box +-3
    pos = _grid(dims, b)
    V = 0.5 * np.sum(pos ** 2, axis=1)
    psi = np.zeros(2 * len(pos))
    psi[0::2] = np.exp(-np.sum((pos - 0.3) ** 2, axis=1) / 1.5)
    dtau = 0.4 / (3.0 / b ** 2 + V.max())
    lf, lf0 = -0.75 / b ** 2 * dtau, 3.0 / b ** 2 * dtau
    Es = []
    for _ in range(6):
        psi, norms = orc.imag_steps(psi, V * dtau, dims, lf, lf0, 0.0, b, 150)
        N, K, P, I = orc.observables(psi, V, dims, b, 0.0)
        assert N == pytest.approx(1.0, abs=1e-13)
        Es.append(K + P)
    assert all(e2 <= e1 + 1e-12 for e1, e2 in zip(Es, Es[1:]))
    assert Es[-1] == pytest.approx(1.5, abs=0.08)
