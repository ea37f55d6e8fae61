"""ctypes access to the parity oracle (oracle/liboracle.so) and to the
reference's own compiled host code (oracle/_ref/libvsref.so).

Test infrastructure only: the oracle is the checker, never the product.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libvsref.so")

_D, _P, _LL, _U32, _I = C.c_double, C.c_void_p, C.c_longlong, C.c_uint32, C.c_int
_DP = C.POINTER(C.c_double)


def _dp(a):
    return None if a is None else a.ctypes.data_as(_DP)


def load_oracle():
    if not os.path.exists(ORACLE_SO):
        raise RuntimeError(f"{ORACLE_SO} missing: run `make -C oracle`")
    o = C.CDLL(ORACLE_SO)
    sig = {
        "or_leapfrog": (_I, [_DP, _DP, _DP, _I, _I, _I, _D, _D, _D, _LL]),
        "or_leapfrog_timed": (_I, [_DP, _DP, _DP, _I, _I, _I, _D, _D, _D, _I, _I, _LL, _DP, C.POINTER(_I)]),
        "or_half_update": (_I, [_DP, _DP, _DP, _I, _I, _I, _D, _D, _D, _I]),
        "or_observables": (_I, [_DP, _DP, _I, _I, _I, _D, _D, _DP]),
        "or_overlap": (_D, [_DP, _DP, _LL]),
        "or_imag_steps": (_I, [_DP, _DP, _I, _I, _I, _D, _D, _D, _D, _LL, _DP]),
        "or_vs_load": (_P, [C.c_char_p]),
        "or_vs_free": (None, [_P]),
        "or_vs_info": (None, [_P, _DP]),
        "or_vs_function": (_D, [_P, _D, _D]),
        "or_vs_psi": (None, [_P, _D, _D, _D, _DP]),
        "or_vs_phase": (None, [_P, _D, _DP]),
        "or_vs_search_max": (_D, [_P]),
        "or_vs_search_max_r": (_D, [_P, _D]),
        "or_vs_search_max_z": (_D, [_P, _D]),
        "or_vs_curvature": (_D, [_P]),
        "or_mt_uints": (None, [_U32, _LL, C.POINTER(C.c_uint32)]),
        "or_mt_circles": (None, [_U32, _LL, _DP]),
        "or_ref_dims": (None, [_D, _DP, _DP, C.POINTER(C.c_uint32), _DP]),
        "or_node_positions": (None, [_DP, _D, _I, _I, _I, _DP]),
        "or_ref_setup": (_I, [_P, _D, _DP, _DP, _D, _D, _U32, _DP, _DP, _DP, _DP]),
        "or_positions": (None, [C.POINTER(C.c_int)]),
        "or_neighbours": (None, [C.POINTER(C.c_int)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(o, name)
        f.restype, f.argtypes = res, args
    return o


def load_ref():
    """The reference's VortexState.cpp, Random.cpp, Picture.cpp and Text.cpp, or None when not built."""
    if not os.path.exists(REF_SO):
        return None
    r = C.CDLL(REF_SO)
    sig = {
        "ref_vs_load": (_P, [C.c_char_p]),
        "ref_vs_free": (None, [_P]),
        "ref_vs_psi": (None, [_P, _D, _D, _D, _DP]),
        "ref_vs_function": (_D, [_P, _D, _D]),
        "ref_vs_phase": (None, [_P, _D, _DP]),
        "ref_vs_search_max": (_D, [_P]),
        "ref_vs_search_max_r": (_D, [_P, _D]),
        "ref_vs_search_max_z": (_D, [_P, _D]),
        "ref_vs_curvature": (_D, [_P]),
        "ref_vs_mu": (_D, [_P]),
        "ref_vs_g": (_D, [_P]),
        "ref_vs_kappa": (_D, [_P]),
        "ref_vs_create": (_P, []),
        "ref_vs_set_physics": (None, [_P, C.c_uint, _D]),
        "ref_vs_set_dimensions": (None, [_P, C.c_uint, _D, _D, C.c_uint, _D]),
        "ref_vs_set_range": (None, [_P, _D, _D, _D, _D, _D]),
        "ref_vs_initialize": (None, [_P, _P]),
        "ref_vs_iterate": (None, [_P, _D, C.c_uint, _D]),
        "ref_vs_error": (_D, [_P, _D]),
        "ref_vs_search_min_r": (_D, [_P, _D]),
        "ref_vs_integrate_z": (_D, [_P, _D]),
        "ref_vs_save": (C.c_int, [_P, C.c_char_p]),
        "ref_picture_save": (C.c_int, [C.c_uint, C.c_uint, _DP, C.c_char_p, C.c_int]),
        "ref_save_volume_map": (C.c_int, [C.c_char_p, C.POINTER(C.c_uint16), C.c_uint, C.c_uint, C.c_uint,
                                          _D, _D, _D]),
        "ref_random_uints": (None, [_U32, _LL, C.POINTER(C.c_uint32)]),
        "ref_random_circles": (None, [_U32, _LL, _DP]),
    }
    for name, (res, args) in sig.items():
        f = getattr(r, name)
        f.restype, f.argtypes = res, args
    return r


class Oracle:
    """Thin numpy wrappers over liboracle.so."""

    def __init__(self):
        self.o = load_oracle()

    def leapfrog(self, even, odd, pot_dt, dims, lapfac, lapfac0, g_dt, nsteps):
        e = np.array(even, dtype=np.float64, copy=True)
        d = np.array(odd, dtype=np.float64, copy=True)
        p = np.ascontiguousarray(pot_dt, dtype=np.float64)
        rc = self.o.or_leapfrog(_dp(e), _dp(d), _dp(p), *map(int, dims), lapfac, lapfac0, g_dt, int(nsteps))
        assert rc == 0
        return e, d

    def leapfrog_timed(self, even, odd, pot_dt, dims, lapfac, lapfac0, g_dt, nwarm, nruns, nsteps):
        """Wall seconds of nruns runs of nsteps leapfrog steps after nwarm untimed ones (fields
        padded once, outside the timing), and the OpenMP threads used."""
        secs = np.zeros(int(nruns))
        th = C.c_int()
        rc = self.o.or_leapfrog_timed(_dp(np.ascontiguousarray(even)), _dp(np.ascontiguousarray(odd)),
                                      _dp(np.ascontiguousarray(pot_dt)), *map(int, dims), lapfac, lapfac0, g_dt,
                                      int(nwarm), int(nruns), int(nsteps), _dp(secs), C.byref(th))
        assert rc == 0
        return secs, th.value

    def half_update(self, nxt, prev, pot_dt, dims, lapfac, lapfac0, g_dt, imag):
        n = np.array(nxt, dtype=np.float64, copy=True)
        p = np.ascontiguousarray(prev, dtype=np.float64)
        v = np.ascontiguousarray(pot_dt, dtype=np.float64)
        rc = self.o.or_half_update(_dp(n), _dp(p), _dp(v), *map(int, dims), lapfac, lapfac0, g_dt, int(imag))
        assert rc == 0
        return n

    def observables(self, psi, V, dims, block_scale, g):
        out = np.zeros(4)
        rc = self.o.or_observables(_dp(np.ascontiguousarray(psi)), _dp(np.ascontiguousarray(V)), *map(int, dims),
                                   block_scale, g, _dp(out))
        assert rc == 0
        return out

    def overlap(self, a, b):
        return self.o.or_overlap(_dp(np.ascontiguousarray(a)), _dp(np.ascontiguousarray(b)), a.size // 2)

    def imag_steps(self, psi, pot_dt, dims, lapfac, lapfac0, g_dt, block_scale, nsteps):
        x = np.array(psi, dtype=np.float64, copy=True)
        norms = np.zeros(int(nsteps))
        rc = self.o.or_imag_steps(_dp(x), _dp(np.ascontiguousarray(pot_dt)), *map(int, dims), lapfac, lapfac0,
                                  g_dt, block_scale, int(nsteps), _dp(norms))
        assert rc == 0
        return x, norms

    def stencil(self):
        out = (C.c_int * 192)()
        self.o.or_neighbours(out)
        return np.array(out, dtype=np.int32).reshape(12, 4, 4)

    def positions(self):
        out = (C.c_int * 36)()
        self.o.or_positions(out)
        return np.array(out, dtype=np.int32).reshape(12, 3)


MESH_SO = os.path.join(ROOT, "oracle", "_ref", "libmeshref.so")
_IP = C.POINTER(C.c_int)


def load_mesh_ref():
    """The reference's mesh library with oracle/mesh_shim.cpp, or None when not built."""
    if not os.path.exists(MESH_SO):
        return None
    m = C.CDLL(MESH_SO)
    m.ref_bcc_stencils.restype = _I
    m.ref_bcc_stencils.argtypes = [_D, _D, _I, _I, _DP, _IP, _DP]
    return m


def mesh_stencils(m, h, ext, max_nodes=20000, max_nb=32):
    """(positions [n,3], counts [n], neighbours [n, max_nb, 4] = (dx, dy, dz, weight))."""
    pos = np.zeros((max_nodes, 3))
    cnt = np.zeros(max_nodes, np.int32)
    nb = np.zeros((max_nodes, max_nb, 4))
    n = m.ref_bcc_stencils(float(h), float(ext), max_nodes, max_nb, _dp(pos), cnt.ctypes.data_as(_IP), _dp(nb))
    assert n > 0, f"mesh has {-n} nodes, more than {max_nodes}"
    return pos[:n], cnt[:n], nb[:n]


class PrimalOracle:
    """numpy wrappers over the primal 14-point functions of liboracle.so
    (fields [k][j][i][s] complex interleaved, dims = (nx, ny, nz))."""

    def __init__(self):
        o = C.CDLL(ORACLE_SO) if os.path.exists(ORACLE_SO) else load_oracle()
        sig = {
            "or_primal_half_update": (_D, [_DP, _DP, _DP, _IP, _DP, _I, _D]),
            "or_primal_leapfrog": (None, [_DP, _DP, _DP, _IP, _DP, _LL]),
            "or_primal_imag_steps": (_D, [_DP, _DP, _DP, _IP, _DP, _D, _LL]),
            "or_primal_observables": (None, [_DP, _DP, _IP, _D, _D, _DP]),
            "or_primal_laplacian": (None, [_DP, _DP, _IP, _D]),
        }
        for name, (res, args) in sig.items():
            f = getattr(o, name)
            f.restype, f.argtypes = res, args
        self.o = o

    @staticmethod
    def _dims(dims):
        return (C.c_int * 3)(*[int(d) for d in dims])

    @staticmethod
    def _coef(coef):
        return np.ascontiguousarray(coef, np.float64)

    def half_update(self, nxt, prev, pot_dt, dims, coef, imag=False, scale=1.0):
        n = np.array(nxt, np.float64, copy=True)
        p = np.ascontiguousarray(prev, np.float64)
        v = np.ascontiguousarray(pot_dt, np.float64)
        c = self._coef(coef)
        s = self.o.or_primal_half_update(_dp(n), _dp(p), _dp(v), self._dims(dims), _dp(c), int(imag), float(scale))
        return n, s

    def leapfrog(self, even, odd, pot_dt, dims, coef, nsteps):
        e = np.array(even, np.float64, copy=True)
        d = np.array(odd, np.float64, copy=True)
        v = np.ascontiguousarray(pot_dt, np.float64)
        c = self._coef(coef)
        self.o.or_primal_leapfrog(_dp(e), _dp(d), _dp(v), self._dims(dims), _dp(c), int(nsteps))
        return e, d

    def imag_steps(self, psi, pot_dt, dims, coef, w, nsteps):
        p = np.array(psi, np.float64, copy=True)
        work = np.zeros_like(p)
        v = np.ascontiguousarray(pot_dt, np.float64)
        c = self._coef(coef)
        N = self.o.or_primal_imag_steps(_dp(p), _dp(work), _dp(v), self._dims(dims), _dp(c), float(w), int(nsteps))
        return p, N

    def observables(self, psi, V, dims, a, g):
        out = np.zeros(4)
        p = np.ascontiguousarray(psi, np.float64)
        v = np.ascontiguousarray(V, np.float64)
        self.o.or_primal_observables(_dp(p), _dp(v), self._dims(dims), float(a), float(g), _dp(out))
        return out

    def laplacian(self, f, dims, a):
        f = np.ascontiguousarray(f, np.float64)
        out = np.zeros_like(f)
        self.o.or_primal_laplacian(_dp(f), _dp(out), self._dims(dims), float(a))
        return out
