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
