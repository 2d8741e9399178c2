"""ctypes wrapper over the CPU oracle (oracle/liboracle.so) and the reference
build (oracle/_ref/libpvref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the cpu_baseline leg of ``bench.py`` as the checker.  The product package never
imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_I = np.int32
_P_INT = C.POINTER(C.c_int)
_P_DBL = C.POINTER(C.c_double)


def _ip(a):
    return a.ctypes.data_as(_P_INT)


def _dp(a):
    return a.ctypes.data_as(_P_DBL)


_lib = None
_ref = {}


def lib(fast: bool = False):
    """The oracle restatement (``fast`` = built with the reference's flags)."""
    global _lib
    name = "liboracle_fast.so" if fast else "liboracle.so"
    if fast:
        return _load(os.path.join(HERE, name))
    if _lib is None:
        _lib = _load(os.path.join(HERE, name))
    return _lib


def _load(path):
    if not os.path.exists(path):
        raise RuntimeError(f"oracle library missing: {path} (run __graft_entry__.build())")
    L = C.CDLL(path)
    L.orc_module_current.argtypes = [C.c_double] * 6 + [_P_DBL] * 3
    return L


def ref(fast: bool = False):
    """The reference's own CPU units (None when oracle/_ref was not built)."""
    name = "libpvref_fast.so" if fast else "libpvref.so"
    if name not in _ref:
        path = os.path.join(HERE, "_ref", name)
        _ref[name] = C.CDLL(path) if os.path.exists(path) else None
    return _ref[name]


def _take(L, ptr, n, free):
    out = np.ctypeslib.as_array(ptr, shape=(n,)).copy() if n > 0 else np.zeros(0, _I)
    free(ptr)
    return out.astype(_I)


def _libc_free():
    libc = C.CDLL(None)
    libc.free.argtypes = [C.c_void_p]
    return lambda p: libc.free(C.cast(p, C.c_void_p))


_free = _libc_free()


# ------------------------------------------------------------------ symbolic
def symbolic_fill(n, rp, ci):
    rp = np.ascontiguousarray(rp, _I)
    ci = np.ascontiguousarray(ci, _I)
    orp, oci = _P_INT(), _P_INT()
    nnz = lib().orc_symbolic_fill(n, _ip(rp), _ip(ci), C.byref(orp), C.byref(oci))
    return _take(None, orp, n + 1, _free), _take(None, oci, nnz, _free)


def inclusive_scan(a):
    a = np.ascontiguousarray(a, _I).copy()
    lib().orc_inclusive_scan(_ip(a), len(a))
    return a


def deps(kind, n, rp, ci, use_ref=False):
    rp = np.ascontiguousarray(rp, _I)
    ci = np.ascontiguousarray(ci, _I)
    dp, di = _P_INT(), _P_INT()
    L = ref() if use_ref else lib()
    fn = L.ref_deps if use_ref else L.orc_deps
    tot = fn(kind, n, _ip(rp), _ip(ci), C.byref(dp), C.byref(di))
    return _take(None, dp, n + 1, _free), _take(None, di, tot, _free)


@dataclass
class Levels:
    start_levels: np.ndarray
    start_chains: np.ndarray
    chains: np.ndarray

    @property
    def nlev(self):
        return len(self.start_levels) - 1


def levels(n, dptr, didx, use_ref=False):
    dptr = np.ascontiguousarray(dptr, _I)
    didx = np.ascontiguousarray(didx, _I)
    sl, sc, cc = _P_INT(), _P_INT(), _P_INT()
    nch = C.c_int()
    L = ref() if use_ref else lib()
    fn = L.ref_levels if use_ref else L.orc_levels
    nlev = fn(n, _ip(dptr), _ip(didx), C.byref(sl), C.byref(sc), C.byref(cc), C.byref(nch))
    if nlev < 0:
        raise ValueError("dependency graph has a cycle")
    return Levels(_take(None, sl, nlev + 1, _free), _take(None, sc, nch.value + 1, _free),
                  _take(None, cc, n, _free))


def max_per_level(lv: Levels, rp):
    rp = np.ascontiguousarray(rp, _I)
    out = np.zeros(lv.nlev, _I)
    lib().orc_max_per_level(lv.nlev, _ip(lv.start_levels), _ip(lv.start_chains),
                            _ip(lv.chains), _ip(rp), _ip(out))
    return out


def ref_max_per_level(n, rp, ci, dptr, didx):
    rp, ci = np.ascontiguousarray(rp, _I), np.ascontiguousarray(ci, _I)
    dptr, didx = np.ascontiguousarray(dptr, _I), np.ascontiguousarray(didx, _I)
    out = np.zeros(n + 1, _I)
    m = ref().ref_max_per_level(n, _ip(rp), _ip(ci), _ip(dptr), _ip(didx), _ip(out))
    return out[:m]


@dataclass
class CsrCsc:
    n: int
    rp: np.ndarray
    ci: np.ndarray
    cp: np.ndarray
    ri: np.ndarray
    csr2csc: np.ndarray
    csc2csr: np.ndarray
    csr_diag: np.ndarray
    csc_diag: np.ndarray


def csr_csc(n, rp, ci) -> CsrCsc:
    rp = np.ascontiguousarray(rp, _I)
    ci = np.ascontiguousarray(ci, _I)
    nnz = int(rp[n])
    cp = np.zeros(n + 1, _I)
    ri = np.zeros(nnz, _I)
    a2c = np.zeros(nnz, _I)
    c2a = np.zeros(nnz, _I)
    rd = np.zeros(n, _I)
    cd = np.zeros(n, _I)
    lib().orc_csr_csc(n, _ip(rp), _ip(ci), _ip(cp), _ip(ri), _ip(a2c), _ip(c2a), _ip(rd), _ip(cd))
    return CsrCsc(n, rp, ci, cp, ri, a2c, c2a, rd, cd)


# ------------------------------------------------------------------ numeric
def glu_csc_sequential(s: CsrCsc, vals, use_ref=False, fast=False):
    """In-place GLU of CSC values; returns (status, bad_col)."""
    assert vals.dtype == np.float64 and vals.flags.c_contiguous
    if use_ref:
        ref(fast).ref_glu_csc_sequential(s.n, _ip(s.rp), _ip(s.ci), _ip(s.csr2csc), _ip(s.cp),
                                         _ip(s.ri), _dp(vals), _ip(s.csr_diag), _ip(s.csc_diag))
        return 0, -1
    bad = C.c_int(-1)
    st = lib(fast).orc_glu_csc_sequential(s.n, _ip(s.rp), _ip(s.ci), _ip(s.csr2csc), _ip(s.cp),
                                          _ip(s.ri), _dp(vals), _ip(s.csr_diag), _ip(s.csc_diag),
                                          C.byref(bad))
    return st, bad.value


def substitute_lu_csc(s: CsrCsc, lu, b, use_ref=False, fast=False):
    """In-place forward/backward solve; b becomes x."""
    assert b.dtype == np.float64 and b.flags.c_contiguous
    if use_ref:
        ref(fast).ref_substitute_lu_csc(s.n, _ip(s.cp), _ip(s.ri), _dp(lu), _ip(s.csc_diag), _dp(b))
    else:
        lib(fast).orc_substitute_lu_csc(s.n, _ip(s.cp), _ip(s.ri), _dp(lu), _ip(s.csc_diag), _dp(b))


def solve_transpose_csc(s: CsrCsc, lu, g):
    assert g.dtype == np.float64 and g.flags.c_contiguous
    lib().orc_solve_transpose_csc(s.n, _ip(s.cp), _ip(s.ri), _dp(lu), _ip(s.csc_diag), _dp(g))


def newton_update(v_old, v_new):
    """Returns (status, v_updated, diff_sum(long double->float), max_abs_delta)."""
    v = np.ascontiguousarray(v_old, np.float64).copy()
    vn = np.ascontiguousarray(v_new, np.float64)
    diff = C.c_longdouble()
    mx = C.c_double()
    st = lib().orc_newton_update(len(v), _dp(v), _dp(vn), C.byref(diff), C.byref(mx))
    return st, v, float(diff.value), mx.value


def apply_permutation(vec, perm):
    v = np.ascontiguousarray(vec, np.float64).copy()
    p = np.ascontiguousarray(perm, _I)
    lib().orc_apply_permutation(len(v), _dp(v), _ip(p))
    return v


def invert_permutation(perm):
    p = np.ascontiguousarray(perm, _I)
    out = np.zeros_like(p)
    lib().orc_invert_permutation(len(p), _ip(p), _ip(out))
    return out


def permute_sym(n, rp, ci, v, P):
    rp, ci, P = (np.ascontiguousarray(x, _I) for x in (rp, ci, P))
    v = np.ascontiguousarray(v, np.float64)
    orp, oci, ov = np.zeros_like(rp), np.zeros_like(ci), np.zeros_like(v)
    lib().orc_permute_sym(n, _ip(rp), _ip(ci), _dp(v), _ip(P), _ip(orp), _ip(oci), _dp(ov))
    return orp, oci, ov


def permute_rows(n, rp, ci, v, p):
    rp, ci, p = (np.ascontiguousarray(x, _I) for x in (rp, ci, p))
    v = np.ascontiguousarray(v, np.float64)
    orp, oci, ov = np.zeros_like(rp), np.zeros_like(ci), np.zeros_like(v)
    lib().orc_permute_rows(n, _ip(rp), _ip(ci), _dp(v), _ip(p), _ip(orp), _ip(oci), _dp(ov))
    return orp, oci, ov


def remove_zeros_diagonal(n, rp, ci, v=None):
    """Returns (status, rp, ci, v, perm) after the row-only swaps of 004b."""
    rp = np.ascontiguousarray(rp, _I).copy()
    ci = np.ascontiguousarray(ci, _I).copy()
    v = np.zeros(len(ci)) if v is None else np.ascontiguousarray(v, np.float64).copy()
    perm = np.arange(n, dtype=_I)
    st = lib().orc_remove_zeros_diagonal(n, _ip(rp), _ip(ci), _dp(v), _ip(perm))
    return st, rp, ci, v, perm


def amd_order(n, rp, ci):
    """The reference's AMD 1.1 (oracle/_ref build).  Returns P or raises."""
    R = ref()
    if R is None:
        raise RuntimeError("oracle/_ref not built")
    rp, ci = np.ascontiguousarray(rp, _I), np.ascontiguousarray(ci, _I)
    P = np.zeros(n, _I)
    st = R.ref_amd_order(n, _ip(rp), _ip(ci), _ip(P))
    if st != 0:
        raise RuntimeError(f"amd_order status {st}")
    return P


def module_current(V, Iph, I0, a, Rs, Rsh):
    I, dV, dI = C.c_double(), C.c_double(), C.c_double()
    lib().orc_module_current(V, Iph, I0, a, Rs, Rsh, C.byref(I), C.byref(dV), C.byref(dI))
    return I.value, dV.value, dI.value
