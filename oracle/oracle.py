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


def ref(fast=False):
    """The reference's own CPU units (None when oracle/_ref was not built).
    fast: False = the parity build (-O3 -ffp-contract=off), True = the
    reference's host flags with --fast-math, "o3" = the same without it."""
    name = {False: "libpvref.so", True: "libpvref_fast.so", "o3": "libpvref_o3.so"}[fast]
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


# ------------------------------------------------------------ PV Newton
LADDER = (100, 10, 1, 0.9, 0.8, 0.5, 0.4, 0.3, 0.2, 0.1, 0.05, -1)  # 000:101-102


class PVNewton:
    """CPU restatement of the reference's Newton loop for a config plant
    (PV/000_PVPP_simulator.cpp:163-235): companion stamping as
    create_matrix_and_vector (PV/002b_sistema_electrico_base_matriz.cpp:141-203),
    symmetric + row permutation, GLU_CSC_CPU_sequential, the CSC substitution,
    un-permutation and the 011 damped update.  Modules use the implicit
    single-diode solve (orc_module_current), not the library's Lambert W.

    form="direct" is the reference's iteration: the companion system
    A(v) x = b(v) is solved for the node voltages themselves.
    form="correction" takes the same Newton step written as a correction,
    A(v) d = b(v) - A(v) x_prev, x = x_prev + d, with the right-hand side
    evaluated branch by branch (the KCL residual of the linearized circuit at
    x_prev: -sum of each node's branch currents g (x_a - x_b), G_N (x - 600),
    f(v) + g (V(x_prev) - v) [+ g_virtual V(x_prev)]).  In exact arithmetic
    both give the same iterates (the first from x_prev = 0 is the direct
    solve); in FP64 the correction form's rounding scales with |d| instead of
    |x|, which lowers the last-stage noise floor of shaded strings from
    ~1e-10 V to ~1e-13 V."""

    RS, I0, A = 0.3, 1e-10, 1.5
    GN, VN = 1e4, 600.0

    def __init__(self, n, pos, neg, rsh, cab_a, cab_b, cab_g, buses, P, Prow, form="direct"):
        if form not in ("direct", "correction"):
            raise ValueError(form)
        self.form = form
        self.n = n
        self.pos, self.neg = np.asarray(pos, _I), np.asarray(neg, _I)
        self.rsh = np.ascontiguousarray(rsh, np.float64)
        self.cab_a, self.cab_b = np.asarray(cab_a, _I), np.asarray(cab_b, _I)
        self.cab_g = np.asarray(cab_g, np.float64)
        self.buses = np.asarray(buses, _I)
        self.P = np.asarray(P, _I)
        Pinv = invert_permutation(self.P)
        self.colmap = Pinv
        self.rowmap = np.asarray(Prow, _I)[Pinv]
        entries = {}
        base = {}

        def add(d, i, j, v):
            d[(i, j)] = d.get((i, j), 0.0) + v

        for a, b, g in zip(cab_a, cab_b, cab_g):
            for (i, j, s) in ((a, a, 1), (b, b, 1), (a, b, -1), (b, a, -1)):
                add(base, int(i), int(j), s * g)
        self.bbase = np.zeros(n)
        for bus in buses:
            add(base, int(bus), int(bus), self.GN)
            self.bbase[bus] += self.GN * self.VN
        for e in range(len(pos)):
            p, m = int(pos[e]), int(neg[e])
            for (i, j) in ((p, p), (m, m), (p, m), (m, p)):
                if i >= 0 and j >= 0:
                    entries[(i, j)] = 1
        for k in base:
            entries[k] = 1
        rows = [set() for _ in range(n)]
        for (i, j) in entries:
            rows[int(self.rowmap[i])].add(int(self.colmap[j]))
        rp = [0]
        ci = []
        for r in rows:
            ci.extend(sorted(r))
            rp.append(len(ci))
        frp, fci = symbolic_fill(n, np.array(rp, _I), np.array(ci, _I))
        self.s = csr_csc(n, frp, fci)
        pos_in_row = {}
        for i in range(n):
            for p_ in range(frp[i], frp[i + 1]):
                pos_in_row[(i, int(fci[p_]))] = int(self.s.csr2csc[p_])

        def slot(i, j):
            return pos_in_row[(int(self.rowmap[i]), int(self.colmap[j]))]

        self.abase = np.zeros(len(fci))
        for (i, j), v in base.items():
            self.abase[slot(i, j)] += v
        # slots of each cable's four stamps (a,a), (b,b), (a,b), (b,a)
        self.cab_slots = np.array([[slot(a, a), slot(b, b), slot(a, b), slot(b, a)]
                                   for a, b in zip(self.cab_a, self.cab_b)], dtype=np.int64)
        sl, sg, el = [], [], []
        for e in range(len(pos)):
            p, m = int(pos[e]), int(neg[e])
            for (i, j, s) in ((p, p, 1), (m, m, 1), (p, m, -1), (m, p, -1)):
                if i >= 0 and j >= 0:
                    sl.append(slot(i, j))
                    sg.append(s)
                    el.append(e)
        self.st_slot = np.array(sl)
        self.st_sign = np.array(sg, float)
        self.st_elem = np.array(el)

    def module(self, V, iph):
        """Rows I, dI/dV, dI/dIph of every module at voltages V."""
        V = np.ascontiguousarray(V, np.float64)
        iph = np.ascontiguousarray(iph, np.float64)
        m = len(V)
        out = np.zeros((3, m))
        lib().orc_module_current_many(m, _dp(V), _dp(iph), C.c_double(self.I0),
                                      C.c_double(self.A), C.c_double(self.RS), _dp(self.rsh),
                                      _dp(out[0]), _dp(out[1]), _dp(out[2]))
        return out

    def element_voltages(self, x):
        """V(pos) - V(neg) per module (voltage_drop_dc, PV/002_base_electric_system.cpp:21-33)."""
        return (np.where(self.pos >= 0, x[np.maximum(self.pos, 0)], 0.0) -
                np.where(self.neg >= 0, x[np.maximum(self.neg, 0)], 0.0))

    def _values(self, g, gv):
        vals = self.abase.copy()
        np.add.at(vals, self.st_slot, self.st_sign * g[self.st_elem])
        if gv > 0:
            np.add.at(vals, self.st_slot, self.st_sign * gv)
        return vals

    def _permute_rows(self, b):
        bp = np.zeros(self.n)
        bp[self.rowmap] = b
        return bp

    def assemble(self, v, iph, gv):
        """Permuted CSC values and right-hand side of the companion system."""
        I, dIdV, _ = self.module(v, iph)
        f, g = -I, -dIdV
        ieq = f - g * v
        b = self.bbase.copy()
        pm = self.pos >= 0
        np.add.at(b, self.pos[pm], -ieq[pm])
        nm = self.neg >= 0
        np.add.at(b, self.neg[nm], ieq[nm])
        return self._values(g, gv), self._permute_rows(b)

    def residual(self, x, v, iph, gv, fg=None):
        """b(v) - A(v) x of the companion system, unpermuted, evaluated branch
        by branch: minus the sum of the branch currents leaving each node
        (cables g (x_a - x_b); Norton buses G_N (x - 600); modules
        f(v) + g (V(x) - v), plus g_virtual V(x) on homotopy stages)."""
        if fg is None:
            I, dIdV, _ = self.module(v, iph)
            fg = (-I, -dIdV)
        f, g = fg
        r = np.zeros(self.n)
        ic = self.cab_g * (x[self.cab_a] - x[self.cab_b])
        np.add.at(r, self.cab_a, -ic)
        np.add.at(r, self.cab_b, ic)
        np.add.at(r, self.buses, -self.GN * (x[self.buses] - self.VN))
        Ve = self.element_voltages(x)
        ie = f + g * (Ve - v)
        if gv > 0:
            ie = ie + gv * Ve
        pm = self.pos >= 0
        np.add.at(r, self.pos[pm], -ie[pm])
        nm = self.neg >= 0
        np.add.at(r, self.neg[nm], ie[nm])
        return r

    def iterate(self, v, iph, gv, xprev=None):
        """One Newton iteration from element voltages v (and, in the
        correction form, the previous node voltages xprev; None = 0).
        Returns (x, v_new, sum|dv|, max|dv|, LU values)."""
        I, dIdV, _ = self.module(v, iph)
        f, g = -I, -dIdV
        vals = self._values(g, gv)
        if self.form == "direct":
            ieq = f - g * v
            b = self.bbase.copy()
            pm = self.pos >= 0
            np.add.at(b, self.pos[pm], -ieq[pm])
            nm = self.neg >= 0
            np.add.at(b, self.neg[nm], ieq[nm])
        else:
            xprev = np.zeros(self.n) if xprev is None else xprev
            b = self.residual(xprev, v, iph, gv, (f, g))
        bp = self._permute_rows(b)
        st, bad = glu_csc_sequential(self.s, vals)
        if st != 0:
            raise RuntimeError(f"oracle GLU status {st} at column {bad}")
        substitute_lu_csc(self.s, vals, bp)
        x = bp[self.colmap]
        if self.form == "correction":
            x = xprev + x
        vn = self.element_voltages(x)
        st, v2, diff, mx = newton_update(v, vn)
        if st != 0:
            raise RuntimeError("oracle Newton: damping below 1e-13")
        return x, v2, diff, mx, vals

    # last-stage noise-floor rule (as the library): 20 iterations with
    # max |delta| at or below the form's FP64 floor
    NOISE = {"direct": 1e-8, "correction": 1e-11}
    STALL = 20

    def solve(self, iph, v0, ladder=True, tol=1e-10, max_iters=5000, x0=None):
        """Returns (x, v, iterations); self.last_max holds the final max |delta|
        (> tol when the iteration stopped at the FP64 noise floor) and
        self.stalled whether it did."""
        v = np.asarray(v0, float).copy()
        it = 0
        stages = LADDER if ladder else LADDER[-1:]
        x = None if x0 is None else np.asarray(x0, float).copy()
        self.stalled = False
        for k, gv in enumerate(stages):
            last = k == len(stages) - 1
            stall = 0
            while True:
                it += 1
                if it > max_iters:
                    raise RuntimeError("oracle Newton did not converge")
                x, v, diff, mx, lu = self.iterate(v, iph, gv, x)
                if not last and diff < 1e-6 * self.n:
                    break
                if last:
                    if mx <= tol:
                        break
                    stall = stall + 1 if mx <= self.NOISE[self.form] else 0
                    if stall >= self.STALL:
                        self.stalled = True
                        break
        self.lu = lu
        self.last_max = mx
        return x, v, it

    def with_cable(self, c, g):
        """A copy whose cable c has conductance g (the other data shared)."""
        import copy
        other = copy.copy(self)
        other.abase = self.abase.copy()
        other.cab_g = self.cab_g.copy()
        other.abase[self.cab_slots[c]] += (g - self.cab_g[c]) * np.array([1.0, 1.0, -1.0, -1.0])
        other.cab_g[c] = g
        return other

    def adjoint(self, x, v, iph):
        """Gradients of P = sum_bus V G_N (V - 600) at a converged state (x, v):
        A(v)^T lam = dP/dx with the oracle's own factors (U^T then L^T), then
        dP/dG_m = 9e-3 dI/dIph (lam[+] - lam[-]) and
        dP/dg_c = -(lam[a] - lam[b]) (x_a - x_b).  Returns (dP/dG, dP/dg)."""
        _, dIdV, dIdIph = self.module(v, iph)
        vals = self._values(-dIdV, -1.0)
        st, bad = glu_csc_sequential(self.s, vals)
        if st != 0:
            raise RuntimeError(f"oracle GLU status {st} at column {bad}")
        gx = np.zeros(self.n)
        gx[self.buses] = self.GN * (2.0 * x[self.buses] - self.VN)
        g = np.zeros(self.n)
        g[self.colmap] = gx
        solve_transpose_csc(self.s, vals, g)
        lam = g[self.rowmap]
        lp = np.where(self.pos >= 0, lam[np.maximum(self.pos, 0)], 0.0)
        ln = np.where(self.neg >= 0, lam[np.maximum(self.neg, 0)], 0.0)
        dG = 9e-3 * dIdIph * (lp - ln)
        dg = -(lam[self.cab_a] - lam[self.cab_b]) * (x[self.cab_a] - x[self.cab_b])
        return dG, dg

    def power(self, x, buses=None):
        V = x[self.buses if buses is None else np.asarray(buses)]
        return float(np.sum(V * self.GN * (V - self.VN)))
