"""CPU restatement of the reference's own PV plant and its Newton loop.

TEST INFRASTRUCTURE ONLY (the product never imports this module).

RefPlantOracle restates, independently of the library's C++ builder:
  * create_full_electric_system / create_block_of_panels_simple /
    add_solar_panel (PV/003_create_electric_system.cpp:87-126,184-382):
    the element lists and the unknowns in the reference's creation order;
  * create_matrix_and_vector (PV/002b_sistema_electrico_base_matriz.cpp:
    158-240) with Matrix_Add_Impedance / _ac / Current_Source /
    Voltage_Source, the nonlinear companion stamps and the inverter
    injections, in the reference's element order;
  * the Newton loop (PV/000_PVPP_simulator.cpp:163-235): ladder, permutation,
    GLU_CSC_CPU_sequential + the CSC substitution (oracle C), un-permutation,
    the 011 update (PV/011_calculate_new_iteration.cpp:37-86) and the
    inverter refresh (:88-109); direct or correction form.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from oracle import oracle as O

_I = np.int32
RS = 0.262 * (175.0 / 400.0)          # PV/003_create_electric_system.cpp:16-25
RSH = 90.85
IPV = 8.11 * 400.0 / 175.0 * 1.05
K0 = -0.45 / 100.0
T0, G0 = 25.0, 1000.0
WIRE = 1e-5                            # PV/main.h:76
VAC = 25000.0                          # PV/main.h:7
ADM, ACADM, ISRC, VSRC, NONLIN, INVERTER = range(6)


class RefPlantOracle:
    def __init__(self, blocks, px, py, volts_per_panel, interconnection=0,
                 irradiance=1000.0, panel_g=None, temperature=25.0, bypass=True,
                 form="direct"):
        self.form = form
        self.T = temperature
        self.node_x = []          # node (1-based) -> unknown
        self.nx = 0
        self.adm, self.acadm, self.isrc, self.vsrc = [], [], [], []
        self.nonlin, self.inv = [], []
        npanels = blocks * px * py
        g = np.full(npanels, float(irradiance)) if panel_g is None else np.asarray(panel_g, float)
        main = []
        v_dc = px * volts_per_panel
        for blk in range(blocks):
            ext = self._block(blk, px, py, interconnection, g, bypass)
            a1 = self._node_ac()
            a2 = self._node_ac()
            main.append((a1, a2))
            self.adm.append((ext[1], 0, 1.0 / WIRE))
            vs = self._vsource(ext[0], ext[1], v_dc)
            self.inv.append((a1, a2, vs, v_dc))
            if blk:
                self.acadm.append((main[blk][0], main[blk - 1][0], 1.0 / WIRE, 0.0))
                self.acadm.append((main[blk][1], main[blk - 1][1], 1.0 / WIRE, 0.0))
        self.pcc = self._node_ac()
        self.pcc_current = self._vsource(self.pcc, 0, VAC)
        self._vsource(self.pcc + 1, 0, 0.0)
        for (na_, nb_) in ((self.pcc, main[0][0]), (0, main[0][1]),
                           (self.pcc, main[-1][0]), (0, main[-1][1])):
            self.acadm.append((na_, nb_, 1.0 / WIRE, 0.0))
        self.n = self.nx
        self.v = np.zeros(len(self.nonlin))
        self.inv_state = np.array([[VAC, 0.0, 0.0] for _ in self.inv]).reshape(-1, 3)
        self.kind = np.array([k for (_, _, k) in self.nonlin], _I)

    # ---------------------------------------------------------- builder
    def _node(self):
        self.node_x.append(self.nx)
        self.nx += 1
        return len(self.node_x)

    def _node_ac(self):
        re = self._node()
        self._node()
        return re

    def _vsource(self, n1, n2, E):
        self.vsrc.append((n1, n2, E, self.nx))
        self.nx += 1
        return self.nx - 1

    def _panel(self, n_in, n_out, irr, bypass):
        internal = self._node()
        self.adm.append((internal, n_in, 1.0 / RS))
        self.adm.append((internal, n_out, 1.0 / RSH))
        self.nonlin.append((internal, n_out, 0))
        self.isrc.append((internal, n_out, IPV * irr / G0 * (1.0 + K0 * (self.T - T0))))
        if bypass:
            self.nonlin.append((n_out, n_in, 1))

    def _block(self, blk, px, py, inter, g, bypass):
        n_in = self._node()
        n_out = self._node()
        rows = []
        for i in range(py):
            row, prev = [], None
            for j in range(px):
                first = n_in if j == px - 1 else self._node()
                second = n_out if j == 0 else prev[0]
                prev = (first, second)
                row.append(prev)
                self._panel(first, second, g[(blk * py + i) * px + j], bypass)
            rows.append(row)
        for i in range(py - 1):
            for j in range(px - 1):
                tie = {0: False, 1: True, 2: (i + j) % 2 == 1,
                       3: ((i % 3) == 0) != ((j % 2) == 0)}[inter]
                if tie:
                    self.adm.append((rows[i][j][0], rows[i + 1][j][0], 1.0 / WIRE))
        e_in = self._node()
        e_out = self._node()
        self.adm.append((e_in, n_in, 1.0 / WIRE))
        self.adm.append((e_out, n_out, 1.0 / WIRE))
        return e_in, e_out

    def x(self, node):
        return self.node_x[node - 1]

    def elements(self):
        """The element table in stamping order, as pvglu_refplant_elements
        returns it: rows (kind, n1, n2, aux, index) and (v0, v1)."""
        E, V = [], []
        for (a, b, g) in self.adm:
            E.append((ADM, a, b, 0, 0)), V.append((g, 0.0))
        for (a, b, re, im) in self.acadm:
            E.append((ACADM, a, b, 0, 0)), V.append((re, im))
        for (a, b, v) in self.isrc:
            E.append((ISRC, a, b, 0, 0)), V.append((v, 0.0))
        for (a, b, v, xx) in self.vsrc:
            E.append((VSRC, a, b, xx, 0)), V.append((v, 0.0))
        for e, (a, b, k) in enumerate(self.nonlin):
            E.append((NONLIN, a, b, k, e)), V.append((self.T, 0.0))
        for i, (a, b, vs, vdc) in enumerate(self.inv):
            E.append((INVERTER, a, b, vs, i)), V.append((vdc, 0.0))
        return np.array(E, _I), np.array(V)

    # --------------------------------------------------------- assembly
    def diodes(self, V):
        V = np.ascontiguousarray(V, float)
        f, df = np.zeros(len(V)), np.zeros(len(V))
        O.lib().orc_ref_diodes(len(V), self.kind.ctypes.data_as(C.POINTER(C.c_int)),
                               O._dp(V), C.c_double(self.T), O._dp(f), O._dp(df))
        return f, df

    def assemble(self, gv):
        """create_matrix_and_vector at the current state: dict (i, j) -> value
        built in the reference's stamping order, and b."""
        A, b = {}, np.zeros(self.n)

        def add(i, j, v):
            A[(i, j)] = A.get((i, j), 0.0) + v

        def imp(n1, n2, g):
            if n1:
                add(self.x(n1), self.x(n1), g)
            if n2:
                add(self.x(n2), self.x(n2), g)
            if n1 and n2:
                add(self.x(n1), self.x(n2), -g)
                add(self.x(n2), self.x(n1), -g)

        def cur(n1, n2, v):
            if n1:
                b[self.x(n1)] += v
            if n2:
                b[self.x(n2)] -= v

        def blk(p, q, re, im):
            if not p or not q:
                return
            i, j = self.x(p), self.x(q)
            add(i, j, re)
            add(i, j + 1, -im)
            add(i + 1, j, im)
            add(i + 1, j + 1, re)

        for (a, c_, g) in self.adm:
            imp(a, c_, g)
        for (a, c_, re, im) in self.acadm:
            blk(a, a, re, im)
            blk(a, c_, -re, -im)
            blk(c_, a, -re, -im)
            blk(c_, c_, re, im)
        for (a, c_, v) in self.isrc:
            cur(a, c_, v)
        for (a, c_, E, xx) in self.vsrc:
            if a:
                add(self.x(a), xx, -1.0)
                add(xx, self.x(a), 1.0)
            if c_:
                add(self.x(c_), xx, 1.0)
                add(xx, self.x(c_), -1.0)
            b[xx] += E
        f, df = self.diodes(self.v)
        for e, (a, c_, _) in enumerate(self.nonlin):
            ieq = f[e] - df[e] * self.v[e]
            imp(a, c_, df[e])
            cur(a, c_, -ieq)
            if gv > 0:
                imp(a, c_, gv)
        for (a, c_, vs, vdc), st in zip(self.inv, self.inv_state):
            den = st[0] * st[0] + st[1] * st[1]
            p0 = vdc * st[2]
            cur(a, c_, -(p0 * st[0] / den))
            cur(a + 1 if a else 0, c_ + 1 if c_ else 0, -(p0 * st[1] / den))
        return A, b

    def drop(self, xx, n1, n2):
        r = xx[self.x(n1)] if n1 else 0.0
        if n2:
            r -= xx[self.x(n2)]
        return r

    def residual(self, xx, gv):
        """b - A x element by element (differences first), at unknowns xx."""
        r = np.zeros(self.n)

        def leave(n1, n2, i):
            if n1:
                r[self.x(n1)] -= i
            if n2:
                r[self.x(n2)] += i

        for (a, c_, g) in self.adm:
            leave(a, c_, g * self.drop(xx, a, c_))
        for (a, c_, re, im) in self.acadm:
            dr = self.drop(xx, a, c_)
            di = self.drop(xx, a + 1 if a else 0, c_ + 1 if c_ else 0)
            leave(a, c_, re * dr - im * di)
            leave(a + 1 if a else 0, c_ + 1 if c_ else 0, im * dr + re * di)
        for (a, c_, v) in self.isrc:
            if a:
                r[self.x(a)] += v
            if c_:
                r[self.x(c_)] -= v
        for (a, c_, E, xv) in self.vsrc:
            leave(a, c_, -xx[xv])
            r[xv] += E - self.drop(xx, a, c_)
        f, df = self.diodes(self.v)
        for e, (a, c_, _) in enumerate(self.nonlin):
            Vx = self.drop(xx, a, c_)
            i = f[e] + df[e] * (Vx - self.v[e])
            if gv > 0:
                i = i + gv * Vx
            leave(a, c_, i)
        for (a, c_, vs, vdc), st in zip(self.inv, self.inv_state):
            den = st[0] * st[0] + st[1] * st[1]
            p0 = vdc * st[2]
            for (p, q, val) in ((a, c_, -(p0 * st[0] / den)),
                                (a + 1 if a else 0, c_ + 1 if c_ else 0, -(p0 * st[1] / den))):
                if p:
                    r[self.x(p)] += val
                if q:
                    r[self.x(q)] -= val
        return r

    def pattern(self):
        A, _ = self.assemble(1.0)
        rows = [[] for _ in range(self.n)]
        for (i, j) in A:
            rows[i].append(j)
        rp, ci = [0], []
        for r in rows:
            ci.extend(sorted(r))
            rp.append(len(ci))
        return np.array(rp, _I), np.array(ci, _I)

    # --------------------------------------------------------- ordering
    def prepare(self, P, Prow):
        """The permuted, filled pattern for the given ordering (reorder_matrix's
        P and Prow) and the CSC slot of every (row, column) of A."""
        n = self.n
        self.P = np.asarray(P, _I)
        self.Pinv = O.invert_permutation(self.P)
        self.Prow = np.asarray(Prow, _I)
        rp, ci = self.pattern()
        prp, pci, _ = O.permute_sym(n, rp, ci, np.zeros(len(ci)), self.P)
        frp, fci, _ = O.permute_rows(n, prp, pci, np.zeros(len(pci)), self._prow_swap())
        frp, fci = O.symbolic_fill(n, frp, fci)
        self.s = O.csr_csc(n, frp, fci)
        return frp, fci

    def _prow_swap(self):
        """orc_permute_rows takes new[p[i]] = old[i]; reorder_matrix's row k
        holds old row Prow[k], i.e. p = Prow^-1 (= Prow for an involution)."""
        return O.invert_permutation(self.Prow)

    def permuted_system(self, A, b):
        """fill_values_CSC_only after interchange_row_columns(Pinv) and
        exchange_rows(Prow) (PV/000_PVPP_simulator.cpp:193-210)."""
        s = self.s
        vals = np.zeros(len(s.ri))
        rowmap = self.Prow[self.Pinv]          # exchange_rows: new[Prow[i]] = old[i]
        pos = {}
        for c in range(self.n):
            for q in range(s.cp[c], s.cp[c + 1]):
                pos[(int(s.ri[q]), c)] = q
        for (i, j), v in A.items():
            vals[pos[(int(rowmap[i]), int(self.Pinv[j]))]] = v
        bp = np.zeros(self.n)
        bp[rowmap] = b
        return vals, bp

    # ----------------------------------------------------------- Newton
    NOISE = {"direct": 1e-8, "correction": 1e-11}

    def refresh(self, xx):
        """011:88-109: (V2 re, V2 im, I1) per inverter; returns (sum, max)."""
        tot, mx = 0.0, 0.0
        for k, (a, c_, vs, _) in enumerate(self.inv):
            old = self.inv_state[k].copy()
            new = np.array([self.drop(xx, a, c_),
                            self.drop(xx, a + 1 if a else 0, c_ + 1 if c_ else 0), xx[vs]])
            self.inv_state[k] = new
            d = np.abs(new - old)
            tot += float(d.sum())
            mx = max(mx, float(d.max()))
        return tot, mx

    def iterate(self, gv, xprev):
        A, b = self.assemble(gv)
        if self.form == "correction":
            b = self.residual(xprev, gv)
        vals, bp = self.permuted_system(A, b)
        st, bad = O.glu_csc_sequential(self.s, vals)
        if st != 0:
            raise RuntimeError(f"oracle GLU status {st} at column {bad}")
        O.substitute_lu_csc(self.s, vals, bp)
        xx = bp[self.Pinv]
        if self.form == "correction":
            xx = xprev + xx
        vn = np.array([self.drop(xx, a, c_) for (a, c_, _) in self.nonlin])
        st, v2, diff, mx = O.newton_update(self.v, vn)
        if st != 0:
            raise RuntimeError("oracle Newton: damping below 1e-13")
        self.v = v2
        s_, m_ = self.refresh(xx)
        return xx, diff + s_, max(mx, m_)

    def solve(self, tol=1e-10, ladder=True, max_iters=5000):
        stages = O.LADDER if ladder else O.LADDER[-1:]
        xx = np.zeros(self.n)
        it = 0
        self.stalled = False
        for k, gv in enumerate(stages):
            last = k == len(stages) - 1
            stall = 0
            while True:
                it += 1
                if it > max_iters:
                    raise RuntimeError("oracle Newton did not converge")
                xx, diff, mx = self.iterate(gv, xx)
                if not last and diff < 1e-6 * self.n:
                    break
                if last:
                    if mx <= tol:
                        break
                    stall = stall + 1 if mx <= self.NOISE[self.form] else 0
                    if stall >= 20:
                        self.stalled = True
                        break
        self.last_max = mx
        return xx, it

    def pcc_power(self, xx):
        V = complex(self.drop(xx, self.pcc, 0), self.drop(xx, self.pcc + 1, 0))
        I = complex(xx[self.pcc_current], xx[self.pcc_current + 1])
        return V * I.conjugate()
