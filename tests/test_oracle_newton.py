"""Pins the oracle's Newton loop (oracle.PVNewton) on the CPU.

The GPU Newton tests compare node voltages against PVNewton, so its
assembly, its residual and its converged state are checked here by what the
circuit laws fix, independently of PVNewton's own formulas:
  * the companion matrix A(v) is the Jacobian of the circuit's KCL currents
    (central differences of an explicit loop over the circuit's branches),
    with and without the homotopy conductance;
  * the right-hand side b(v) makes A(v) x - b(v) equal those KCL currents
    when the modules are linearized at the voltages of x;
  * the branch-wise residual of the correction form equals b(v) - A(v) x
    formed densely;
  * the converged state satisfies Kirchhoff's current law with the exact
    (implicit-equation) module currents, in both forms;
  * both forms, and the ladder and last-stage-only paths, reach the same
    state; the correction form's last-stage updates fall below 1e-12 V
    where the direct form stalls at its noise floor.
"""
import numpy as np
import pytest

import phdthesis_albertojimenezruiz_amd as pv
from oracle import oracle as O

RS, I0, A_N, GN, VN = 0.3, 1e-10, 1.5, 1e4, 600.0


@pytest.fixture(scope="module")
def plant():
    return pv.PV(4, 10, 20, seed=0)


def _oracle(plant, form):
    pos, neg, rsh = plant.modules()
    a, b, g = plant.cables()
    rp, ci = plant.pattern()
    P, Prow = pv.reorder(plant.n, rp, ci)
    return O.PVNewton(plant.n, pos, neg, rsh, a, b, g, plant.buses(), P, Prow, form=form)


def kcl_currents(plant, x, iph, gv=-1.0, v_lin=None):
    """Net current leaving every node through its branches, written as an
    explicit loop over the circuit (cables, Norton buses, modules), with the
    module currents from the implicit diode equation at V = x[+] - x[-]; with
    v_lin given, each module is linearized at v_lin instead (I(v) + I'(v)
    (V - v)).  gv > 0 adds the homotopy conductance across every module."""
    pos, neg, rsh = plant.modules()
    a, b, gc = plant.cables()
    out = np.zeros(plant.n)
    for k in range(plant.ncab):
        i = gc[k] * (x[a[k]] - x[b[k]])
        out[a[k]] += i
        out[b[k]] -= i
    for bus in plant.buses():
        out[bus] += GN * (x[bus] - VN)
    for e in range(plant.nmod):
        V = (x[pos[e]] if pos[e] >= 0 else 0.0) - (x[neg[e]] if neg[e] >= 0 else 0.0)
        if v_lin is None:
            I = O.module_current(V, iph[e], I0, A_N, RS, rsh[e])[0]
        else:
            I0_, dI, _ = O.module_current(v_lin[e], iph[e], I0, A_N, RS, rsh[e])
            I = I0_ + dI * (V - v_lin[e])
        f = -I                     # element current flows + -> - as -I(V)
        if gv > 0:
            f = f + gv * V
        if pos[e] >= 0:
            out[pos[e]] += f
        if neg[e] >= 0:
            out[neg[e]] -= f
    return out


def _dense(on, vals):
    """Unpermuted dense matrix of the oracle's permuted CSC values."""
    s = on.s
    D = np.zeros((on.n, on.n))
    for c in range(on.n):
        for q in range(s.cp[c], s.cp[c + 1]):
            D[s.ri[q], c] = vals[q]
    return D[np.ix_(on.rowmap, on.colmap)]


def _state(plant, seed):
    rng = np.random.default_rng(seed)
    iph = 9e-3 * plant.irradiance(seed, 0.0)
    # node voltages around a string's operating range, incl. reverse bias
    x = np.zeros(plant.n)
    pos, neg, _ = plant.modules()
    for e in range(plant.nmod):
        lo = x[neg[e]] if neg[e] >= 0 else 0.0
        x[pos[e]] = lo + rng.uniform(-8.0, 34.0)
    for bus, k in zip(plant.buses(), range(plant.I)):
        x[bus] = 600.0 + rng.uniform(-5, 5)
    return x, iph


@pytest.mark.parametrize("gv", [-1.0, 0.5])
def test_companion_matrix_is_the_kcl_jacobian(plant, gv):
    on = _oracle(plant, "direct")
    x, iph = _state(plant, 1)
    v = on.element_voltages(x)
    vals, _ = on.assemble(v, iph, gv)
    A = _dense(on, vals)
    # Richardson-extrapolated central differences (h = 2e-3, 1e-3 V): the
    # truncation error of the diode exponential stays below 1e-9 relative and
    # the rounding of the ~600 V node voltages below 1e-9 S
    cols = np.random.default_rng(2).choice(plant.n, 60, replace=False)

    def fd(j, h):
        e = np.zeros(plant.n)
        e[j] = h
        return (kcl_currents(plant, x + e, iph, gv) - kcl_currents(plant, x - e, iph, gv)) / (2 * h)

    for j in cols:
        J = (4 * fd(j, 1e-3) - fd(j, 2e-3)) / 3
        scale = np.abs(A[:, j]).max()
        assert np.allclose(J, A[:, j], rtol=1e-7, atol=1e-7 * scale), j


@pytest.mark.parametrize("gv", [-1.0, 0.5])
def test_companion_rhs_and_branch_residual(plant, gv):
    on = _oracle(plant, "correction")
    x, iph = _state(plant, 3)
    v = on.element_voltages(x) + np.random.default_rng(4).uniform(-0.5, 0.5, plant.nmod)
    vals, bp = on.assemble(v, iph, gv)
    A = _dense(on, vals)
    b = np.zeros(plant.n)
    b[:] = bp[on.rowmap]
    # A x - b = the KCL currents with the modules linearized at v
    want = kcl_currents(plant, x, iph, gv, v_lin=v)
    scale = np.abs(A).max() * np.abs(x).max()
    assert np.abs((A @ x - b) - want).max() <= 1e-12 * scale
    # the branch-wise residual is b - A x, without the dense cancellation
    r = on.residual(x, v, iph, gv)
    assert np.abs(r - (b - A @ x)).max() <= 1e-12 * scale
    assert np.abs(r + want).max() <= 1e-12 * np.abs(want).max() + 1e-12


@pytest.mark.parametrize("form", ["direct", "correction"])
def test_converged_state_satisfies_kcl(plant, form):
    """Each node's current imbalance, divided by the node's total conductance
    (the voltage error it amounts to), is within the form's accuracy: 1e-12 V
    for the correction form; 1e-9 V for the direct form, whose node voltages
    carry its noise floor.  (One ulp of a 600 V bus voltage is 1.1e-13 V, i.e.
    1.1e-9 A through the 1e4 S Norton source, so absolute currents would not
    do.)"""
    on = _oracle(plant, form)
    for scen in (-1, 1):
        iph = 9e-3 * plant.irradiance(scen, 0.0)
        x, v, it = on.solve(iph, plant.v0(), ladder=False, tol=1e-12)
        kcl = kcl_currents(plant, x, iph)
        vals, _ = on.assemble(v, iph, -1.0)
        G = np.abs(_dense(on, vals)).sum(axis=1)
        err_v = np.abs(kcl) / G
        bound = 1e-9 if form == "direct" else 1e-12
        assert err_v.max() <= bound, (scen, err_v.max())
        assert np.abs(v - on.element_voltages(x)).max() <= (1e-9 if form == "direct" else 1e-12)


def test_forms_and_ladder_reach_the_same_state(plant):
    iph = 9e-3 * plant.irradiance(2, 0.0)
    sols = {}
    for form in ("direct", "correction"):
        on = _oracle(plant, form)
        for ladder in (True, False):
            x, _, it = on.solve(iph, plant.v0(), ladder=ladder, tol=1e-12)
            sols[(form, ladder)] = (x, it, on.last_max, on.stalled)
    xc = sols[("correction", False)][0]
    for k, (x, it, last, stalled) in sols.items():
        assert np.abs(x - xc).max() <= (2e-10 if k[0] == "direct" else 1e-11), k
    # the ladder takes more iterations than the last stage alone
    assert sols[("correction", True)][1] > sols[("correction", False)][1]


def test_correction_form_lowers_the_noise_floor(plant):
    """Shaded cloud scenarios of the 4 x 10 x 20 plant: the direct form's
    last-stage updates stall above 1e-12 V on every one of them (the solve's
    rounding of the node voltages), the correction form's converge below it."""
    on_d, on_c = _oracle(plant, "direct"), _oracle(plant, "correction")
    for scen in range(4):
        iph = 9e-3 * plant.irradiance(scen, 0.0)
        xd, _, _ = on_d.solve(iph, plant.v0(), ladder=False, tol=1e-12)
        assert on_d.stalled and on_d.last_max > 1e-12
        xc, _, _ = on_c.solve(iph, plant.v0(), ladder=False, tol=1e-12)
        assert not on_c.stalled and on_c.last_max <= 1e-12
        assert np.abs(xd - xc).max() <= 2e-10


def test_correction_form_first_iterate_is_the_direct_solve(plant):
    """From x_prev = 0 the correction right-hand side is b itself."""
    on_d, on_c = _oracle(plant, "direct"), _oracle(plant, "correction")
    iph = 9e-3 * plant.irradiance(0, 0.0)
    v = plant.v0()
    xd, vd, *_ = on_d.iterate(v, iph, 100.0)
    xc, vc, *_ = on_c.iterate(v, iph, 100.0, np.zeros(plant.n))
    assert np.array_equal(xd, xc) and np.array_equal(vd, vc)


def _inverter_power(plant, x, iph, inv):
    """P of inverter inv from its strings' bottom-module currents (equal to
    V_bus G_N (V_bus - 600) by KCL, without the cancellation in V_bus - 600)."""
    pos, _, rsh = plant.modules()
    S, M = plant.S, plant.M
    tot = sum(O.module_current(x[pos[inv * S * M + s * M]], iph[inv * S * M + s * M],
                               I0, A_N, RS, rsh[inv * S * M + s * M])[0] for s in range(S))
    return x[plant.buses()[inv]] * tot


def test_oracle_adjoint_against_central_differences(plant):
    """PVNewton.adjoint (its own factors, U^T then L^T) against Richardson
    central differences of the owning inverter's power, with the oracle
    Newton re-converged to 1e-13 V at each perturbed parameter: modules
    (h = 0.1, 0.05 W/m^2) and cables (PVNewton.with_cable, h = 0.02, 0.01 S)."""
    on = _oracle(plant, "correction")
    iph = 9e-3 * plant.irradiance(1, 0.0)
    x, v, _ = on.solve(iph, plant.v0(), ladder=False, tol=1e-13)
    dG, dg = on.adjoint(x, v, iph)
    S, M = plant.S, plant.M
    rng = np.random.default_rng(8)
    for m in rng.choice(plant.nmod, 4, replace=False):
        inv = m // (S * M)

        def fd(h):
            out = []
            for sg in (1, -1):
                i2 = iph.copy()
                i2[m] += sg * 9e-3 * h
                xx, _, _ = on.solve(i2, v, ladder=False, tol=1e-13)
                out.append(_inverter_power(plant, xx, i2, inv))
            return (out[0] - out[1]) / (2 * h)

        want = (4 * fd(0.05) - fd(0.1)) / 3
        assert abs(dG[m] - want) <= 1e-6 * abs(want) + 1e-8, (m, dG[m], want)
    a, b, gc = plant.cables()
    for c in rng.choice(plant.ncab, 3, replace=False):
        inv = c // S

        def fdc(h):
            out = []
            for sg in (1, -1):
                o2 = on.with_cable(c, gc[c] + sg * h)
                xx, _, _ = o2.solve(iph, v, ladder=False, tol=1e-13)
                out.append(_inverter_power(plant, xx, iph, inv))
            return (out[0] - out[1]) / (2 * h)

        want = (4 * fdc(0.01) - fdc(0.02)) / 3
        assert abs(dg[c] - want) <= 1e-6 * abs(want) + 3e-8, (c, dg[c], want)
    # with_cable leaves the original untouched
    assert np.array_equal(on.cab_g, gc)
