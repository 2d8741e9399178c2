"""The reference's own plant on the host (CPU): the library's builder against
the oracle's independent restatement of PV/003_create_electric_system.cpp,
the reorder (AMD + zero-diagonal row swaps, Q1) against the reference's AMD
and the oracle's 004b, and the symbolic plan (006-009) against the
reference's own units on the plant's unsymmetric filled pattern (Q2)."""
import numpy as np
import pytest

import phdthesis_albertojimenezruiz_amd as pv
from oracle import oracle as O
from oracle.refplant import RefPlantOracle

needs_ref = pytest.mark.skipif(O.ref() is None, reason="oracle/_ref not built (reference absent)")

# (blocks, panels_x, panels_y, volts per panel, interconnection, bypass)
CASES = [(1, 4, 3, 23.0, "nothing", True), (2, 5, 4, 23.0, "honeycomb", True),
         (3, 3, 3, 23.0, "full", True), (2, 6, 4, 23.0, "bridge", True),
         (2, 4, 2, 23.0, "nothing", False), (1, 50, 50, 23.0, "nothing", True)]


def _pair(case, panel_g=None):
    b, x, y, v, inter, byp = case
    rp = pv.RefPlant(b, x, y, v, inter, panel_irradiance=panel_g, bypass=byp)
    ro = RefPlantOracle(b, x, y, v, interconnection=pv.INTERCONNECTION[inter],
                        panel_g=panel_g, bypass=byp)
    return rp, ro


@pytest.mark.parametrize("case", CASES[:5])
def test_builder_matches_oracle_restatement(case):
    rng = np.random.default_rng(1)
    g = rng.uniform(100.0, 1000.0, case[0] * case[1] * case[2])
    rp, ro = _pair(case, g)
    assert rp.n == ro.n
    E, V = rp.elements()
    Eo, Vo = ro.elements()
    assert np.array_equal(E, Eo) and np.array_equal(V, Vo)
    assert np.array_equal(rp.node_unknowns(), np.array(ro.node_x))
    rpp, ci = rp.pattern()
    orp, oci = ro.pattern()
    assert np.array_equal(rpp, orp) and np.array_equal(ci, oci)
    assert rp.info("inverters") == case[0]
    assert rp.info("vsources") == case[0] + 2           # one per inverter, PCC re + im
    assert rp.info("nonlinear") == case[0] * case[1] * case[2] * (2 if case[5] else 1)


def test_builder_sizes_of_the_reference_default():
    """The reference's main.cpp plant: 1 inverter, 50 x 50 panels, 23 V per
    panel, no cross-ties: 2500 internal nodes + 2450 string nodes + 4 block
    nodes + 2 AC nodes (4 unknowns) + the PCC (2) + 3 source currents."""
    rp = pv.RefPlant(1, 50, 50, 23.0)
    assert rp.n == 2500 + 50 * 49 + 4 + 4 + 2 + 3
    assert rp.info("panels") == 2500


@needs_ref
@pytest.mark.parametrize("case", CASES)
def test_reorder_swaps_rows_and_is_an_involution(case):
    rp, ro = _pair(case)
    n = rp.n
    rpp, ci = rp.pattern()
    P, Prow = pv.reorder(n, rpp, ci)
    assert np.array_equal(P, O.amd_order(n, rpp, ci))          # AMD 1.1 bit-exact
    orp, oci, _ = O.permute_sym(n, rpp, ci, np.zeros(len(ci)), P)
    st, _, _, _, perm = O.remove_zeros_diagonal(n, orp, oci)
    assert st == 0 and np.array_equal(Prow, perm)               # 004b
    assert not np.array_equal(Prow, np.arange(n))               # voltage sources swap rows
    assert np.array_equal(Prow[Prow], np.arange(n))             # Q1: an involution


@needs_ref
@pytest.mark.parametrize("case", CASES)
def test_plan_of_the_unsymmetric_filled_pattern_matches_reference(case):
    rp, ro = _pair(case)
    n = rp.n
    rpp, ci = rp.pattern()
    P, Prow = pv.reorder(n, rpp, ci)
    frp, fci = ro.prepare(P, Prow)
    assert len(fci) > len(ci)                                   # real fill-in
    S = [set(fci[frp[i]:frp[i + 1]].tolist()) for i in range(n)]
    assert any(j not in S[i] or i not in S[j] for i in range(n) for j in S[i])   # unsymmetric
    plan = pv.Plan(n, frp, fci)
    for kind, pre in ((0, "lu"), (1, "l"), (2, "u")):
        rdp, rdi = O.deps(kind, n, frp, fci, use_ref=True)
        rlv = O.levels(n, rdp, rdi, use_ref=True)
        assert np.array_equal(plan.array(pre + "_start_levels"), rlv.start_levels), pre
        assert np.array_equal(plan.array(pre + "_start_chains"), rlv.start_chains), pre
        assert np.array_equal(plan.array(pre + "_chains"), rlv.chains), pre
        if kind == 0:
            assert np.array_equal(plan.array("max_per_level"),
                                  O.ref_max_per_level(n, frp, fci, rdp, rdi))


def test_residual_is_b_minus_A_x():
    """The oracle's element-wise residual equals b - A x formed from its own
    assembly (voltage sources, AC blocks, inverter injections, diodes)."""
    rp, ro = _pair((2, 5, 4, 23.0, "honeycomb", True))
    rng = np.random.default_rng(2)
    ro.v = rng.uniform(-1.0, 0.8, len(ro.v))
    ro.inv_state = np.array([[24000.0, 300.0, 40.0], [25100.0, -200.0, 35.0]])
    x = rng.uniform(-100.0, 1200.0, ro.n)
    for gv in (-1.0, 0.3):
        A, b = ro.assemble(gv)
        Ax = np.zeros(ro.n)
        for (i, j), v in A.items():
            Ax[i] += v * x[j]
        r = ro.residual(x, gv)
        scale = max(abs(v) for v in A.values()) * np.abs(x).max()
        assert np.abs(r - (b - Ax)).max() <= 1e-13 * scale
