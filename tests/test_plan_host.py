"""Host half of the symbolic plan (CPU): the product's 006-010 restatement
against the oracle and, where it builds, the reference's own units."""
import numpy as np
import pytest

import phdthesis_albertojimenezruiz_amd as pv
from oracle import oracle as O
from tests._patterns import (HUB_CASES, HUB_STAR, csr_from_sets, hub_pattern, plant_pattern,
                             random_pattern)


def _cases():
    rng = np.random.default_rng(21)
    out = []
    for t in range(8):
        n = int(rng.integers(10, 150))
        rows = random_pattern(rng, n, float(rng.uniform(0.01, 0.06)), symmetric=bool(t % 2))
        rp, ci = csr_from_sets(rows)
        out.append((n, *O.symbolic_fill(n, rp, ci)))
    for I, S, M in ((1, 2, 3), (4, 10, 20)):
        n, rows = plant_pattern(I, S, M)
        rp, ci = csr_from_sets(rows)
        if O.ref() is not None:          # the reference pipeline's AMD ordering
            P = O.amd_order(n, rp, ci)
            rp, ci, _ = O.permute_sym(n, rp, ci, np.zeros(len(ci)), P)
        out.append((n, *O.symbolic_fill(n, rp, ci)))
    return out


CASES = _cases()


@pytest.mark.parametrize("case", range(len(CASES)))
def test_plan_matches_oracle(case):
    n, rp, ci = CASES[case]
    plan = pv.Plan(n, rp, ci)
    s = O.csr_csc(n, rp, ci)
    assert plan.nnz == len(ci)
    for name, want in (("csr_rowptr", rp), ("csr_colidx", ci), ("csc_colptr", s.cp),
                       ("csc_rowidx", s.ri), ("csr2csc", s.csr2csc), ("csc2csr", s.csc2csr),
                       ("csr_diag", s.csr_diag), ("csc_diag", s.csc_diag)):
        assert np.array_equal(plan.array(name), want), name
    for kind, pre in ((0, "lu"), (1, "l"), (2, "u")):
        dp, di = O.deps(kind, n, rp, ci)
        lv = O.levels(n, dp, di)
        assert np.array_equal(plan.array(pre + "_start_levels"), lv.start_levels)
        assert np.array_equal(plan.array(pre + "_start_chains"), lv.start_chains)
        assert np.array_equal(plan.array(pre + "_chains"), lv.chains)
        if kind == 0:
            assert np.array_equal(plan.array("max_per_level"), O.max_per_level(lv, rp))
        if O.ref() is not None:
            rdp, rdi = O.deps(kind, n, rp, ci, use_ref=True)
            rlv = O.levels(n, rdp, rdi, use_ref=True)
            assert np.array_equal(plan.array(pre + "_chains"), rlv.chains)
            assert np.array_equal(plan.array(pre + "_start_chains"), rlv.start_chains)
            assert np.array_equal(plan.array(pre + "_start_levels"), rlv.start_levels)


def test_plan_plant_structure():
    n, rows = plant_pattern(40, 24, 28)
    rp, ci = csr_from_sets(rows)
    plan = pv.Plan(n, rp, ci)
    assert n == 26920 and plan.nnz == 80680
    assert plan.info("ncomp") == 40
    assert plan.info("fused_lsolve") == 1
    assert plan.info("nlev_lu") == 2 and plan.info("nchain_lu") == 1000


def test_plan_rejects_unclosed_and_bad_patterns():
    # (2,0),(0,1) present but (2,1) missing: not closed under elimination
    rp, ci = csr_from_sets([{0, 1}, {1}, {0, 2}])
    with pytest.raises(pv.PvgluError) as e:
        pv.Plan(3, rp, ci)
    assert e.value.code == pv.E_ARG
    rp, ci = csr_from_sets([{0}, {0}, {2}])            # no diagonal in row 1
    with pytest.raises(pv.PvgluError) as e:
        pv.Plan(3, rp, ci)
    assert e.value.code == pv.E_NODIAG
    with pytest.raises(pv.PvgluError):
        pv.Plan(2, np.array([0, 2, 1], np.int32), np.array([0, 1], np.int32))


def test_numeric_entry_points_fail_without_device():
    if pv.lib().pvglu_device_count() > 0:
        pytest.skip("a device is present")
    n, rows = plant_pattern(1, 2, 3)
    rp, ci = csr_from_sets(rows)
    plan = pv.Plan(n, rp, ci)
    with pytest.raises(pv.PvgluError) as e:
        plan.refactor(1, 2, 1)
    assert e.value.code == pv.E_HIP


def _plan_of(n, rows, amd=False):
    rp, ci = csr_from_sets(rows)
    if amd:
        P, _ = pv.reorder(n, rp, ci)
        rp, ci, _ = O.permute_sym(n, rp, ci, np.zeros(len(ci)), P)
    rp, ci = O.symbolic_fill(n, rp, ci)
    return pv.Plan(n, rp, ci), rp, ci


@pytest.mark.parametrize("case", range(len(HUB_CASES)))
def test_star_column_classification_hubs(case):
    """Column flag bit 1 (star column: updates only on the diagonal, at most
    32, strided operands) is set on the hub exactly when its strings have
    equal lengths and there are at most 32 of them.  A string column is a
    one-update star (its pivot update), the bottom column of a string has no
    update and is not.  Bit 0 (diagonal and forward row share their L
    operands) holds on every column of these trees."""
    lengths, tail = HUB_CASES[case]
    n, rows = hub_pattern(lengths, tail)
    plan, _, _ = _plan_of(n, rows)
    flags = plan.array("col_flags")
    hub = sum(lengths)
    assert bool(flags[hub] & 2) == HUB_STAR[case]
    bottoms = set(np.cumsum([0] + lengths[:-1]).tolist())
    for c in range(n):
        if c != hub:
            assert bool(flags[c] & 2) == (c not in bottoms), c
    assert np.all(flags & 1)


def test_star_columns_and_tridiagonal_segments_of_the_plant():
    """40x24x28 plant in AMD order: the star columns are exactly the 40 DC
    bus columns (25 entries: 24 combiners and the bus), and the strings'
    columns run as tridiagonal segments forwarding the previous column's L
    entry (segment flags bits 2 and 3, entry 2)."""
    I, S, M = 40, 24, 28
    n, rows = plant_pattern(I, S, M)
    plan, rp, ci = _plan_of(n, rows, amd=True)
    flags = plan.array("col_flags")
    cp = plan.array("csc_colptr")
    nnz_col = np.diff(cp)
    buses = np.nonzero(nnz_col == S + 1)[0]
    assert len(buses) == I
    assert np.all(flags[buses] & 2)
    # no other column with more than the register path's 4 entries
    assert np.array_equal(np.nonzero(nnz_col > 4)[0], buses)
    seg = plan.array("lu_segments").reshape(-1, 4)
    starts = seg[seg[:, 0] > 0]
    tri = starts[(starts[:, 3] & 12) == 12]
    # forwarded operand: entry 1 of the bottom column (diagonal, L), entry 2
    # of a tridiagonal column (U, diagonal, L), entry 3 of the bus column's
    # register window (its last four entries, the L entry last)
    k = (tri[:, 3] >> 4) & 3
    assert set(k.tolist()) == {1, 2, 3}
    assert np.count_nonzero(k == 3) == I
    # every string column but the bottom one is a tridiagonal step, except
    # one per inverter whose forwarded operand is in memory: the last
    # string's column after its bottom node (eliminated a level earlier)
    assert tri[:, 0].sum() == I * S * (M - 1) - I
