"""Pins the CPU oracle (oracle/pvglu_oracle.c) before it is trusted.

Each check is independent of the oracle's own formulation:
  * the reference's own CPU units compiled in oracle/_ref (006-009, the CSC
    GLU and the CSC triangular solve), bit-exact where they exist;
  * dense textbook elimination (symbolic by numeric non-cancellation, numeric
    by Doolittle without pivoting) and dense solves;
  * closed forms (scipy.special.lambertw for the diode, the halving
    recurrence for the Newton update) and hand-worked examples.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests._patterns import (csr_from_sets, dd_values, dense_from_csr, dense_lu_nopivot,
                             plant_pattern, random_pattern)

needs_ref = pytest.mark.skipif(O.ref() is None, reason="oracle/_ref not built (reference absent)")


def _filled(rows):
    rp, ci = csr_from_sets(rows)
    n = len(rows)
    return n, O.symbolic_fill(n, rp, ci)


# --------------------------------------------------------------- 005a / 005b
@pytest.mark.parametrize("seed", range(12))
def test_symbolic_fill_matches_numeric_elimination(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(5, 48))
    rows = random_pattern(rng, n, float(rng.uniform(0.03, 0.2)), symmetric=bool(seed % 2))
    rp, ci = csr_from_sets(rows)
    frp, fci = O.symbolic_fill(n, rp, ci)
    # Dense numeric elimination with generic values: nonzeros of L+U are the
    # fill pattern (cancellation has probability zero).
    vals = dd_values(rng, n, rp, ci)
    LU = dense_lu_nopivot(dense_from_csr(n, rp, ci, vals))
    want = [set(np.nonzero(np.abs(LU[i]) > 0)[0].tolist()) | set(ci[rp[i]:rp[i + 1]].tolist())
            for i in range(n)]
    got = [set(fci[frp[i]:frp[i + 1]].tolist()) for i in range(n)]
    assert got == want
    # sorted rows, original pattern contained
    for i in range(n):
        row = fci[frp[i]:frp[i + 1]]
        assert np.all(np.diff(row) > 0)


def test_symbolic_fill_closure_invariant_large():
    rng = np.random.default_rng(7)
    n = 400
    rows = random_pattern(rng, n, 0.004)
    rp, ci = csr_from_sets(rows)
    frp, fci = O.symbolic_fill(n, rp, ci)
    S = [set(fci[frp[i]:frp[i + 1]].tolist()) for i in range(n)]
    for i in range(n):
        for k in S[i]:
            if k < i:
                assert {c for c in S[k] if c > k} <= S[i]


def test_plant_pattern_has_zero_fill_in_natural_order():
    n, rows = plant_pattern(3, 4, 5)
    rp, ci = csr_from_sets(rows)
    frp, fci = O.symbolic_fill(n, rp, ci)
    assert np.array_equal(frp, rp) and np.array_equal(fci, ci)


def test_inclusive_scan_matches_cumsum():
    rng = np.random.default_rng(1)
    for m in (1, 2, 3, 1023, 1024, 1025, 70001):
        a = rng.integers(0, 50, m).astype(np.int32)
        assert np.array_equal(O.inclusive_scan(a), np.cumsum(a).astype(np.int32))


# ------------------------------------------------------ 006/007/008/009 vs ref
def _patterns_for_levels():
    rng = np.random.default_rng(3)
    out = []
    for t in range(10):
        n = int(rng.integers(8, 120))
        out.append(random_pattern(rng, n, float(rng.uniform(0.01, 0.08)), symmetric=bool(t % 2)))
    out.append(plant_pattern(2, 3, 4)[1])
    out.append(plant_pattern(4, 10, 20)[1])
    return out


@needs_ref
@pytest.mark.parametrize("case", range(12))
@pytest.mark.parametrize("kind", [0, 1, 2])
def test_deps_levels_match_reference(case, kind):
    rows = _patterns_for_levels()[case]
    n, (frp, fci) = _filled(rows)
    dp, di = O.deps(kind, n, frp, fci)
    rdp, rdi = O.deps(kind, n, frp, fci, use_ref=True)
    assert np.array_equal(dp, rdp) and np.array_equal(di, rdi)
    lv = O.levels(n, dp, di)
    rlv = O.levels(n, rdp, rdi, use_ref=True)
    assert np.array_equal(lv.start_levels, rlv.start_levels)
    assert np.array_equal(lv.start_chains, rlv.start_chains)
    assert np.array_equal(lv.chains, rlv.chains)
    if kind == 0:
        assert np.array_equal(O.max_per_level(lv, frp), O.ref_max_per_level(n, frp, fci, dp, di))


@pytest.mark.parametrize("case", range(12))
def test_levels_invariants(case):
    rows = _patterns_for_levels()[case]
    n, (frp, fci) = _filled(rows)
    for kind in (0, 1, 2):
        dp, di = O.deps(kind, n, frp, fci)
        lv = O.levels(n, dp, di)
        assert sorted(lv.chains.tolist()) == list(range(n))      # partition
        where = np.empty(n, np.int64)                             # (level, chain, position)
        lvl_of, ch_of, pos_of = {}, {}, {}
        for L in range(lv.nlev):
            for c in range(lv.start_levels[L], lv.start_levels[L + 1]):
                for t in range(lv.start_chains[c], lv.start_chains[c + 1]):
                    col = int(lv.chains[t])
                    lvl_of[col], ch_of[col], pos_of[col] = L, c, t
        for col in range(n):
            for d in di[dp[col]:dp[col + 1]]:
                d = int(d)
                # a dependency is in an earlier level, or earlier in the same chain
                assert lvl_of[d] < lvl_of[col] or (ch_of[d] == ch_of[col] and pos_of[d] < pos_of[col])
        del where


def test_plant_levels_shape():
    """Natural (postorder) numbering: one level of S*I chains of M columns,
    then one level of I bus columns (both for GLU and the L solve)."""
    I, S, M = 3, 4, 6
    n, rows = plant_pattern(I, S, M)
    rp, ci = csr_from_sets(rows)
    dp, di = O.deps(0, n, rp, ci)
    lv = O.levels(n, dp, di)
    assert lv.nlev == 2
    assert np.diff(lv.start_levels).tolist() == [I * S, I]
    assert set(np.diff(lv.start_chains[:I * S + 1]).tolist()) == {M}
    dp, di = O.deps(2, n, rp, ci)
    lu = O.levels(n, dp, di)
    assert np.diff(lu.start_levels).tolist() == [I, I * S]


# --------------------------------------------------------------- 010 CSR/CSC
@pytest.mark.parametrize("case", range(12))
def test_csr_csc_round_trip(case):
    rows = _patterns_for_levels()[case]
    n, (rp, ci) = _filled(rows)
    s = O.csr_csc(n, rp, ci)
    nnz = int(rp[n])
    row_of = np.repeat(np.arange(n), np.diff(rp))
    col_of_csc = np.repeat(np.arange(n), np.diff(s.cp))
    assert np.array_equal(s.ri[s.csr2csc], row_of)
    assert np.array_equal(col_of_csc[s.csr2csc], ci)
    assert np.array_equal(s.csc2csr[s.csr2csc], np.arange(nnz))
    for j in range(n):
        col = s.ri[s.cp[j]:s.cp[j + 1]]
        assert np.all(np.diff(col) > 0)
        below = col > j
        first = s.cp[j] + (np.argmax(below) if below.any() else len(col))
        assert s.csc_diag[j] == first
        assert s.ri[s.csc_diag[j] - 1] == j          # diagonal sits just before
        row = ci[rp[j]:rp[j + 1]]
        right = row > j
        assert s.csr_diag[j] == rp[j] + (np.argmax(right) if right.any() else len(row))


# ---------------------------------------------------------------- numerics
def _lu_case(seed, rows):
    rng = np.random.default_rng(seed)
    n, (rp, ci) = _filled(rows)
    s = O.csr_csc(n, rp, ci)
    vals_csr = dd_values(rng, n, rp, ci)
    vals = np.zeros(len(ci))
    vals[s.csr2csc] = vals_csr
    return rng, n, s, vals


@pytest.mark.parametrize("case", range(12))
def test_glu_matches_dense_lu_and_reference(case):
    rows = _patterns_for_levels()[case]
    rng, n, s, vals = _lu_case(case, rows)
    A = dense_from_csr(n, s.rp, s.ci, vals[s.csr2csc])
    lu = vals.copy()
    st, _ = O.glu_csc_sequential(s, lu)
    assert st == 0
    want = dense_lu_nopivot(A)
    got = dense_from_csr(n, s.rp, s.ci, lu[s.csr2csc])
    assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want)
    L = np.tril(got, -1) + np.eye(n)
    U = np.triu(got)
    assert np.linalg.norm(L @ U - A) <= 1e-12 * np.linalg.norm(A)
    if O.ref() is not None:
        lr = vals.copy()
        O.glu_csc_sequential(s, lr, use_ref=True)
        assert np.array_equal(lr, lu)                  # bit-exact


@pytest.mark.parametrize("case", range(12))
def test_triangular_solves(case):
    rows = _patterns_for_levels()[case]
    rng, n, s, vals = _lu_case(100 + case, rows)
    A = dense_from_csr(n, s.rp, s.ci, vals[s.csr2csc])
    lu = vals.copy()
    assert O.glu_csc_sequential(s, lu)[0] == 0
    b = rng.normal(size=n)
    x = b.copy()
    O.substitute_lu_csc(s, lu, x)
    assert np.linalg.norm(A @ x - b) <= 1e-12 * np.linalg.norm(b) * np.linalg.cond(A)
    if O.ref() is not None:
        xr = b.copy()
        O.substitute_lu_csc(s, lu, xr, use_ref=True)
        assert np.array_equal(xr, x)
    g = rng.normal(size=n)
    y = g.copy()
    O.solve_transpose_csc(s, lu, y)
    assert np.allclose(y, np.linalg.solve(A.T, g), rtol=1e-11, atol=1e-12)


def test_glu_error_codes():
    n, rows = 3, [{0, 1}, {0, 1, 2}, {1, 2}]
    rp, ci = csr_from_sets(rows)
    s = O.csr_csc(n, rp, ci)
    vals = np.ones(len(ci))
    vals[s.csr2csc[0]] = 1e-9                         # tiny pivot in column 0
    st, bad = O.glu_csc_sequential(s, vals.copy())
    assert (st, bad) == (-1, 0)
    rows = [{0}, {0, 2}, {2}]                          # column 1 empty
    rp, ci = csr_from_sets(rows)
    s = O.csr_csc(n, rp, ci)
    st, bad = O.glu_csc_sequential(s, np.ones(len(ci)))
    assert (st, bad) == (-2, 1)


# -------------------------------------------------------------- 011 update
def test_newton_update_halving():
    old = np.array([0.0, 0.0, 5.0, 2.0, -1.0])
    new = np.array([0.5, 3.0, 5.0 - 7.9, 2.0 + 1.0, -1.0 + 1.0000001])
    st, v, diff, mx = O.newton_update(old, new)
    assert st == 0
    # m = 1, 1/4, 1/8, 1, 1/2 respectively
    want = np.array([0.5, 0.75, 5.0 * (1 - 0.125) + (5.0 - 7.9) * 0.125, 3.0,
                     -1.0 * 0.5 + (-1.0 + 1.0000001) * 0.5])
    assert np.array_equal(v, want)
    assert diff == pytest.approx(np.abs(want - old).sum(), rel=1e-15)
    assert mx == np.abs(want - old).max()


def test_newton_update_errors():
    assert O.newton_update(np.array([0.0]), np.array([np.nan]))[0] == -1
    # 2^-43 * 1e20 > 1 -> damping would fall below 1e-13
    assert O.newton_update(np.array([0.0]), np.array([1e20]))[0] == -1
    # 2^-43 * 8.0e12 < 1 -> still accepted at the last admissible factor
    st, v, _, _ = O.newton_update(np.array([0.0]), np.array([8.0e12]))
    assert st == 0 and v[0] == 8.0e12 * 2.0 ** -43


# ------------------------------------------------------------ permutations
def test_permutations_round_trip():
    rng = np.random.default_rng(5)
    n = 50
    p = rng.permutation(n).astype(np.int32)
    v = rng.normal(size=n)
    w = O.apply_permutation(v, p)
    assert np.array_equal(w[p], v)                      # out[p[i]] = in[i]
    pinv = O.invert_permutation(p)
    assert np.array_equal(pinv[p], np.arange(n))
    assert np.array_equal(O.apply_permutation(w, pinv), v)
    rows = random_pattern(rng, n, 0.08)
    rp, ci = csr_from_sets(rows)
    vals = rng.normal(size=len(ci))
    A = dense_from_csr(n, rp, ci, vals)
    orp, oci, ov = O.permute_sym(n, rp, ci, vals, p)
    assert np.array_equal(dense_from_csr(n, orp, oci, ov), A[np.ix_(p, p)])
    orp, oci, ov = O.permute_rows(n, rp, ci, vals, p)
    B = np.zeros_like(A)
    B[p] = A
    assert np.array_equal(dense_from_csr(n, orp, oci, ov), B)


def test_remove_zeros_diagonal_examples():
    # row 0 lacks a diagonal and nothing fills it: swap with row 1
    st, rp, ci, v, perm = O.remove_zeros_diagonal(2, *csr_from_sets([{1}, {0, 1}]))
    assert st == 0 and perm.tolist() == [1, 0]
    assert [ci[rp[i]:rp[i + 1]].tolist() for i in range(2)] == [[0, 1], [1]]
    # row 1 lacks a diagonal but (1,0),(0,1) will fill it: untouched
    rows = [{0, 1}, {0}, {1, 2}]
    st, rp, ci, v, perm = O.remove_zeros_diagonal(3, *csr_from_sets(rows))
    assert st == 0 and perm.tolist() == [0, 1, 2]
    # MNA voltage-source shape: row 2 is the source equation (no diagonal)
    rows = [{0, 2}, {1}, {0}]
    st, rp, ci, v, perm = O.remove_zeros_diagonal(3, *csr_from_sets(rows))
    # row 0 has a diagonal; row 2: no c<2 with (2,c),(c,2) both... (2,0),(0,2) -> fills
    assert st == 0 and perm.tolist() == [0, 1, 2]
    rows = [{2}, {1}, {0, 2}]
    st, rp, ci, v, perm = O.remove_zeros_diagonal(3, *csr_from_sets(rows))
    assert st == 0 and perm.tolist() == [2, 1, 0]
    assert [ci[rp[i]:rp[i + 1]].tolist() for i in range(3)] == [[0, 2], [1], [2]]
    # (1,0) and (0,1) will fill the missing (1,1): no swap, no error
    st, *_ = O.remove_zeros_diagonal(2, *csr_from_sets([{0, 1}, {0}]))
    assert st == 0
    # row 1 lacks a diagonal and column 1 has no entry below it: error
    st, *_ = O.remove_zeros_diagonal(2, *csr_from_sets([{0}, {0}]))
    assert st == -1


# ------------------------------------------------------------- module model
def test_module_current_against_lambert_w():
    from scipy.special import lambertw
    rng = np.random.default_rng(11)
    for _ in range(200):
        Iph = rng.uniform(0.5, 9.0)
        Rsh = np.exp(rng.uniform(np.log(200), np.log(600)))
        Rs, I0, a = 0.3, 1e-10, 1.5
        V = rng.uniform(-5.0, 38.0)
        I, dIdV, dIdIph = O.module_current(V, Iph, I0, a, Rs, Rsh)
        th = (Rs * Rsh * I0) / (a * (Rs + Rsh)) * np.exp(Rsh * (Rs * (Iph + I0) + V) / (a * (Rs + Rsh)))
        W = lambertw(th).real
        Iw = (Rsh * (Iph + I0) - V) / (Rs + Rsh) - a / Rs * W
        assert I == pytest.approx(Iw, rel=1e-11, abs=1e-12)
        # implicit equation residual
        r = Iph - I0 * (np.exp((V + I * Rs) / a) - 1) - (V + I * Rs) / Rsh - I
        assert abs(r) < 1e-12
        h = 1e-6
        fd = (O.module_current(V + h, Iph, I0, a, Rs, Rsh)[0] -
              O.module_current(V - h, Iph, I0, a, Rs, Rsh)[0]) / (2 * h)
        assert dIdV == pytest.approx(fd, rel=1e-5, abs=1e-8)
        fd = (O.module_current(V, Iph + h, I0, a, Rs, Rsh)[0] -
              O.module_current(V, Iph - h, I0, a, Rs, Rsh)[0]) / (2 * h)
        assert dIdIph == pytest.approx(fd, rel=1e-5, abs=1e-8)


@needs_ref
def test_reference_amd_on_plant_has_zero_fill():
    n, rows = plant_pattern(4, 10, 20)
    rp, ci = csr_from_sets(rows)
    P = O.amd_order(n, rp, ci)
    assert sorted(P.tolist()) == list(range(n))
    orp, oci, _ = O.permute_sym(n, rp, ci, np.zeros(len(ci)), P)
    frp, fci = O.symbolic_fill(n, orp, oci)
    assert frp[-1] == orp[-1]
