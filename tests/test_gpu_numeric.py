"""GPU parity of the HIP path against the CPU oracle (through the C-ABI).

Bar: the refactor, both substitutions and the transposed solves follow the
sequential CPU GLU operation order entry by entry, so results are compared
bit-exactly (np.array_equal) against oracle/liboracle.so, which is itself
bit-exact against the reference's own GLU_CSC_CPU_sequential /
Substitute_LU_Vector_b_CSC_CPU (tests/test_oracle.py).  Pattern work (fill,
scan) is integer and bit-exact by nature.
"""
import numpy as np
import pytest

import phdthesis_albertojimenezruiz_amd as pv
from oracle import oracle as O
from tests._patterns import (HUB_CASES, csr_from_sets, dd_values, hub_pattern,
                             plant_pattern, plant_values, random_pattern)

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda:0")


def _to_dev(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _oracle_lu_solve(s, vals_csc, b):
    lu = vals_csc.copy()
    st, bad = O.glu_csc_sequential(s, lu)
    assert st == 0, (st, bad)
    x = b.copy()
    O.substitute_lu_csc(s, lu, x)
    return lu, x


def _random_case(seed, n=None):
    rng = np.random.default_rng(seed)
    n = n or int(rng.integers(20, 160))
    rows = random_pattern(rng, n, float(rng.uniform(0.01, 0.05)), symmetric=bool(seed % 2))
    rp, ci = csr_from_sets(rows)
    frp, fci = O.symbolic_fill(n, rp, ci)
    return rng, n, frp, fci


SCHEDS = [(pv.SCHED_LEVEL, 1), (pv.SCHED_LEVEL, 64), (pv.SCHED_LEVEL, 16),
          (pv.SCHED_COMPONENT, 1), (pv.SCHED_COMPONENT, 8), (pv.SCHED_COMPONENT, 64),
          (pv.SCHED_HYBRID, 1), (pv.SCHED_HYBRID, 4), (pv.SCHED_HYBRID, 64)]


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("sched", SCHEDS)
def test_refactor_solve_random_patterns_bitexact(dev, seed, sched):
    rng, n, rp, ci = _random_case(seed)
    plan = pv.Plan(n, rp, ci)
    plan.set_schedule(*sched)
    s = O.csr_csc(n, rp, ci)
    B = 70                                     # a ragged last instance tile
    A = np.zeros((len(ci), B))
    bb = np.zeros((n, B))
    want_lu = np.zeros_like(A)
    want_x = np.zeros_like(bb)
    for k in range(B):
        v = dd_values(rng, n, rp, ci)
        A[:, k] = v[s.csc2csr]
        bb[:, k] = rng.normal(size=n)
        want_lu[:, k], want_x[:, k] = _oracle_lu_solve(s, A[:, k].copy(), bb[:, k].copy())
    dA, db = _to_dev(A, dev), _to_dev(bb, dev)
    dlu, dx = torch.empty_like(dA), torch.empty_like(db)
    plan.status_reset()
    plan.refactor_solve(dA, dlu, db, dx, B)
    torch.cuda.synchronize()
    assert plan.status()[0] == 0
    assert np.array_equal(dlu.cpu().numpy(), want_lu)
    assert np.array_equal(dx.cpu().numpy(), want_x)
    # separate refactor + solve, in place (a == lu, x == b)
    dA2, db2 = dA.clone(), db.clone()
    plan.refactor(dA2, dA2, B)
    plan.solve(dA2, db2, db2, B)
    torch.cuda.synchronize()
    assert np.array_equal(dA2.cpu().numpy(), want_lu)
    assert np.array_equal(db2.cpu().numpy(), want_x)


@pytest.mark.parametrize("sched", SCHEDS)
def test_solve_transpose_bitexact(dev, sched):
    rng, n, rp, ci = _random_case(9, n=120)
    plan = pv.Plan(n, rp, ci)
    plan.set_schedule(*sched)
    s = O.csr_csc(n, rp, ci)
    B = 5
    LU = np.zeros((len(ci), B))
    G = rng.normal(size=(n, B))
    want = np.zeros_like(G)
    for k in range(B):
        v = dd_values(rng, n, rp, ci)[s.csc2csr]
        st, _ = O.glu_csc_sequential(s, v)
        assert st == 0
        LU[:, k] = v
        g = G[:, k].copy()
        O.solve_transpose_csc(s, v, g)
        want[:, k] = g
    dlu, dg = _to_dev(LU, dev), _to_dev(G, dev)
    dy = torch.empty_like(dg)
    plan.solve_transpose(dlu, dg, dy, B)
    torch.cuda.synchronize()
    assert np.array_equal(dy.cpu().numpy(), want)


def _multi_component_case(seed):
    """Block-diagonal pattern (6 random blocks of unequal size, one of them a
    single column) under a random symmetric permutation, so independent
    components interleave in column order; filled by the oracle."""
    rng = np.random.default_rng(seed)
    sizes = [1] + [int(v) for v in rng.integers(5, 60, 5)]
    n = sum(sizes)
    rows, off = [set() for _ in range(n)], 0
    for m in sizes:
        blk = random_pattern(rng, m, float(rng.uniform(0.05, 0.2)), symmetric=True)
        for i, r in enumerate(blk):
            rows[off + i] = {off + j for j in r}
        off += m
    perm = rng.permutation(n)
    prow = [set() for _ in range(n)]
    for i, r in enumerate(rows):
        prow[perm[i]] = {int(perm[j]) for j in r}
    rp, ci = csr_from_sets(prow)
    frp, fci = O.symbolic_fill(n, rp, ci)
    return rng, n, frp, fci


@pytest.mark.parametrize("variant", [pv.SCHED_LEVEL, pv.SCHED_HYBRID])
@pytest.mark.parametrize("groups", [2, 4])
@pytest.mark.parametrize("width", [1, 64])
def test_component_groups_bitexact(dev, variant, groups, width):
    """Component groups on side streams (fork/join by events): refactor+solve,
    refactor then solve, and the transposed solve against the oracle."""
    rng, n, rp, ci = _multi_component_case(11)
    plan = pv.Plan(n, rp, ci)
    assert plan.info("ncomp") >= 6
    plan.set_schedule(variant, width)
    plan.set_stream_groups(groups)
    assert plan.get_stream_groups() == groups
    s = O.csr_csc(n, rp, ci)
    B = 70
    A = np.stack([dd_values(rng, n, rp, ci)[s.csc2csr] for _ in range(B)], axis=1)
    bb = rng.normal(size=(n, B))
    G = rng.normal(size=(n, B))
    dA, db, dg = _to_dev(A, dev), _to_dev(bb, dev), _to_dev(G, dev)
    dlu, dx, dy = torch.empty_like(dA), torch.empty_like(db), torch.empty_like(dg)
    stream = torch.cuda.Stream()
    stream.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(stream):
        plan.refactor_solve(dA, dlu, db, dx, B, stream)
        plan.solve_transpose(dlu, dg, dy, B, stream)
    stream.synchronize()
    lu, x, y = dlu.cpu().numpy(), dx.cpu().numpy(), dy.cpu().numpy()
    for k in (0, 33, B - 1):
        wlu, wx = _oracle_lu_solve(s, A[:, k].copy(), bb[:, k].copy())
        assert np.array_equal(lu[:, k], wlu) and np.array_equal(x[:, k], wx)
        g = G[:, k].copy()
        O.solve_transpose_csc(s, wlu, g)
        assert np.array_equal(y[:, k], g)
    dlu2, dx2 = torch.empty_like(dA), torch.empty_like(db)
    plan.refactor(dA, dlu2, B)
    plan.solve(dlu2, db, dx2, B)
    torch.cuda.synchronize()
    assert torch.equal(dlu2, dlu) and torch.equal(dx2, dx)


# (variant, width, component groups on side streams)
PLANT_SCHEDS = [(pv.SCHED_AUTO, 0, 1), (pv.SCHED_LEVEL, 0, 1), (pv.SCHED_COMPONENT, 0, 1),
                (pv.SCHED_HYBRID, 0, 1), (pv.SCHED_LEVEL, 0, 2), (pv.SCHED_HYBRID, 0, 2),
                (pv.SCHED_HYBRID, 0, 4)]


@pytest.mark.parametrize("cfg", [(4, 10, 20), (40, 24, 28)])
@pytest.mark.parametrize("B", [1, 64, 130])
@pytest.mark.parametrize("sched", PLANT_SCHEDS)
def test_plant_refactor_solve(dev, cfg, B, sched):
    """Config plants (cfg0 4x10x20, cfg1 40x24x28) in the reference pipeline's
    AMD ordering when oracle/_ref is built (natural order otherwise), under
    every schedule (the DC bus chains mix register runs with generic columns;
    HYBRID merges the narrow bus/substitution levels into one launch)."""
    I, S, M = cfg
    rng = np.random.default_rng(I * 100 + B)
    n, rp0, ci0 = plant_pattern(I, S, M)[0], *csr_from_sets(plant_pattern(I, S, M)[1])
    P = O.amd_order(n, rp0, ci0) if O.ref() is not None else np.arange(n, dtype=np.int32)
    rp, ci, _ = O.permute_sym(n, rp0, ci0, np.zeros(len(ci0)), P)
    plan = pv.Plan(n, rp, ci)
    plan.set_schedule(*sched[:2])
    plan.set_stream_groups(sched[2])
    s = O.csr_csc(n, rp, ci)
    Pinv = O.invert_permutation(P)
    _, _, _, V, bv = plant_values(rng, I, S, M, B)
    # symmetric permutation of the values: new(Pinv[i], Pinv[j]) = old(i, j)
    order = O.permute_sym(n, rp0, ci0, np.arange(len(ci0), dtype=float), P)[2].astype(np.int64)
    A = np.ascontiguousarray(V[order][s.csc2csr])
    bb = np.zeros((n, B))
    bb[Pinv] = bv
    check = sorted(set([0, B - 1, B // 2]))
    want = {k: _oracle_lu_solve(s, A[:, k].copy(), bb[:, k].copy()) for k in check}
    dA, db = _to_dev(A, dev), _to_dev(bb, dev)
    dlu, dx = torch.empty_like(dA), torch.empty_like(db)
    plan.status_reset()
    plan.refactor_solve(dA, dlu, db, dx, B)
    torch.cuda.synchronize()
    assert plan.status()[0] == 0
    lu, x = dlu.cpu().numpy(), dx.cpu().numpy()
    for k in check:
        assert np.array_equal(lu[:, k], want[k][0])
        assert np.array_equal(x[:, k], want[k][1])
    # the drop-in sequence: refactor alone (register chains without the fused
    # forward substitution), then the two substitution sweeps
    dlu2, dx2 = torch.empty_like(dA), torch.empty_like(db)
    plan.refactor(dA, dlu2, B)
    plan.solve(dlu2, db, dx2, B)
    torch.cuda.synchronize()
    assert plan.status()[0] == 0
    assert torch.equal(dlu2, dlu) and torch.equal(dx2, dx)


def _hub_case(lengths, tail):
    n, rows = hub_pattern(lengths, tail)
    rp, ci = csr_from_sets(rows)
    frp, fci = O.symbolic_fill(n, rp, ci)
    assert len(fci) == len(ci)                    # a tree in this order: no fill
    return n, frp, fci


@pytest.mark.parametrize("case", range(len(HUB_CASES)))
@pytest.mark.parametrize("sched", [(pv.SCHED_HYBRID, 64), (pv.SCHED_HYBRID, 1),
                                   (pv.SCHED_LEVEL, 64), (pv.SCHED_COMPONENT, 64)])
@pytest.mark.parametrize("B", [1, 70])
def test_star_and_tridiagonal_paths_bitexact(dev, case, sched, B):
    """Star columns with 1, 2, 23, 32 updates (33: over the register budget,
    generic gather; unequal strings: operands not strided, generic gather),
    the tridiagonal segments before and after them, fused and unfused
    refactor, both sweeps and the transposed sweeps, against the oracle."""
    n, rp, ci = _hub_case(*HUB_CASES[case])
    rng = np.random.default_rng(100 + case)
    plan = pv.Plan(n, rp, ci)
    plan.set_schedule(*sched)
    s = O.csr_csc(n, rp, ci)
    A = np.stack([dd_values(rng, n, rp, ci)[s.csc2csr] for _ in range(B)], axis=1)
    bb = rng.normal(size=(n, B))
    want_lu, want_x, want_t = np.zeros_like(A), np.zeros_like(bb), np.zeros_like(bb)
    for k in range(B):
        want_lu[:, k], want_x[:, k] = _oracle_lu_solve(s, A[:, k].copy(), bb[:, k].copy())
        g = bb[:, k].copy()
        O.solve_transpose_csc(s, want_lu[:, k].copy(), g)
        want_t[:, k] = g
    dA, db = _to_dev(A, dev), _to_dev(bb, dev)
    dlu, dx, dy = torch.empty_like(dA), torch.empty_like(db), torch.empty_like(db)
    plan.status_reset()
    plan.refactor_solve(dA, dlu, db, dx, B)
    plan.solve_transpose(dlu, db, dy, B)
    torch.cuda.synchronize()
    assert plan.status()[0] == 0
    assert np.array_equal(dlu.cpu().numpy(), want_lu)
    assert np.array_equal(dx.cpu().numpy(), want_x)
    assert np.array_equal(dy.cpu().numpy(), want_t)
    dlu2, dx2 = torch.empty_like(dA), torch.empty_like(db)
    plan.refactor(dA, dlu2, B)
    plan.solve(dlu2, db, dx2, B)
    torch.cuda.synchronize()
    assert torch.equal(dlu2, dlu) and torch.equal(dx2, dx)


@pytest.mark.parametrize("sched", [(pv.SCHED_HYBRID, 64), (pv.SCHED_LEVEL, 64),
                                   (pv.SCHED_COMPONENT, 16), (pv.SCHED_LEVEL, 1)])
def test_singular_pivot_in_a_string_segment(dev, sched):
    """Two pivots below 1e-8 in one string (columns 3 and 4 of the first
    string, instance 37 of 70): the first is reported, whichever kernel path
    (tridiagonal segment, generic gather) the schedule takes."""
    n, rp, ci = _hub_case([6, 6, 6], 5)
    rng = np.random.default_rng(8)
    plan = pv.Plan(n, rp, ci)
    plan.set_schedule(*sched)
    s = O.csr_csc(n, rp, ci)
    B = 70
    A = np.stack([dd_values(rng, n, rp, ci)[s.csc2csr] for _ in range(B)], axis=1)
    for c in (3, 4):
        q0, qd = s.cp[c], s.csc_diag[c] - 1            # U entry (c-1, c), pivot
        assert qd == q0 + 1
        A[q0, 37] = 0.0                                  # pivot = its A value
        A[qd, 37] = 1e-9
    dA = _to_dev(A, dev)
    plan.status_reset()
    plan.refactor(dA, dA, B)
    torch.cuda.synchronize()
    assert plan.status() == (pv.E_SINGULAR, 3, 37)
    db = _to_dev(rng.normal(size=(n, B)), dev)
    dA = _to_dev(A, dev)
    plan.status_reset()
    plan.refactor_solve(dA, torch.empty_like(dA), db, torch.empty_like(db), B)
    torch.cuda.synchronize()
    assert plan.status() == (pv.E_SINGULAR, 3, 37)


@pytest.mark.parametrize("sched", [(pv.SCHED_HYBRID, 64), (pv.SCHED_LEVEL, 64),
                                   (pv.SCHED_HYBRID, 32)])
def test_streaming_cache_policy_bitexact(dev, sched):
    """The width-32/64 kernels with the streaming cache policy (non-temporal
    hints, chosen by default only for batches far beyond the Infinity Cache)
    forced on a 130-instance batch of the 40x24x28 plant: refactor, both
    sweeps and the transposed sweeps bit-exact vs the oracle and equal to
    the default policy's results."""
    I, S, M = 40, 24, 28
    rng = np.random.default_rng(31)
    n, rows = plant_pattern(I, S, M)
    rp0, ci0 = csr_from_sets(rows)
    P = O.amd_order(n, rp0, ci0) if O.ref() is not None else np.arange(n, dtype=np.int32)
    rp, ci, _ = O.permute_sym(n, rp0, ci0, np.zeros(len(ci0)), P)
    s = O.csr_csc(n, rp, ci)
    B = 130
    _, _, _, V, bv = plant_values(rng, I, S, M, B)
    order = O.permute_sym(n, rp0, ci0, np.arange(len(ci0), dtype=float), P)[2].astype(np.int64)
    A = np.ascontiguousarray(V[order][s.csc2csr])
    bb = np.zeros((n, B))
    bb[O.invert_permutation(P)] = bv
    out = {}
    for policy in (pv.CACHE_STREAMING, pv.CACHE_DEFAULT):
        plan = pv.Plan(n, rp, ci)
        plan.set_schedule(*sched)
        plan.set_cache_policy(policy)
        dA, db = _to_dev(A, dev), _to_dev(bb, dev)
        dlu, dx, dy = torch.empty_like(dA), torch.empty_like(db), torch.empty_like(db)
        plan.status_reset()
        plan.refactor_solve(dA, dlu, db, dx, B)
        plan.solve_transpose(dlu, db, dy, B)
        dlu2, dx2 = torch.empty_like(dA), torch.empty_like(db)
        plan.refactor(dA, dlu2, B)
        plan.solve(dlu2, db, dx2, B)
        torch.cuda.synchronize()
        assert plan.status()[0] == 0
        assert torch.equal(dlu2, dlu) and torch.equal(dx2, dx)
        out[policy] = (dlu.cpu().numpy(), dx.cpu().numpy(), dy.cpu().numpy())
    lu, x, y = out[pv.CACHE_STREAMING]
    for a, b_ in zip(out[pv.CACHE_STREAMING], out[pv.CACHE_DEFAULT]):
        assert np.array_equal(a, b_)
    for k in (0, 65, B - 1):
        wlu, wx = _oracle_lu_solve(s, A[:, k].copy(), bb[:, k].copy())
        assert np.array_equal(lu[:, k], wlu) and np.array_equal(x[:, k], wx)
        g = bb[:, k].copy()
        O.solve_transpose_csc(s, wlu.copy(), g)
        assert np.array_equal(y[:, k], g)


def test_autotune_choice_is_bitexact_and_cached(dev):
    """pvglu_autotune times every (schedule, width, groups, cache policy) on
    scratch values, keeps the fastest, and caches it per (pattern, batch);
    the tuned launch stays bit-exact vs the oracle."""
    I, S, M = 4, 10, 20
    rng = np.random.default_rng(5)
    n, rows = plant_pattern(I, S, M)
    rp, ci = csr_from_sets(rows)
    s = O.csr_csc(n, rp, ci)
    B = 70
    _, _, _, V, bv = plant_values(rng, I, S, M, B)
    A = np.ascontiguousarray(V[s.csc2csr])
    plan = pv.Plan(n, rp, ci)
    best = plan.autotune(B, reps=2)
    assert 0.0 < best < 100.0
    choice = plan.get_schedule()
    assert choice[0] in (pv.SCHED_LEVEL, pv.SCHED_COMPONENT, pv.SCHED_HYBRID)
    assert choice[1] in (1, 2, 4, 8, 16, 32, 64)
    plan2 = pv.Plan(n, rp, ci)                         # same pattern: cached
    plan2.autotune(B, reps=2, cached=True)
    assert plan2.get_schedule() == choice
    dA, db = _to_dev(A, dev), _to_dev(bv, dev)
    dlu, dx = torch.empty_like(dA), torch.empty_like(db)
    plan.status_reset()
    plan.refactor_solve(dA, dlu, db, dx, B)
    torch.cuda.synchronize()
    assert plan.status()[0] == 0
    lu, x = dlu.cpu().numpy(), dx.cpu().numpy()
    for k in (0, 33, B - 1):
        wlu, wx = _oracle_lu_solve(s, A[:, k].copy(), bv[:, k].copy())
        assert np.array_equal(lu[:, k], wlu) and np.array_equal(x[:, k], wx)


def test_singular_pivot_is_reported(dev):
    rng, n, rp, ci = _random_case(3, n=40)
    plan = pv.Plan(n, rp, ci)
    s = O.csr_csc(n, rp, ci)
    B = 3
    A = np.stack([dd_values(rng, n, rp, ci)[s.csc2csr] for _ in range(B)], axis=1)
    A[s.csc_diag[0] - 1, 2] = 1e-9                       # instance 2, column 0
    dA = _to_dev(A, dev)
    plan.status_reset()
    plan.refactor(dA, dA, B)
    code, col, inst = plan.status()
    assert (code, col, inst) == (pv.E_SINGULAR, 0, 2)
    plan.status_reset()
    assert plan.status()[0] == 0


@pytest.mark.parametrize("n,ne,B", [(50, 80, 33), (40, 1000, 1), (40, 700, 32),
                                    (30, 90, 300)])
def test_newton_update_matches_oracle(dev, n, ne, B):
    # instance tiles of 64 (B=33: half-empty), 1, 32 and 256 + a partial one;
    # element counts that end inside a block
    rng = np.random.default_rng(4)
    pos = rng.integers(-1, n, ne).astype(np.int32)
    neg = rng.integers(-1, n, ne).astype(np.int32)
    X = rng.normal(scale=3.0, size=(n, B))
    V = rng.normal(scale=3.0, size=(ne, B))
    plan = pv.Plan(1, np.array([0, 1], np.int32), np.array([0], np.int32))
    dpos, dneg = _to_dev(pos, dev), _to_dev(neg, dev)
    dX, dV = _to_dev(X, dev), _to_dev(V, dev)
    dsum = torch.zeros(B, dtype=torch.float64, device=dev)
    dmax = torch.zeros(B, dtype=torch.float64, device=dev)
    plan.status_reset()
    plan.newton_update(ne, B, dpos, dneg, dX, dV, dsum, dmax)
    torch.cuda.synchronize()
    assert plan.status()[0] == 0
    got = dV.cpu().numpy()
    for k in range(B):
        vn = np.where(pos >= 0, X[np.maximum(pos, 0), k], 0.0) - \
            np.where(neg >= 0, X[np.maximum(neg, 0), k], 0.0)
        st, v, diff, mx = O.newton_update(V[:, k], vn)
        assert st == 0
        assert np.array_equal(got[:, k], v)
        assert dmax[k].item() == mx
        assert dsum[k].item() == pytest.approx(diff, rel=1e-13)


def test_newton_update_reports_nonconvergence(dev):
    plan = pv.Plan(1, np.array([0, 1], np.int32), np.array([0], np.int32))
    pos = torch.tensor([0], dtype=torch.int32, device=dev)
    neg = torch.tensor([-1], dtype=torch.int32, device=dev)
    X = torch.tensor([[float("nan")]], dtype=torch.float64, device=dev)
    V = torch.zeros((1, 1), dtype=torch.float64, device=dev)
    dsum = torch.zeros(1, dtype=torch.float64, device=dev)
    dmax = torch.zeros(1, dtype=torch.float64, device=dev)
    plan.status_reset()
    plan.newton_update(1, 1, pos, neg, X, V, dsum, dmax)
    assert plan.status()[0] == pv.E_NOCONV
    assert V.item() == 0.0


@pytest.mark.parametrize("seed", range(6))
def test_obtain_fillin_matches_oracle(dev, seed):
    rng = np.random.default_rng(50 + seed)
    n = int(rng.integers(30, 400))
    rows = random_pattern(rng, n, float(rng.uniform(0.002, 0.03)), symmetric=bool(seed % 2))
    rp, ci = csr_from_sets(rows)
    frp, fci = pv.obtain_fillin(n, rp, ci)
    orp, oci = O.symbolic_fill(n, rp, ci)
    assert np.array_equal(frp, orp) and np.array_equal(fci, oci)


def test_obtain_fillin_plant_zero_fill_and_wide_window(dev):
    n, rows = plant_pattern(40, 24, 28)
    rp, ci = csr_from_sets(rows)
    frp, fci = pv.obtain_fillin(n, rp, ci)
    assert np.array_equal(frp, rp) and np.array_equal(fci, ci)
    # arrow pattern wider than one 32768-column window: last row/col dense
    n = 70000
    rows = [{i, n - 1} for i in range(n)]
    rows[n - 1] = set(range(0, n, 7)) | {5, n - 1}
    rows[5].add(n - 2)            # (5, n-2) with (n-1,5) -> fills (n-1, n-2)
    rp, ci = csr_from_sets(rows)
    frp, fci = pv.obtain_fillin(n, rp, ci)
    orp, oci = O.symbolic_fill(n, rp, ci)
    assert np.array_equal(frp, orp) and np.array_equal(fci, oci)


@pytest.mark.parametrize("m", [1, 5, 4096, 4097, 270001])
def test_device_inclusive_scan(dev, m):
    rng = np.random.default_rng(m)
    a = rng.integers(0, 100, m).astype(np.int32)
    d = _to_dev(a, dev)
    pv.inclusive_scan_device(d)
    torch.cuda.synchronize()
    assert np.array_equal(d.cpu().numpy(), np.cumsum(a).astype(np.int32))


# (inverters, batch) per group; the smallest batch decides the path: below
# 16 the component-style work list, 16..31 HYBRID steps at width 16, else at
# width 64; a random pattern next to a plant has another step sequence and
# takes the work list whatever the batch
MULTI_CASES = [((1, 7), (2, 20), (5, 3)), ((1, 17), (2, 20), (5, 30)),
               ((1, 40), (2, 33), (5, 64)), ((2, 20), ("random", 24))]


@pytest.mark.parametrize("case", range(len(MULTI_CASES)))
def test_multi_pattern_batch(dev, case):
    """Mixed plant sizes (cfg3 shape, scaled: 1/2/5 inverters x 24 x 28) in
    one launch per step, forward and transposed, bit-exact per instance."""
    rng = np.random.default_rng(77)
    groups = []
    for I, B in MULTI_CASES[case]:
        if I == "random":
            _, n, rp, ci = _random_case(5, n=90)
            s = O.csr_csc(n, rp, ci)
            A = np.stack([dd_values(rng, n, rp, ci)[s.csc2csr] for _ in range(B)], axis=1)
            bv = rng.normal(size=(n, B))
            groups.append((pv.Plan(n, rp, ci), s, A, bv, B))
            continue
        n, rows = plant_pattern(I, 24, 28)
        rp, ci = csr_from_sets(rows)
        plan = pv.Plan(n, rp, ci)
        s = O.csr_csc(n, rp, ci)
        _, _, _, V, bv = plant_values(rng, I, 24, 28, B)
        A = np.ascontiguousarray(V[s.csc2csr])
        groups.append((plan, s, A, np.ascontiguousarray(bv), B))
    m = pv.Multi([g[0] for g in groups], [g[4] for g in groups])
    dA = [_to_dev(g[2], dev) for g in groups]
    db = [_to_dev(g[3], dev) for g in groups]
    dlu = [torch.empty_like(t) for t in dA]
    dx = [torch.empty_like(t) for t in db]
    m.refactor_solve(dA, dlu, db, dx)
    dy = [torch.empty_like(t) for t in db]
    m.solve_transpose(dlu, db, dy)
    torch.cuda.synchronize()
    for (plan, s, A, bv, B), lu_d, x_d, y_d in zip(groups, dlu, dx, dy):
        lu_h, x_h, y_h = lu_d.cpu().numpy(), x_d.cpu().numpy(), y_d.cpu().numpy()
        for k in range(B):
            lu, x = _oracle_lu_solve(s, A[:, k].copy(), bv[:, k].copy())
            assert np.array_equal(lu_h[:, k], lu) and np.array_equal(x_h[:, k], x)
            y = bv[:, k].copy()
            O.solve_transpose_csc(s, lu, y)
            assert np.array_equal(y_h[:, k], y)
