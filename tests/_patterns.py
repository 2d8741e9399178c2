"""Pattern generators shared by the tests (test-side, independent of the
product's plant builder).

``plant_pattern`` lays out the config plant topology with the node numbering
the product uses: inverter i, string s, module node k (1..M) at
``i*(S*M+1) + s*M + (k-1)``, the inverter's DC bus at ``i*(S*M+1) + S*M``.
Module k sits between node k-1 and node k (node 0 is ground); node M is the
string's combiner, tied to the bus by a cable; the bus has a Norton
conductance to ground.
"""
import numpy as np


def csr_from_sets(rows):
    rp = [0]
    ci = []
    for r in rows:
        ci.extend(sorted(r))
        rp.append(len(ci))
    return np.array(rp, np.int32), np.array(ci, np.int32)


def plant_pattern(I, S, M):
    n = I * (S * M + 1)
    rows = [set([i]) for i in range(n)]
    for inv in range(I):
        base = inv * (S * M + 1)
        bus = base + S * M
        for s in range(S):
            for k in range(1, M + 1):
                node = base + s * M + (k - 1)
                if k > 1:
                    rows[node].add(node - 1)
                    rows[node - 1].add(node)
            comb = base + s * M + (M - 1)
            rows[comb].add(bus)
            rows[bus].add(comb)
    return n, rows


def hub_pattern(lengths, tail):
    """A hub node (a DC bus) tied to the top node of len(lengths) strings
    numbered before it (each from its leaf up), then a tail string numbered
    after it from its top down.  In natural order: zero fill, the hub
    column's updates all on its diagonal (one per string before it), the
    strings' columns tridiagonal steps.  Returns (n, row sets)."""
    rows, tops = [], []

    def node():
        rows.append(set([len(rows)]))
        return len(rows) - 1

    def link(a, c):
        rows[a].add(c)
        rows[c].add(a)

    for L in lengths:
        prev = None
        for _ in range(L):
            k = node()
            if prev is not None:
                link(prev, k)
            prev = k
        tops.append(prev)
    hub = node()
    for t in tops:
        link(t, hub)
    prev = hub
    for _ in range(tail):
        k = node()
        link(prev, k)
        prev = k
    return len(rows), rows


# (string lengths before the hub, tail length): star columns with 1, 2, 23
# and 32 updates; 33 updates (over the star register budget) and unequal
# strings (operands not strided) take the generic gather
HUB_CASES = [([5], 0), ([5, 5], 7), ([4] * 23, 27), ([3] * 32, 6), ([3] * 33, 6),
             ([3, 5, 4, 6], 5)]
HUB_STAR = [True, True, True, True, False, False]


def random_pattern(rng, n, density, symmetric=False, diag=True):
    rows = [set() for _ in range(n)]
    m = rng.random((n, n)) < density
    if symmetric:
        m = m | m.T
    for i in range(n):
        for j in np.nonzero(m[i])[0]:
            rows[i].add(int(j))
        if diag:
            rows[i].add(i)
    return rows


def dense_from_csr(n, rp, ci, vals):
    A = np.zeros((n, n))
    for i in range(n):
        for p in range(rp[i], rp[i + 1]):
            A[i, ci[p]] = vals[p]
    return A


def dense_lu_nopivot(A):
    """Textbook Doolittle elimination without pivoting (dense, rank-1 form)."""
    A = A.copy()
    n = A.shape[0]
    for k in range(n):
        A[k + 1:, k] /= A[k, k]
        A[k + 1:, k + 1:] -= np.outer(A[k + 1:, k], A[k, k + 1:])
    return A


def dd_values(rng, n, rp, ci):
    """Random values on a pattern, diagonally dominant by rows."""
    vals = rng.uniform(-1.0, 1.0, len(ci))
    for i in range(n):
        s = 0.0
        d = -1
        for p in range(rp[i], rp[i + 1]):
            if ci[p] == i:
                d = p
            else:
                s += abs(vals[p])
        if d >= 0:
            vals[d] = (s + 1.0) * (1 if rng.random() < 0.5 else -1)
    return vals


def plant_values(rng, I, S, M, B=None):
    """MNA values on ``plant_pattern(I, S, M)``: per module a conductance
    (diode companion, 1e-3..5 S), per cable 5..50 S, per bus a 1e4 S Norton
    conductance, and a right-hand side.  Returns (n, rp, ci, vals, b) with
    CSR values; with ``B`` the values are [nnz][B] and b is [n][B]
    (independent draws per instance)."""
    n, rows = plant_pattern(I, S, M)
    rp, ci = csr_from_sets(rows)
    where = {}
    for i in range(n):
        for p in range(rp[i], rp[i + 1]):
            where[(i, int(ci[p]))] = p
    nb = 1 if B is None else B
    nm = I * S * M
    node = np.arange(nm) + np.arange(nm) // (S * M)
    first = (np.arange(nm) % M) == 0
    prev = np.where(first, -1, node - 1)
    comb = np.array([inv * (S * M + 1) + s * M + M - 1 for inv in range(I) for s in range(S)])
    bus = np.array([inv * (S * M + 1) + S * M for inv in range(I) for s in range(S)])
    buses = np.array([inv * (S * M + 1) + S * M for inv in range(I)])
    gm = np.exp(rng.uniform(np.log(1e-3), np.log(5.0), (nm, nb)))
    gc = np.exp(rng.uniform(np.log(5.0), np.log(50.0), (len(comb), nb)))
    ieq = rng.uniform(-9.0, 9.0, (nm, nb))
    vals = np.zeros((len(ci), nb))
    bvec = np.zeros((n, nb))
    idx = lambda a, b: np.array([where[(int(x), int(y))] for x, y in zip(a, b)], dtype=np.int64)
    np.add.at(vals, idx(node, node), gm)
    has = ~first
    np.add.at(vals, idx(prev[has], prev[has]), gm[has])
    np.add.at(vals, idx(node[has], prev[has]), -gm[has])
    np.add.at(vals, idx(prev[has], node[has]), -gm[has])
    np.add.at(vals, idx(comb, comb), gc)
    np.add.at(vals, idx(bus, bus), gc)
    np.add.at(vals, idx(comb, bus), -gc)
    np.add.at(vals, idx(bus, comb), -gc)
    np.add.at(vals, idx(buses, buses), 1e4)
    np.add.at(bvec, node, ieq)
    np.add.at(bvec, prev[has], -ieq[has])
    bvec[buses] += 1e4 * 600.0
    if B is None:
        return n, rp, ci, vals[:, 0], bvec[:, 0]
    return n, rp, ci, vals, bvec
