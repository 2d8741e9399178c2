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
