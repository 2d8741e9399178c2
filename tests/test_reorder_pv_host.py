"""Host-side ordering and plant harness (CPU).

* pvglu_amd_order (the library's AMD 1.1 restatement) against the reference's
  own AMD compiled in oracle/_ref: bit-exact permutations.
* pvglu_reorder's row swaps against the oracle's 004b restatement.
* the config plant generator: sizes, pattern, element distributions, and
  the module model against the oracle's implicit-equation solve.
"""
import numpy as np
import pytest

import phdthesis_albertojimenezruiz_amd as pv
from oracle import oracle as O
from tests._patterns import csr_from_sets, plant_pattern, random_pattern

needs_ref = pytest.mark.skipif(O.ref() is None, reason="oracle/_ref not built")


@needs_ref
@pytest.mark.parametrize("cfg", [(1, 2, 3), (4, 10, 20), (40, 24, 28), (10, 24, 28)])
def test_amd_matches_reference_on_plants(cfg):
    n, rows = plant_pattern(*cfg)
    rp, ci = csr_from_sets(rows)
    assert np.array_equal(pv.amd_order(n, rp, ci), O.amd_order(n, rp, ci))


@needs_ref
@pytest.mark.parametrize("seed", range(40))
def test_amd_matches_reference_on_random(seed):
    rng = np.random.default_rng(1000 + seed)
    n = int(rng.integers(1, 400))
    rows = random_pattern(rng, n, float(rng.uniform(0.001, 0.3)),
                          symmetric=bool(seed % 3 == 0), diag=bool(seed % 2))
    # a dense row and column to exercise the dense-row path (> 10 sqrt(n))
    if seed % 5 == 0 and n > 20:
        rows[0] = set(range(n))
        for r in rows:
            r.add(0)
    rp, ci = csr_from_sets(rows)
    assert np.array_equal(pv.amd_order(n, rp, ci), O.amd_order(n, rp, ci))


@pytest.mark.parametrize("seed", range(10))
def test_reorder_row_swaps_match_oracle(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(3, 60))
    rows = random_pattern(rng, n, 0.1, diag=False)
    for i in range(n):            # keep every column nonempty below some row
        rows[(i + 1) % n].add(i)
    rp, ci = csr_from_sets(rows)
    try:
        P, Prow = pv.reorder(n, rp, ci)
    except pv.PvgluError as e:
        assert e.code == pv.E_ARG
        Pr = pv.amd_order(n, rp, ci)
        orp, oci, _ = O.permute_sym(n, rp, ci, np.zeros(len(ci)), Pr)
        assert O.remove_zeros_diagonal(n, orp, oci)[0] == -1
        return
    assert np.array_equal(P, pv.amd_order(n, rp, ci))
    orp, oci, _ = O.permute_sym(n, rp, ci, np.zeros(len(ci)), P)
    st, _, _, _, perm = O.remove_zeros_diagonal(n, orp, oci)
    assert st == 0 and np.array_equal(Prow, perm)


@pytest.mark.parametrize("cfg", [(4, 10, 20), (40, 24, 28), (1, 1, 1)])
def test_plant_generator(cfg):
    I, S, M = cfg
    plant = pv.PV(I, S, M, seed=0)
    assert plant.n == I * (S * M + 1)
    assert plant.nmod == I * S * M and plant.ncab == I * S
    rp, ci = plant.pattern()
    n, rows = plant_pattern(I, S, M)
    erp, eci = csr_from_sets(rows)
    assert np.array_equal(rp, erp) and np.array_equal(ci, eci)
    pos, neg, rsh = plant.modules()
    a, b, g = plant.cables()
    assert rsh.min() >= 200 and rsh.max() <= 600 and g.min() >= 5 and g.max() <= 50
    assert np.array_equal(plant.buses(), [k * (S * M + 1) + S * M for k in range(I)])
    # module k of a string: + at node k, - at node k-1 (ground for k = 1)
    assert np.all(pos == np.arange(plant.nmod) + np.arange(plant.nmod) // (S * M))
    first = (np.arange(plant.nmod) % M) == 0
    assert np.all(neg[first] == -1) and np.all(neg[~first] == pos[~first] - 1)


def test_plant_distributions_are_log_uniform():
    plant = pv.PV(40, 24, 28, seed=0)
    _, _, rsh = plant.modules()
    u = (np.log(rsh) - np.log(200)) / (np.log(600) - np.log(200))
    assert abs(u.mean() - 0.5) < 0.01 and abs(u.var() - 1 / 12) < 0.005
    _, _, g = plant.cables()
    u = (np.log(g) - np.log(5)) / (np.log(50) - np.log(5))
    assert abs(u.mean() - 0.5) < 0.05
    # same seed -> same plant; other seed -> other draws
    assert np.array_equal(pv.PV(40, 24, 28, seed=0).modules()[2], rsh)
    assert not np.array_equal(pv.PV(40, 24, 28, seed=1).modules()[2], rsh)


def test_cloud_field():
    plant = pv.PV(40, 24, 28, seed=0)
    assert np.all(plant.irradiance(-1) == 1000.0)
    g0 = plant.irradiance(0, 0.0)
    g1 = plant.irradiance(0, 1.0)
    assert g0.min() >= 0.0 and g0.max() <= 1000.0 and g0.min() < 900.0
    assert 0 < np.abs(g1 - g0).max() < 200.0          # drift of 5..15 m over 1 s
    assert not np.array_equal(plant.irradiance(1, 0.0), g0)
    x, y = plant.positions()
    assert x.min() > 0 and x.max() < 1000 and y.max() < 1000


def test_module_current_matches_oracle_implicit_solve():
    rng = np.random.default_rng(3)
    for _ in range(300):
        V = rng.uniform(-50.0, 60.0)
        Iph = rng.uniform(0.0, 9.0)
        rsh = np.exp(rng.uniform(np.log(200), np.log(600)))
        I, dV, dI = pv.module_current(V, Iph, rsh)
        oI, odV, odI = O.module_current(V, Iph, 1e-10, 1.5, 0.3, rsh)
        assert I == pytest.approx(oI, rel=1e-12, abs=1e-12)
        assert dV == pytest.approx(odV, rel=1e-10, abs=1e-14)
        assert dI == pytest.approx(odI, rel=1e-10, abs=1e-14)


def test_module_current_against_mpmath_lambert_w():
    """The library's explicit module current (single-precision start, one
    double-precision Fritsch-Shafer-Crowley step) against the same formula
    evaluated with a 30-digit Lambert W, over ln(theta) from about -60 (the
    small-argument series) to about 2600."""
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 30
    Rs, I0, a = mp.mpf("0.3"), mp.mpf("1e-10"), mp.mpf("1.5")
    rng = np.random.default_rng(5)
    Vs = np.concatenate([rng.uniform(-60.0, 60.0, 300), rng.uniform(60.0, 4000.0, 60)])
    for V in Vs:
        Iph = rng.uniform(0.0, 9.0)
        rsh = float(np.exp(rng.uniform(np.log(200), np.log(600))))
        I, dV, _ = pv.module_current(V, Iph, rsh)
        R, Vm, Ip = mp.mpf(rsh), mp.mpf(float(V)), mp.mpf(float(Iph))
        lnth = mp.log(Rs * R * I0 / (a * (Rs + R))) + R * (Rs * (Ip + I0) + Vm) / (a * (Rs + R))
        W = mp.lambertw(mp.e ** lnth).real
        Ie = (R * (Ip + I0) - Vm) / (Rs + R) - a / Rs * W
        dVe = -1 / (Rs + R) - R / (Rs * (Rs + R)) * W / (1 + W)
        # the two terms of I, each rounded at its own magnitude
        scale = 1.0 + abs(float((R * (Ip + I0) - Vm) / (Rs + R))) + abs(float(a / Rs * W))
        assert abs(I - float(Ie)) <= 2e-14 * scale, (V, Iph, rsh)
        assert abs(dV - float(dVe)) <= 1e-14 * (1.0 + abs(float(dVe))), (V, Iph, rsh)
