"""On-disk plan format (pvglu_plan_save / pvglu_plan_load), CPU only."""
import numpy as np
import pytest

import phdthesis_albertojimenezruiz_amd as pv
from oracle import oracle as O
from tests._patterns import csr_from_sets, plant_pattern, random_pattern


def _plant_plan():
    n, rows = plant_pattern(4, 10, 20)
    rp, ci = csr_from_sets(rows)
    P, Prow = pv.reorder(n, rp, ci)
    prp, pci, _ = O.permute_sym(n, rp, ci, np.zeros(len(ci)), P)
    frp, fci = O.symbolic_fill(n, prp, pci)
    return pv.Plan(n, frp, fci), P, Prow


def test_round_trip_keeps_every_array_and_the_ordering(tmp_path):
    plan, P, Prow = _plant_plan()
    f = tmp_path / "plant.pvglu"
    plan.save(f, P, Prow)
    back, P2, Prow2 = pv.Plan.load(f)
    for k in pv.ARRAYS:
        assert np.array_equal(back.array(k), plan.array(k)), k
    for k in pv.INFO:
        assert back.info(k) == plan.info(k), k
    assert np.array_equal(P2, P) and np.array_equal(Prow2, Prow)


def test_ordering_defaults_to_identity(tmp_path):
    rng = np.random.default_rng(4)
    n = 60
    rp, ci = csr_from_sets(random_pattern(rng, n, 0.05))
    frp, fci = O.symbolic_fill(n, rp, ci)
    plan = pv.Plan(n, frp, fci)
    f = tmp_path / "r.pvglu"
    plan.save(f)
    _, P, Prow = pv.Plan.load(f)
    assert np.array_equal(P, np.arange(n)) and np.array_equal(Prow, np.arange(n))


def test_corruption_is_detected(tmp_path):
    plan, P, Prow = _plant_plan()
    f = tmp_path / "p.pvglu"
    plan.save(f, P, Prow)
    raw = bytearray(f.read_bytes())
    raw[100] ^= 1
    bad = tmp_path / "bad.pvglu"
    bad.write_bytes(bytes(raw))
    with pytest.raises(pv.PvgluError, match="checksum"):
        pv.Plan.load(bad)
    trunc = tmp_path / "trunc.pvglu"
    trunc.write_bytes(f.read_bytes()[:40])
    with pytest.raises(pv.PvgluError):
        pv.Plan.load(trunc)
    junk = tmp_path / "junk.bin"
    junk.write_bytes(b"x" * 64)
    with pytest.raises(pv.PvgluError, match="not a plan file"):
        pv.Plan.load(junk)
    with pytest.raises(pv.PvgluError, match="cannot open"):
        pv.Plan.load(tmp_path / "missing.pvglu")


def _fnv(data):
    h = 1469598103934665603
    for c in data:
        h ^= c
        h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def test_arrays_that_disagree_with_the_pattern_are_rejected(tmp_path):
    """A well-formed file whose level arrays were altered (checksum redone)
    fails: loading rebuilds the plan from the pattern and compares."""
    plan, P, Prow = _plant_plan()
    f = tmp_path / "p.pvglu"
    plan.save(f, P, Prow)
    raw = bytearray(f.read_bytes()[:-8])
    words = np.frombuffer(bytes(raw[24:]), dtype=np.int32).copy()
    # walk the sections to the LU chains array and swap two of its entries
    pos, target = 0, pv.ARRAYS["lu_chains"]
    while True:
        sid, ln = words[pos], words[pos + 1]
        if sid == target:
            words[pos + 2], words[pos + 3] = words[pos + 3], words[pos + 2]
            break
        pos += 2 + ln
    raw[24:] = words.tobytes()
    raw += _fnv(bytes(raw)).to_bytes(8, "little")
    bad = tmp_path / "alt.pvglu"
    bad.write_bytes(bytes(raw))
    with pytest.raises(pv.PvgluError, match="differs"):
        pv.Plan.load(bad)


def _rewrite_section(raw_file, sid_target, edit):
    """Applies edit(words_of_section) to one section and redoes the checksum."""
    raw = bytearray(raw_file[:-8])
    words = np.frombuffer(bytes(raw[24:]), dtype=np.int32).copy()
    pos = 0
    while True:
        sid, ln = words[pos], words[pos + 1]
        if sid == sid_target:
            edit(words[pos + 2:pos + 2 + ln])
            break
        pos += 2 + ln
    raw[24:] = words.tobytes()
    raw += _fnv(bytes(raw)).to_bytes(8, "little")
    return bytes(raw)


@pytest.mark.parametrize("sid,edit,match", [
    (100, lambda w: w.__setitem__(0, w[1]), "P of .* is not a permutation"),     # duplicate
    (100, lambda w: w.__setitem__(3, len(w)), "P of .* is not a permutation"),   # out of range
    (101, lambda w: w.__setitem__(slice(0, 3), [1, 2, 0]), "not an involution"),  # 3-cycle
    (101, lambda w: w.__setitem__(0, -1), "Prow of .* is not a permutation"),
])
def test_crafted_orderings_are_rejected(tmp_path, sid, edit, match):
    """The checksum catches accidents only: a file whose ordering section was
    altered and re-checksummed fails on the ordering's own checks (P a
    permutation of [0, n); Prow a permutation that is its own inverse)."""
    plan, P, Prow = _plant_plan()
    f = tmp_path / "p.pvglu"
    plan.save(f, P, Prow)
    bad = tmp_path / "crafted.pvglu"
    bad.write_bytes(_rewrite_section(f.read_bytes(), sid, edit))
    with pytest.raises(pv.PvgluError, match=match):
        pv.Plan.load(bad)
    # the unaltered file still loads
    pv.Plan.load(f)
