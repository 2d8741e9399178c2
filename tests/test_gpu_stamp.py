"""Stamping on load (pvglu_pv_set_stamp_on_load): the batched Newton's
refactor forms the series-string entries of the companion matrix from the
module conductances instead of reading a stamped A.  The results must be
bit-identical with the switch off (every entry stamped by k_stamp_a): node
voltages, module voltages, power and per-instance iteration counts, on the
LEVEL and HYBRID schedules at several widths, with and without the homotopy
ladder (whose virtual conductance the stamped values carry), and with the
batch compaction on or off.  Prints the time per replay both ways.
"""
import numpy as np
import pytest

import phdthesis_albertojimenezruiz_amd as pv

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def plant40():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return pv.PV(40, 24, 28, seed=0)


def _solve(plant, G, ladder, stamp, skip):
    plant.set_stamp_on_load(stamp)
    plant.set_skip_frozen(skip)
    x, v, p, stats = plant.newton_batched(G, tol=1e-12, ladder=ladder)
    its, last = plant.newton_info(G.shape[1])
    return x, v, p, stats, its, last


@pytest.mark.parametrize("sched,width,ladder,skip", [
    (pv.SCHED_HYBRID, 64, False, True),
    (pv.SCHED_HYBRID, 64, True, True),
    (pv.SCHED_HYBRID, 16, False, True),
    (pv.SCHED_LEVEL, 32, True, False),
])
def test_stamp_on_load_bit_identical(plant40, sched, width, ladder, skip):
    B = 160
    G = np.stack([plant40.irradiance(s, 0.0) for s in range(B)], axis=1)
    plant40.plan().set_schedule(sched, width)
    try:
        a = _solve(plant40, G, ladder, False, skip)
        b = _solve(plant40, G, ladder, True, skip)
    finally:
        plant40.set_stamp_on_load(True)
        plant40.set_skip_frozen(True)
        plant40.plan().set_schedule(pv.SCHED_AUTO, 0)
    print(f"sched {sched}:{width} ladder={ladder} skip={skip}: {b[3][0]:.0f} replays, "
          f"ms per replay {a[3][2]:.3f} (A stamped) -> {b[3][2]:.3f} (stamped on load)")
    for u, w in zip(a[:3] + a[4:], b[:3] + b[4:]):
        assert np.array_equal(u, w)
