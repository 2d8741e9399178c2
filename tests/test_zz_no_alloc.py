"""pvglu_obtain_fillin_noalloc (the allocation-free fill-in discovery)
against pvglu_obtain_fillin and the oracle's symbolic closure, through the
C-ABI: two calls in a row on one device, the second with a smaller input
(the file-scope device array keeps the first call's contents), bit-exact
filled patterns; the capacity and size limits fail cleanly."""
import ctypes as C

import numpy as np
import pytest

import phdthesis_albertojimenezruiz_amd as pv
from oracle import oracle as O
from tests._patterns import csr_from_sets

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

_PI = C.POINTER(C.c_int)


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda:0")


def _fn():
    f = pv.lib().pvglu_obtain_fillin_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_int, _PI, _PI, _PI, _PI, C.c_int, _PI, C.c_void_p]
    return f


def _noalloc(n, rp, ci, capacity):
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    orp = np.zeros(n + 1, np.int32)
    oci = np.zeros(max(capacity, 1), np.int32)
    nnz = C.c_int(-1)
    code = _fn()(n, rp.ctypes.data_as(_PI), ci.ctypes.data_as(_PI), orp.ctypes.data_as(_PI),
                 oci.ctypes.data_as(_PI), capacity, C.byref(nnz), None)
    return code, orp, oci[:max(nnz.value, 0)], nnz.value


def band_pattern(rng, n, half, density):
    """Random structurally symmetric band pattern with a full diagonal: its
    no-pivot closure stays inside the band but takes several passes."""
    rows = [set([i]) for i in range(n)]
    for i in range(n):
        for j in range(i + 1, min(n, i + half + 1)):
            if rng.random() < density:
                rows[i].add(j)
                rows[j].add(i)
    return csr_from_sets(rows)


def test_two_calls_match_the_allocating_entry_point(dev):
    rng = np.random.default_rng(17)
    for n, half, dens in ((20000, 24, 0.08), (5000, 12, 0.1)):     # second call smaller
        rp, ci = band_pattern(rng, n, half, dens)
        want_rp, want_ci = pv.obtain_fillin(n, rp, ci)
        assert len(want_ci) > len(ci)                                # real fill
        code, got_rp, got_ci, nnz = _noalloc(n, rp, ci, len(want_ci))
        assert code == 0, pv.lib().pvglu_last_error()
        assert nnz == len(want_ci)
        assert np.array_equal(got_rp, want_rp) and np.array_equal(got_ci, want_ci)
        orp, oci = O.symbolic_fill(n, rp, ci)
        assert np.array_equal(got_rp, orp) and np.array_equal(got_ci, oci)


def test_capacity_and_limits(dev):
    rng = np.random.default_rng(3)
    n = 3000
    rp, ci = band_pattern(rng, n, 10, 0.1)
    want_rp, want_ci = pv.obtain_fillin(n, rp, ci)
    code, _, _, nnz = _noalloc(n, rp, ci, len(want_ci) - 1)
    assert code == pv.E_ARG and nnz == len(want_ci)
    assert b"output holds" in pv.lib().pvglu_last_error()
    code, got_rp, got_ci, _ = _noalloc(n, rp, ci, nnz)
    assert code == 0 and np.array_equal(got_ci, want_ci) and np.array_equal(got_rp, want_rp)
    big = (1 << 19) + 1
    code, *_ = _noalloc(big, np.zeros(big + 1, np.int32), np.zeros(1, np.int32), 0)
    assert code == pv.E_ARG and b"maximum" in pv.lib().pvglu_last_error()
