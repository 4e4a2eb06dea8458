"""ob_bunch_moments_noalloc (the moments with the library's fixed device scratch instead of
a per-call allocation) against ob_bunch_moments, through the C ABI: the same 28 sums bit
for bit, on two calls in a row with the second bunch smaller, with and without
per-particle macrosizes, after queued maps, and for an empty bunch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sums(fn, b):
    import _obnative as N
    out = np.full(28, np.nan)
    N.call(fn, b._h, N.dptr(out))
    return out


def test_noalloc_moments_equal_allocating_moments():
    import teapot_base as TPB
    from workloads import gaussian_coords, make_bunch
    big = make_bunch(gaussian_coords(3_000_001, seed=71))
    small = make_bunch(gaussian_coords(1000, seed=72))
    small.setMacroSizes(np.random.default_rng(3).uniform(0.5, 1.5, 1000))
    for b in (big, small, big):  # the scratch keeps the last call's contents
        TPB.drift(b, 0.25)      # queued map: both entry points apply it first
        want = _sums("ob_bunch_moments", b)
        got = _sums("ob_bunch_moments_noalloc", b)
        assert np.array_equal(got, want)
        assert got[0] == pytest.approx(b.getSize() if b is big else b.getMacroSizes().sum(),
                                       rel=1e-12)
    from bunch import Bunch
    empty = Bunch()
    assert np.array_equal(_sums("ob_bunch_moments_noalloc", empty), np.zeros(28))


def test_noalloc_moments_refuse_a_batched_bunch():
    from batch import BunchBatch
    from _obnative import OrbitError
    from workloads import gaussian_coords, make_bunch
    bunches = [make_bunch(gaussian_coords(2000, seed=s)) for s in (1, 2)]
    bb = BunchBatch(bunches)
    with pytest.raises(OrbitError, match="belongs to a batch"):
        _sums("ob_bunch_moments_noalloc", bunches[0])
    del bb
