"""BunchTwissAnalysis on the device (ob_bunch_moments) against the oracle on the same
coordinates: the 28 sums to FP64 summation-order accuracy, then every getter."""
import ctypes as C

import numpy as np
import pytest

import twiss_oracle as T

pytestmark = pytest.mark.gpu


def _coords(n, seed):
    r = np.random.default_rng(seed)
    c = np.empty((6, n))
    c[0] = r.normal(1e-3, 3e-3, n)
    c[1] = 0.2 * c[0] + r.normal(0, 5e-4, n)
    c[2] = r.normal(-2e-3, 2e-3, n)
    c[3] = -0.1 * c[2] + r.normal(0, 4e-4, n)
    c[4] = r.uniform(-80, 80, n)
    c[5] = r.normal(0, 1e-3, n)
    c[0] += 1.5 * c[5]  # dispersion
    return c


def _device_sums(b):
    import _obnative as N
    b._push_attrs()
    s = np.zeros(28)
    N.call("ob_bunch_moments", b._h, N.dptr(s))
    return s


def _scale(c, w):
    """sum |w c_a c_b| for each of the 28 sums: the bound FP64 rounding is relative to."""
    return T.sums(np.abs(c), np.abs(w) if w is not None else None)


def _check_getters(tw, ref):
    for ic in range(3):
        assert tw.getEmittance(ic) == pytest.approx(ref.emittance(ic), rel=1e-9)
        assert tw.getAlpha(ic) == pytest.approx(ref.alpha(ic), rel=1e-8, abs=1e-12)
        assert tw.getBeta(ic) == pytest.approx(ref.beta_fn(ic), rel=1e-9)
        assert tw.getGamma(ic) == pytest.approx(ref.gamma_fn(ic), rel=1e-9)
        e, a, be, g = ref.effective(ic)
        assert tw.getEffectiveEmittance(ic) == pytest.approx(e, rel=1e-9)
        assert tw.getEffectiveAlpha(ic) == pytest.approx(a, rel=1e-8, abs=1e-12)
        assert tw.getEffectiveBeta(ic) == pytest.approx(be, rel=1e-9)
        assert tw.getEffectiveGamma(ic) == pytest.approx(g, rel=1e-9)
    for ic in range(2):
        assert tw.getDispersion(ic) == pytest.approx(ref.dispersion(ic), rel=1e-8, abs=1e-12)
        assert tw.getDispersionDerivative(ic) == pytest.approx(ref.dispersion_prime(ic),
                                                               rel=1e-8, abs=1e-12)
    for i in range(6):
        assert tw.getAverage(i) == pytest.approx(ref.avg[i], rel=1e-10, abs=1e-15)
        for j in range(6):
            assert tw.getCorrelation(i, j) == pytest.approx(ref.correlation(i, j), rel=1e-8,
                                                            abs=1e-15)


@pytest.mark.parametrize("n,weights", [(1, False), (1000, False), (200_000, False),
                                       (200_000, True), (3_000_001, True)])
def test_sums_match_oracle(n, weights):
    from workloads import make_bunch
    c = _coords(n, seed=n)
    b = make_bunch(c)
    w = None
    if weights:
        w = np.random.default_rng(1).uniform(0.5, 2.0, n)
        b.setMacroSizes(w)
    s = _device_sums(b)
    ref = T.sums(c, w)
    assert np.all(np.abs(s - ref) <= 1e-12 * _scale(c, w) + 1e-300), np.abs(s - ref).max()


def test_analysis_after_queued_maps_matches_oracle():
    """Maps queued on the bunch are applied before the sums are taken."""
    import teapot_base as TPB
    from bunch import BunchTwissAnalysis
    from workloads import make_bunch
    c = _coords(100_000, seed=5)
    b = make_bunch(c, kin_energy=1.0)
    TPB.drift(b, 0.8)
    TPB.quad1(b, 0.3, 0.6)
    TPB.drift(b, 0.4)
    tw = BunchTwissAnalysis()
    tw.analyzeBunch(b)
    after = b.getCoordinates()
    sp = b.getSyncParticle()
    ref = T.Twiss(after, p=sp.momentum(), beta=sp.beta(), gamma=sp.gamma())
    assert tw.getGlobalCount() == 100_000
    assert tw.getGlobalMacrosize() == pytest.approx(100_000.0, rel=0, abs=0)
    _check_getters(tw, ref)
    assert tw.getEmittanceNormalized(0) == pytest.approx(
        ref.emittance(0) * sp.gamma() * sp.beta(), rel=1e-9)


def test_analysis_weights_and_deleted_particles():
    """Deleted particles are removed (the analysis compresses first); macrosize weights."""
    from bunch import BunchTwissAnalysis
    from workloads import make_bunch
    c = _coords(20_000, seed=9)
    b = make_bunch(c)
    w = np.random.default_rng(2).uniform(0.1, 3.0, c.shape[1])
    b.setMacroSizes(w)
    dead = np.arange(0, c.shape[1], 7)
    for i in dead:
        b.deleteParticleFast(int(i))
    keep = np.setdiff1d(np.arange(c.shape[1]), dead)
    tw = BunchTwissAnalysis()
    tw.analyzeBunch(b)
    ref = T.Twiss(c[:, keep], w[keep], p=b.getSyncParticle().momentum(),
                  beta=b.getSyncParticle().beta())
    assert tw.getGlobalCount() == keep.size
    assert tw.getGlobalMacrosize() == pytest.approx(w[keep].sum(), rel=1e-12)
    _check_getters(tw, ref)


def test_empty_bunch():
    from bunch import Bunch, BunchTwissAnalysis
    b = Bunch()
    tw = BunchTwissAnalysis()
    tw.analyzeBunch(b)
    assert tw.getGlobalCount() == 0
    assert tw.getGlobalMacrosize() == 0.0
    assert tw.getAverage(0) == 0.0 and tw.getCorrelation(0, 0) == 0.0
