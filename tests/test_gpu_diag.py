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


# ---- extrema, Grid1D binning, computeBunchMoments, diagnostics nodes, profiles ----------

def _oracle_bunch(c, w=None, macrosize=1.0):
    import oracle as O
    return O.Bunch(c, macrosize=macrosize, kin_energy=1.0, msize=w)


@pytest.mark.parametrize("n", [1, 777, 1_000_003])
def test_extrema_match_oracle_exactly(n):
    import teapot_base as TPB
    from orbit_utils import BunchExtremaCalculator
    from workloads import make_bunch
    c = _coords(n, seed=n + 3)
    b = make_bunch(c)
    TPB.drift(b, 0.7)  # queued: applied before the scan
    calc = BunchExtremaCalculator()
    xyz, pde = calc.extremaXYZ(b), calc.extremaXpYpdE(b)
    ob = _oracle_bunch(c)
    ob.apply("drift", 0.7)
    e = ob.extrema6()
    assert xyz == tuple(e[[0, 1, 4, 5, 8, 9]]) and pde == tuple(e[[2, 3, 6, 7, 10, 11]])
    assert calc.extremaZ(b) == (e[8], e[9])
    assert calc.extremaR(b) == pytest.approx(ob.extrema_r(), rel=1e-14)


def test_extrema_of_empty_bunch():
    from bunch import Bunch
    from orbit_utils import BunchExtremaCalculator
    big = np.finfo(float).max
    assert BunchExtremaCalculator().extremaXYZ(Bunch()) == (big, -big) * 3
    assert BunchExtremaCalculator().extremaR(Bunch()) == 0.0


@pytest.mark.parametrize("nz", [1, 2, 100, 4096, 4097])
@pytest.mark.parametrize("mode", ["macrosize", "attribute", "particle"])
def test_grid1d_bin_bunch_matches_oracle(nz, mode):
    import spacecharge
    from workloads import make_bunch
    n = 300_000
    c = _coords(n, seed=nz)
    b = make_bunch(c, intensity=1e10)
    w = None
    if mode == "attribute":
        w = np.random.default_rng(nz).uniform(0.5, 2.0, n)
        b.setMacroSizes(w)
    ob = _oracle_bunch(c, w, macrosize=1e10 / n)
    for axis in range(6):
        lo, hi = np.percentile(c[axis], [3, 96])  # particles outside are skipped
        g = spacecharge.Grid1D(nz, lo, hi)
        g.setValue(5.0, nz - 1)  # binning adds to what the grid holds
        if mode == "particle":
            g.binBunchByParticle(b, axis)
        else:
            g.binBunch(b, axis)
        want = ob.bin_axis(nz, lo, hi, axis, by_particle=(mode == "particle"))
        want[nz - 1] += 5.0
        got = g.array()
        # the oracle adds up to n values into one bin in sequence: ~n eps relative
        assert np.all(np.abs(got - want) <= 1e-10 * np.abs(want) + 1e-300), (axis, nz, mode)


def _moment_scale(c, w, order, prm):
    """sum |terms| / total weight for every entry: the bound summation order is relative to."""
    x_avg, y_avg, D, em, b2, xb, yb, disp = prm
    dx = np.abs((c[0] - (D * c[5] / em / b2 if disp > 0 else 0.0)) - x_avg)
    dy = np.abs(c[2] - y_avg)
    m = np.ones(c.shape[1]) if w is None else w
    s = np.zeros((order + 1, order + 1))
    for j in range(order):
        for i in range(order + 1 - j):
            t = (m ** (i + j) if w is not None else 1.0) * dx ** i * dy ** j
            t = t / (xb ** i * yb ** j) + (t if w is not None else 0.0)
            s[i, j] = t.sum() / m.sum()
    if w is None and order >= 2:
        s[2, order] = np.sum(dx * dx * dy ** order)
    return s


@pytest.mark.parametrize("n,order,weights,disp,norm", [
    (100_000, 4, False, -1, -1), (100_000, 6, True, 1, 1), (1_000_001, 12, False, 1, -1),
    (50_000, 1, True, -1, 1), (50_000, 2, False, 1, 1), (1, 3, False, -1, -1)])
def test_compute_bunch_moments_matches_oracle(n, order, weights, disp, norm):
    import teapot_base as TPB
    from bunch import BunchTwissAnalysis
    from workloads import make_bunch
    c = _coords(n, seed=order)
    c[0] *= 1e3  # mm-scale coordinates keep high powers inside FP64's range
    c[2] *= 1e3
    b = make_bunch(c, kin_energy=1.3)
    w = np.random.default_rng(3).uniform(0.5, 1.5, n) if weights else None
    if weights:
        b.setMacroSizes(w)
    TPB.drift(b, 0.2)
    tw = BunchTwissAnalysis()
    tw.computeBunchMoments(b, order, disp, norm)
    after = b.getCoordinates()
    sp = b.getSyncParticle()
    em = sp.kinEnergy() + sp.mass()
    if norm > 0:
        xb = np.sqrt(tw.getBeta(0) * tw.getEmittance(0))
        yb = np.sqrt(tw.getBeta(1) * tw.getEmittance(1))
    else:
        xb, yb = np.sqrt(tw.getBeta(0)), np.sqrt(tw.getBeta(1))
    prm = (tw.getAverage(0), tw.getAverage(2), tw.getDispersion(0), em, sp.beta() ** 2, xb, yb,
           disp)
    ob = _oracle_bunch(after, w)
    want = ob.compute_moments(order, disp, *prm[:7])
    got = np.array([[tw.getBunchMoment(i, j) for j in range(order + 1)]
                    for i in range(order + 1)])
    tol = 1e-11 * _moment_scale(after, w, order, prm) + 1e-300
    tol[0, 0] = tol[1, 0] = tol[0, 1] = 0.0
    ok = (np.abs(got - want) <= tol) | (np.isnan(got) & np.isnan(want))  # n = 1: beta is 0/0
    assert np.all(ok), (got[~ok], want[~ok])
    assert tw.getBunchMoment(order + 1, 0) == 0.0 and tw.getBunchMoment(-1, 0) == 0.0


def test_compute_bunch_moments_refuses_order_zero():
    from _obnative import OrbitError
    from bunch import BunchTwissAnalysis
    from workloads import make_bunch
    with pytest.raises(OrbitError):
        BunchTwissAnalysis().computeBunchMoments(make_bunch(_coords(10, 1)), 0, 1, 1)


def _diag_lattice():
    from orbit.teapot import TEAPOT_Lattice, DriftTEAPOT, QuadTEAPOT
    lat = TEAPOT_Lattice()
    d1, q, d2 = DriftTEAPOT("d1"), QuadTEAPOT("q"), DriftTEAPOT("d2")
    d1.setLength(1.0)
    q.setLength(0.4)
    q.addParam("kq", 0.6)
    d2.setLength(2.0)
    for nd in (d1, q, d2):
        lat.addNode(nd)
    lat.initialize()
    return lat, q


def test_diagnostics_nodes_in_a_lattice(tmp_path):
    """StatLats at 0.5 m (in the first drift), Moments at 2.4 m (in the second), a BPM at
    the quad's entrance (after the quad's entrance fringe, as the reference orders entrance
    children): each sees the bunch as the oracle has it at that point."""
    from orbit.diagnostics import (TeapotStatLatsNode, TeapotMomentsNode, TeapotBPMSignalNode,
                                   addTeapotDiagnosticsNode, addTeapotDiagnosticsNodeAsChild)
    from workloads import make_bunch
    lat, q = _diag_lattice()
    st = TeapotStatLatsNode(str(tmp_path / "stat.dat"))
    mo = TeapotMomentsNode(str(tmp_path / "mom.dat"), 3)
    bpm = TeapotBPMSignalNode()
    addTeapotDiagnosticsNode(lat, 0.5, st)
    addTeapotDiagnosticsNode(lat, 2.4, mo)
    addTeapotDiagnosticsNodeAsChild(lat, q, bpm)
    assert [nd.getName() for nd in lat.getNodes()] == \
        ["d1", "statlats no name", "d1", "q", "d2", "moments no name", "d2"]
    c = _coords(50_000, seed=77)
    b = make_bunch(c)
    lat.trackBunch(b)
    st.closeStatLats()
    mo.closeMoments()
    sp = b.getSyncParticle()
    ob = _oracle_bunch(c)
    ob.apply("drift", 0.5)
    t1 = T.Twiss(ob.coords, p=sp.momentum(), beta=sp.beta(), gamma=sp.gamma())
    row = [float(v) for v in open(tmp_path / "stat.dat").read().split()]
    assert len(row) == 10 and row[0] == 0.5
    want = [t1.emittance(0), t1.emittance(1), t1.beta_fn(0), t1.beta_fn(1), t1.alpha(0),
            t1.alpha(1), t1.dispersion(0), t1.dispersion_prime(0)]
    np.testing.assert_allclose(row[2:], want, rtol=1e-8, atol=1e-14)
    ob.apply("drift", 0.5)
    ob.apply("quadfringeIN", 0.6)  # entrance children run after the quad's own fringe
    xb_yb = T.Twiss(ob.coords, p=sp.momentum(), beta=sp.beta(), gamma=sp.gamma())
    assert bpm.getSignal() == pytest.approx((xb_yb.avg[0], xb_yb.avg[2]), rel=1e-10, abs=1e-16)
    lines = open(tmp_path / "mom.dat").read().splitlines()
    assert len(lines) == 1 and lines[0].endswith("\t")
    vals = [float(v) for v in lines[0].split()]
    assert vals[0] == 2.4 and len(vals) == 2 + 10
    assert vals[2] == 1.0  # <x^0 y^0>


def test_profiles_match_oracle(tmp_path):
    from orbit.diagnostics import profiles
    from orbit.utils import OrbitFinalizeError
    from workloads import make_bunch
    c = _coords(200_000, seed=4)
    b = make_bunch(c, intensity=2e12)
    ob = _oracle_bunch(c, macrosize=2e12 / 200_000)
    e = ob.extrema6()
    g = profiles(b, "px", str(tmp_path / "h.dat"), steps=64)
    want = ob.bin_axis(64, e[2], e[3], 1)
    assert (g.getMinZ(), g.getMaxZ()) == (e[2], e[3])
    assert np.all(np.abs(g.array() - want) <= 1e-12 * want + 1e-300)
    lines = open(tmp_path / "h.dat").read().splitlines()
    assert len(lines) == 65 and lines[0].startswith("%  Min, Max =")
    assert float(lines[0].split("hist_sum =")[1]) == pytest.approx(want.sum(), rel=1e-7)
    z, v = (float(t) for t in lines[10].split())
    assert z == pytest.approx(g.getGridZ(9), rel=1e-7) and v == pytest.approx(want[9], rel=1e-7)
    g2 = profiles(b, "dE", str(tmp_path / "h2.dat"), steps=10, Min=-1e-3, Max=1e-3)
    assert np.all(np.abs(g2.array() - ob.bin_axis(10, -1e-3, 1e-3, 5)) <=
                  1e-12 * ob.bin_axis(10, -1e-3, 1e-3, 5) + 1e-300)
    with pytest.raises(OrbitFinalizeError):
        profiles(b, "t", str(tmp_path / "h3.dat"))
    one = make_bunch(c[:, :1])
    profiles(one, "x", str(tmp_path / "h4.dat"))
    assert open(tmp_path / "h4.dat").read().startswith("bunch has only number of particles =1")
