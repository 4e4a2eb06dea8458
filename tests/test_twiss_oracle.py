"""Pins the BunchTwissAnalysis oracle by closed forms: particles built through the
Courant-Snyder normalization from a base set with exactly zero mean, unit variances and no
correlations must return the (alpha, beta, emittance, dispersion) they were built with."""
import numpy as np
import pytest

import twiss_oracle as T


def _base(m):
    """4^m points in (u, v, a, b) from the product of {-1, +1} sets: every coordinate has
    mean 0 and variance 1, every pair is uncorrelated, exactly."""
    pts = np.array(np.meshgrid(*[[-1.0, 1.0]] * 4, indexing="ij")).reshape(4, -1)
    return np.tile(pts, (1, m))


def _build(alpha, beta, eps, ay, by, ey, D, Dp, sdE, p, bb):
    u, v, a, b = _base(3)
    n = u.size
    c = np.zeros((6, n))
    # dE is uncorrelated with (u, v, a, b): +-sdE alternating over a doubled set
    c = np.concatenate([c, c], axis=1)
    dE = np.concatenate([np.full(n, sdE), np.full(n, -sdE)])
    u, v, a, b = (np.concatenate([q, q]) for q in (u, v, a, b))
    k = 1.0 / (p * bb)  # x_disp = D * dE / (p beta), the form getDispersion inverts
    c[0] = np.sqrt(eps * beta) * u + D * k * dE
    c[1] = np.sqrt(eps / beta) * (v - alpha * u) + Dp * k * dE
    c[2] = np.sqrt(ey * by) * a
    c[3] = np.sqrt(ey / by) * (b - ay * a)
    c[4] = 0.3 * u  # z correlated with x only; the z plane reads (z, dE)
    c[5] = dE
    return c


def test_courant_snyder_and_dispersion_recovered():
    p, bb = 1.696, 0.875
    c = _build(alpha=-1.3, beta=7.5, eps=2e-6, ay=0.4, by=3.2, ey=1e-6, D=2.1, Dp=-0.15,
               sdE=1e-3, p=p, bb=bb)
    t = T.Twiss(c, p=p, beta=bb, gamma=1.0)
    assert t.emittance(0) == pytest.approx(2e-6, rel=1e-12)
    assert t.alpha(0) == pytest.approx(-1.3, rel=1e-12)
    assert t.beta_fn(0) == pytest.approx(7.5, rel=1e-12)
    assert t.gamma_fn(0) == pytest.approx((1 + 1.3 ** 2) / 7.5, rel=1e-12)
    assert t.dispersion(0) == pytest.approx(2.1, rel=1e-12)
    assert t.dispersion_prime(0) == pytest.approx(-0.15, rel=1e-12)
    assert t.emittance(1) == pytest.approx(1e-6, rel=1e-12)
    assert t.alpha(1) == pytest.approx(0.4, rel=1e-12)
    assert t.beta_fn(1) == pytest.approx(3.2, rel=1e-12)
    assert t.dispersion(1) == pytest.approx(0.0, abs=1e-12)
    # effective values include the dispersive spread: <x^2> = eps beta + (D k sdE)^2
    ke = 2.1 / (p * bb) * 1e-3
    kp = -0.15 / (p * bb) * 1e-3
    x2 = 2e-6 * 7.5 + ke ** 2
    xp2 = 2e-6 * (1 + 1.3 ** 2) / 7.5 + kp ** 2
    xxp = 2e-6 * 1.3 + ke * kp
    e = np.sqrt(x2 * xp2 - xxp ** 2)
    assert t.effective(0)[0] == pytest.approx(e, rel=1e-12)
    assert t.effective(0)[2] == pytest.approx(x2 / e, rel=1e-12)
    # z plane: <z^2> = 0.09, <dE^2> = sdE^2, <z dE> = 0 -> emittance 0.3 sdE
    assert t.emittance(2) == pytest.approx(0.3e-3, rel=1e-12)
    assert t.avg == pytest.approx(np.zeros(6), abs=1e-15)


def test_weights_equal_replicated_particles():
    r = np.random.default_rng(3)
    c = r.normal(size=(6, 50))
    w = r.integers(1, 4, 50).astype(float)
    rep = np.repeat(c, w.astype(int), axis=1)
    a, b = T.sums(c, w), T.sums(rep)
    assert np.allclose(a, b, rtol=1e-13, atol=1e-12)
    ta, tb = T.Twiss(c, w), T.Twiss(rep)
    for ic in range(3):
        assert ta.emittance(ic) == pytest.approx(tb.emittance(ic), rel=1e-12)


def test_zero_total_keeps_raw_sums():
    c = np.ones((6, 4))
    t = T.Twiss(c, w=np.array([1.0, -1.0, 2.0, -2.0]))
    assert t.total == 0.0
    assert np.all(t.avg == 0.0)
