"""Pins of the denoiser oracle (oracle/denoise_oracle.py) by what the mathematics fixes:
scipy's 2D correlation, the adjoint identities of pooling and upsampling, central finite
differences of the loss, and closed-form moments of the training distributions."""
import numpy as np
import pytest
import scipy.signal

import denoise_oracle as D


def test_conv3x3_matches_scipy_correlate():
    r = np.random.default_rng(0)
    x = r.normal(size=(2, 8, 12, 3))
    w = r.normal(size=(9, 3, 4))
    b = r.normal(size=4)
    out = D.conv3x3(x, w, b)
    k = w.reshape(3, 3, 3, 4)  # [ky][kx][ci][co]
    for bi in range(2):
        for co in range(4):
            ref = b[co] + sum(scipy.signal.correlate2d(x[bi, :, :, ci], k[:, :, ci, co], mode="same")
                              for ci in range(3))
            assert np.allclose(out[bi, :, :, co], ref, atol=1e-12)


def test_pool_and_up_gradients_are_adjoints():
    r = np.random.default_rng(1)
    x = r.normal(size=(2, 8, 8, 3))
    y = r.normal(size=(2, 4, 4, 3))
    assert np.isclose(np.sum(D.pool(x) * y), np.sum(x * D.pool_grad(y)))
    assert np.isclose(np.sum(D.up(y) * x), np.sum(y * D.up_grad(x)))
    w = r.normal(size=(9, 3, 5))
    g = r.normal(size=(2, 8, 8, 5))
    assert np.isclose(np.sum(D.conv3x3(x, w) * g), np.sum(x * D.conv3x3_dgrad(g, w)))


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.14159265])
    got = D.bf16_round(v)
    # ties go to even: 1 + 2^-8 -> 1, 1 + 3*2^-8 -> 1 + 4*2^-8
    assert list(got[:3]) == [1.0, 1.0, 1.0 + 4 * 2 ** -8]
    assert got[3] == 1.0 and got[4] == -2.5
    assert abs(got[5] - 3.14159265) < 2 ** -7 * 3.2


def test_backward_matches_finite_differences():
    r = np.random.default_rng(2)
    p = D.he_init(3)
    p[-17:] = r.normal(0, 0.5, 17)                 # non-trivial head
    off = 0                                        # small random biases
    for ci, co, _ in D.LAYERS:
        off += 9 * ci * co
        p[off:off + co] = r.normal(0, 0.1, co)
        off += co
    x = r.gamma(2.0, 0.5, size=(2, 8, 8))
    y = r.gamma(2.0, 0.5, size=(2, 8, 8))
    lam = (4.0, 16.0)  # the multi-scale terms too
    net = D.Net(p, bf16=False)
    loss, g = net.loss_grad(x, y, lam)
    # a sample of parameters from every layer and the head
    idx, off = [], 0
    for ci, co, _ in D.LAYERS:
        n = 9 * ci * co + co
        idx += list(r.choice(np.arange(off, off + n), 3, replace=False))
        off += n
    idx += [off + 3, off + 16]
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (D.Net(pp, bf16=False).loss_grad(x, y, lam)[0] -
              D.Net(pm, bf16=False).loss_grad(x, y, lam)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-7 + 1e-4 * abs(g[i]), (i, fd, g[i])
    assert loss > 0


@pytest.mark.parametrize("family,r4", [(0, 8.0), (1, 16.0 / 3.0), (2, 6.0),
                                       (3, 1.714985851 ** 4 * (1 - 0.36 ** 3) / (3 * 0.64))])
def test_pair_generator_moments(family, r4):
    # rms 1 per axis before scaling: <r^2> = 2; <r^4> fixes the radial profile of each family
    xs, ys = D.pair_particles((family, 1.0, 1.0, 1.0, 0.0), seed=11, k=family, which=1, n=20000)
    r2 = xs ** 2 + ys ** 2
    assert abs(xs.mean()) < 0.03 and abs(ys.mean()) < 0.03
    assert np.mean(r2) == pytest.approx(2.0, rel=0.03)
    assert np.mean(r2 ** 2) == pytest.approx(r4, rel=0.05)


def test_pair_generator_aspect_and_rotation():
    th = 0.7
    sa, sb = np.sqrt(2.0), 1 / np.sqrt(2.0)
    xs, ys = D.pair_particles((0, sa, sb, np.cos(th), np.sin(th)), seed=5, k=0, which=0, n=20000)
    cov = np.cov(np.stack([xs, ys]))
    R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    want = R @ np.diag([sa ** 2, sb ** 2]) @ R.T
    assert np.allclose(cov, want, atol=0.06)


def test_tsc_deposit_mean_and_mass():
    xs, ys = D.pair_particles((1, 1.0, 1.0, 1.0, 0.0), seed=3, k=0, which=0, n=3000)
    geo = D.square_grid(xs, ys, 16, 16)
    rho = D.tsc_deposit(xs, ys, 16, 16, geo)
    assert rho.mean() == pytest.approx(1.0, rel=1e-12)  # every particle lies in the box
    assert rho.min() >= 0.0
