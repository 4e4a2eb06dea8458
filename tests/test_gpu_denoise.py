"""The learned denoiser on the device against the numpy oracle (forward, loss, every
parameter gradient), the device pair generator against the oracle's restatement, and a
short training run that must beat the plain low-N deposit on held-out pairs."""
import numpy as np
import pytest

import denoise_oracle as D
from helpers import rel_l2

pytestmark = pytest.mark.gpu


def _params(seed):
    r = np.random.default_rng(seed)
    p = D.he_init(seed)
    off = 0
    for ci, co, _ in D.LAYERS:
        off += 9 * ci * co
        p[off:off + co] = r.normal(0, 0.05, co)
        off += co
    p[-17:-1] = r.normal(0, 0.3, 16)
    p[-1] = 0.1
    return p.astype(np.float32)


def _images(seed, B=2, H=64, W=64):
    r = np.random.default_rng(seed)
    x = r.gamma(1.5, 1 / 1.5, size=(B, H, W)).astype(np.float32)
    y = (0.5 * x + 0.5 * r.gamma(4.0, 0.25, size=(B, H, W))).astype(np.float32)
    return x, y


@pytest.mark.parametrize("H,W,B", [(64, 64, 2), (256, 256, 2), (128, 192, 1)])
def test_forward_matches_oracle(H, W, B):
    """(256, 256) is the size of the denoised workloads' grids."""
    from denoiser import Denoiser
    p = _params(1)
    x, _ = _images(2, B=B, H=H, W=W)
    dn = Denoiser(H, W, max_batch=B)
    dn.setParams(p)
    out = dn(x)
    ref, _ = D.Net(p, bf16=True).forward(x)
    # the net's correction (out - x); BF16 activations: differences only from FP32 order
    assert rel_l2(out - x, ref - x) < 1e-2
    assert np.max(np.abs(out - ref)) < 0.05 * np.max(np.abs(ref - x))


@pytest.mark.parametrize("scales,H,B", [(0, 64, 2), (3, 64, 2), (3, 256, 1), (0, 256, 2)])
def test_loss_and_gradients_match_oracle(scales, H, B):
    from denoiser import Denoiser
    p = _params(3)
    x, y = _images(4, B=B, H=H, W=H)
    dn = Denoiser(H, H, max_batch=B)
    dn.setParams(p)
    dn.setLoss(scales, 0.5)
    loss, g = dn.lossAndGrad(x, y)
    rloss, rg = D.Net(p, bf16=True).loss_grad(x, y, tuple(float(v) for v in dn.loss_lam))
    assert loss == pytest.approx(rloss, rel=1e-3)
    off = 0
    for l, (ci, co, _) in enumerate(D.LAYERS):
        nw = 9 * ci * co
        assert rel_l2(g[off:off + nw], rg[off:off + nw]) < 3e-2, ("W", l)
        assert rel_l2(g[off + nw:off + nw + co], rg[off + nw:off + nw + co]) < 3e-2, ("b", l)
        off += nw + co
    assert rel_l2(g[off:], rg[off:]) < 1e-3  # head: FP32 throughout


@pytest.mark.parametrize("H,n_low,n_high", [(64, 3000, 6000), (256, 100_000, 200_000)])
def test_pair_generator_matches_oracle(H, n_low, n_high):
    """(256, 100k) is the low-N side of the denoiser's training pairs."""
    from denoiser import PairSet
    W = H
    ps = PairSet(H, W, npairs=4, n_low=n_low, n_high=n_high, seed=21,
                 families=(0, 1, 2, 3))
    xs, ys = ps.host(0, 4)
    for k in range(4):
        spec = tuple(float(v) for v in ps.specs[k])
        spec = (int(spec[0]),) + spec[1:]
        lx, ly = D.pair_particles(spec, 21, k, 0, n_low)
        geo = D.square_grid(lx, ly, H, W)
        want_x = D.tsc_deposit(lx, ly, H, W, geo)
        hx, hy = D.pair_particles(spec, 21, k, 1, n_high)
        want_y = D.tsc_deposit(hx, hy, H, W, geo)
        # FP32 device arithmetic: a few particles may fall on the other side of a node
        assert rel_l2(xs[k], want_x) < 3e-2, k
        assert rel_l2(ys[k], want_y) < 3e-2, k
        assert xs[k].mean() == pytest.approx(1.0, rel=1e-3)


def _poisson_field(rho):
    """E = -grad phi of the free-space 2D potential of rho (unit grid spacing), by zero-
    padded FFT convolution with -ln r, on the nodes."""
    H, W = rho.shape
    ii = np.arange(2 * H)
    jj = np.arange(2 * W)
    di = np.minimum(ii, 2 * H - ii)[:, None]
    dj = np.minimum(jj, 2 * W - jj)[None, :]
    r2 = di ** 2 + dj ** 2
    G = np.where(r2 > 0, -0.5 * np.log(np.maximum(r2, 1)), 0.0)
    pad = np.zeros((2 * H, 2 * W))
    pad[:H, :W] = rho
    phi = np.fft.irfft2(np.fft.rfft2(pad) * np.fft.rfft2(G), s=(2 * H, 2 * W))[:H, :W]
    ex, ey = np.gradient(-phi)
    return np.stack([ex, ey])


def test_training_beats_plain_low_n_deposit():
    """Short training on 64 x 64 pairs (1 particle per node on average against 100): on
    held-out pairs the denoised rho and its field are much closer to the high-N ones than
    the plain low-N deposit is."""
    from denoiser import Denoiser, PairSet
    H = W = 64
    train = PairSet(H, W, npairs=2048, n_low=4096, n_high=409600, seed=100)
    test = PairSet(H, W, npairs=32, n_low=4096, n_high=409600, seed=999)
    dn = Denoiser(H, W, max_batch=16)
    dn.setParams(Denoiser.initParams(5))
    dn.setLoss(4, 1.0)
    hist = dn.train(train, epochs=3, batch=16, lr=2e-3, lr_final=2e-4, log_every=64, log=None)
    assert hist[-1][1] < 0.5 * hist[0][1]
    x, y = test.host(0, 32)
    out = dn(x)
    mse_plain = float(np.mean((x - y) ** 2))
    mse_dn = float(np.mean((out - y) ** 2))
    assert mse_dn < 0.5 * mse_plain, (mse_dn, mse_plain)
    ef = [rel_l2(_poisson_field(out[k]), _poisson_field(y[k])) for k in range(32)]
    ep = [rel_l2(_poisson_field(x[k]), _poisson_field(y[k])) for k in range(32)]
    assert np.median(ef) < 0.9 * np.median(ep), (np.median(ef), np.median(ep))


def test_pair_from_tracked_bunch_matches_oracle():
    """capture(): every stride-th particle of a bunch against all of them, on the square
    box around the sample (oracle deposit of the same particles)."""
    from denoiser import PairSet
    from workloads import gaussian_coords, make_bunch
    c = gaussian_coords(12000, seed=44)
    b = make_bunch(c)
    ps = PairSet(64, 64, capacity=2)
    ps.capture(b, 4)
    x, y = ps.host(0, 1)
    lx, ly = c[0, ::4].astype(np.float32).astype(np.float64), c[2, ::4].astype(np.float32).astype(np.float64)
    geo = D.square_grid(lx, ly, 64, 64)
    ax, ay = c[0].astype(np.float32).astype(np.float64), c[2].astype(np.float32).astype(np.float64)
    assert rel_l2(x[0], D.tsc_deposit(lx, ly, 64, 64, geo)) < 1e-2
    assert rel_l2(y[0], D.tsc_deposit(ax, ay, 64, 64, geo)) < 1e-2
    assert ps.npairs == 1


def test_gathered_batch_step_equals_contiguous_step():
    from denoiser import Denoiser, PairSet
    data = PairSet(64, 64, npairs=8, n_low=4096, n_high=40960, seed=3)
    d1, d2 = Denoiser(64, 64, max_batch=4), Denoiser(64, 64, max_batch=4)
    p = Denoiser.initParams(2)
    d1.setParams(p)
    d2.setParams(p)
    d1.trainStep(data, 4, 4, 1e-3)
    d2.trainStepIndices(data, [4, 5, 6, 7], 1e-3)
    # same batch: same loss; Adam's first step is lr * sign(gradient), so the updates agree
    # except where an FP32-summation-order difference flips the sign of a ~0 gradient
    assert d2.lastLoss() == pytest.approx(d1.lastLoss(), rel=1e-5)
    u1, u2 = d1.getParams() - p, d2.getParams() - p
    assert np.mean(np.abs(u1 - u2) > 5e-4) < 1e-3
