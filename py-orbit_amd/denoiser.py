"""denoiser -- learned denoiser of deposited charge density (no reference counterpart).

    from denoiser import Denoiser, PairSet
    dn = Denoiser(256, 256, max_batch=32)
    dn.setParams(Denoiser.initParams(seed=1))
    data = PairSet(256, 256, npairs=20000, n_low=100_000, n_high=16_000_000, seed=7)
    dn.train(data, epochs=4, batch=32, lr=1e-3)
    rho_clean = dn(rho_noisy)                        # [B, H, W] float32, mean-1 images

The network (a 3-level U-Net, 16/32/64 channels, 191,841 parameters) runs on the device:
BF16 weights and activations, FP32 accumulation on the matrix cores, FP32 master weights
and Adam, all through the library's own kernels (ob_dn_* in include/orbit_mi355x.h).
Training pairs are made on the device too: seeded particles of four distribution
families (Gaussian, KV, waterbag, hollow shell) with random aspect ratio and rotation,
deposited with the TSC weights of the 2.5D space-charge grid at a low and a high particle
count on the same grid.
"""
import ctypes as C

import numpy as np

import _obnative as N

# (Cin, Cout, level) of the conv3x3 layers, in parameter order
LAYERS = [(1, 16, 0), (16, 16, 0), (16, 32, 1), (32, 32, 1), (32, 64, 2), (64, 64, 2),
          (64, 64, 2), (64, 64, 2), (96, 32, 1), (32, 32, 1), (48, 16, 0), (16, 16, 0)]
FAMILIES = ("gaussian", "kv", "waterbag", "hollow")


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class DeviceArray:
    """A float32 buffer in device memory owned by the library (ob_dn_alloc)."""

    def __init__(self, n):
        self.n = int(n)
        p = C.c_void_p()
        N.call("ob_dn_alloc", self.n, C.byref(p))
        self.ptr = p

    def upload(self, host, offset=0):
        h = _f32(host).ravel()
        dst = C.c_void_p(self.ptr.value + 4 * offset)
        N.call("ob_dn_copy", dst, h.ctypes.data_as(C.c_void_p), h.size, 0)

    def download(self, count=None, offset=0):
        count = self.n - offset if count is None else count
        h = np.empty(count, dtype=np.float32)
        src = C.c_void_p(self.ptr.value + 4 * offset)
        N.call("ob_dn_copy", h.ctypes.data_as(C.c_void_p), src, count, 1)
        return h

    def at(self, offset):
        return C.c_void_p(self.ptr.value + 4 * offset)

    def __del__(self):
        p = getattr(self, "ptr", None)
        if p is not None and p.value and N._lib is not None:
            N._lib.ob_dn_free(p)
            self.ptr = None


def pair_specs(npairs, seed, families=(0, 1, 2, 3), aspect=(0.5, 2.0)):
    """Per-pair distribution: family, x/y scales for an aspect ratio drawn log-uniformly in
    `aspect`, rotation uniform in [0, pi).  Returns float32 [npairs, 5]."""
    r = np.random.default_rng(seed)
    fam = r.choice(np.asarray(families), npairs)
    ar = np.exp(r.uniform(np.log(aspect[0]), np.log(aspect[1]), npairs))
    th = r.uniform(0.0, np.pi, npairs)
    return np.stack([fam, np.sqrt(ar), 1.0 / np.sqrt(ar), np.cos(th), np.sin(th)],
                    axis=1).astype(np.float32)


class PairSet:
    """Training pairs resident on the device: x[k] (low-N deposit) and y[k] (high-N deposit
    on the same grid), both [H][W] float32 of mean 1.

    PairSet(H, W, npairs, n_low, n_high, seed) generates npairs pairs of the four
    distribution families.  PairSet(H, W, capacity=C) makes an empty set to which
    generate() adds generated pairs and capture() adds pairs taken from a bunch being
    tracked (its every stride-th particle against all of them)."""

    def __init__(self, H, W, npairs=0, n_low=None, n_high=None, seed=0, capacity=None,
                 chunk=256, **spec_kw):
        self.H, self.W = H, W
        self.capacity = capacity if capacity is not None else npairs
        if self.capacity < 1:
            raise ValueError("a pair set needs room for at least one pair")
        img = H * W
        self.x = DeviceArray(self.capacity * img)
        self.y = DeviceArray(self.capacity * img)
        self.npairs = 0
        self.specs = np.zeros((0, 5), dtype=np.float32)
        if npairs:
            self.generate(npairs, n_low, n_high, seed, chunk=chunk, **spec_kw)

    def _room(self, n):
        if self.npairs + n > self.capacity:
            raise ValueError(f"pair set full ({self.capacity} pairs)")

    def generate(self, npairs, n_low, n_high, seed, chunk=256, **spec_kw):
        self._room(npairs)
        img = self.H * self.W
        specs = pair_specs(npairs, seed, **spec_kw)
        for k0 in range(0, npairs, chunk):
            k1 = min(npairs, k0 + chunk)
            sp = _f32(specs[k0:k1])
            at = (self.npairs + k0) * img
            # pair j of a chunk draws its particles with (seed + 7919 k0, j)
            N.call("ob_dn_make_pairs", self.H, self.W, k1 - k0, N.fptr(sp),
                   int(seed + 7919 * k0) & (2 ** 64 - 1), int(n_low), int(n_high),
                   self.x.at(at), self.y.at(at))
        self.specs = np.concatenate([self.specs, specs])
        self.npairs += npairs

    def capture(self, bunch, stride):
        """Add the pair (every stride-th particle, all particles) of a bunch as it is now."""
        self._room(1)
        at = self.npairs * self.H * self.W
        N.call("ob_dn_pair_from_bunch", self.H, self.W, bunch._h, int(stride), self.x.at(at),
               self.y.at(at))
        self.specs = np.concatenate([self.specs, np.full((1, 5), -1, dtype=np.float32)])
        self.npairs += 1

    def host(self, k0, k1):
        img = self.H * self.W
        xs = self.x.download((k1 - k0) * img, k0 * img).reshape(k1 - k0, self.H, self.W)
        ys = self.y.download((k1 - k0) * img, k0 * img).reshape(k1 - k0, self.H, self.W)
        return xs, ys


class CapturingCalc:
    """Stands in for a SpaceChargeCalc2p5Drb in a lattice: every `every`-th trackBunch call
    first adds a training pair of the bunch to `pairs`, then runs the calculator."""

    def __init__(self, calc, pairs, stride, every=1):
        self.calc, self.pairs, self.stride, self.every = calc, pairs, stride, every
        self.calls = 0

    def trackBunch(self, bunch, *args):
        if self.calls % self.every == 0 and self.pairs.npairs < self.pairs.capacity:
            self.pairs.capture(bunch, self.stride)
        self.calls += 1
        return self.calc.trackBunch(bunch, *args)


class Denoiser:
    def __init__(self, H=256, W=256, max_batch=32):
        h = C.c_void_p()
        N.call("ob_dn_create", int(H), int(W), int(max_batch), C.byref(h))
        self._h = h
        self.H, self.W, self.max_batch = H, W, max_batch
        n = C.c_int64()
        N.call("ob_dn_num_params", self._h, C.byref(n))
        self.nparams = n.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and N._lib is not None:
            N._lib.ob_dn_destroy(h)
            self._h = None

    @staticmethod
    def initParams(seed=0):
        """He-normal conv weights, zero biases and a zero head: the untrained net returns
        its input unchanged."""
        rng = np.random.default_rng(seed)
        parts = []
        for ci, co, _ in LAYERS:
            parts.append(rng.normal(0.0, np.sqrt(2.0 / (9 * ci)), 9 * ci * co))
            parts.append(np.zeros(co))
        parts.append(np.zeros(17))
        return np.concatenate(parts).astype(np.float32)

    def setParams(self, p):
        p = _f32(p)
        if p.size != self.nparams:
            raise ValueError(f"expected {self.nparams} parameters, got {p.size}")
        N.call("ob_dn_set_params", self._h, N.fptr(p), p.size)

    def getParams(self):
        p = np.empty(self.nparams, dtype=np.float32)
        N.call("ob_dn_get_params", self._h, N.fptr(p), p.size)
        return p

    def save(self, path):
        np.savez(path, params=self.getParams(), H=self.H, W=self.W)

    def load(self, path):
        with np.load(path, allow_pickle=False) as f:
            self.setParams(f["params"])

    def __call__(self, x):
        """Denoise mean-1 images x [B, H, W] (host arrays)."""
        x = _f32(x)
        single = x.ndim == 2
        if single:
            x = x[None]
        out = np.empty_like(x)
        for b0 in range(0, x.shape[0], self.max_batch):
            xb = _f32(x[b0:b0 + self.max_batch])
            yb = np.empty_like(xb)
            N.call("ob_dn_forward", self._h, N.fptr(xb), N.fptr(yb), xb.shape[0])
            out[b0:b0 + xb.shape[0]] = yb
        return out[0] if single else out

    def setLoss(self, scales=0, weight=1.0):
        """Loss = mean e^2 + sum_{s=1..scales} weight * 4^s * mean of the squared 2^s x 2^s
        block means of e = out - target.  Block means of white noise shrink as 2^-s, so
        each scale then weighs like the pixel term: long wavelengths, which carry the
        field error, count as much as short ones.  scales = 0: plain MSE."""
        lam = np.array([weight * 4.0 ** s for s in range(1, scales + 1)], dtype=np.float32)
        N.call("ob_dn_set_loss", self._h, int(scales), N.fptr(lam) if scales else None)
        self.loss_lam = lam

    def lossAndGrad(self, x, y):
        """MSE loss and the gradient of every parameter for host batches x, y [B, H, W]."""
        x, y = _f32(x), _f32(y)
        loss = np.zeros(1, dtype=np.float32)
        grad = np.empty(self.nparams, dtype=np.float32)
        N.call("ob_dn_loss_grad", self._h, N.fptr(x), N.fptr(y), x.shape[0], N.fptr(loss),
               N.fptr(grad))
        return float(loss[0]), grad

    def trainStep(self, data, k0, batch, lr):
        img = self.H * self.W
        N.call("ob_dn_train_step_dev", self._h, data.x.at(k0 * img), data.y.at(k0 * img),
               int(batch), float(lr))

    def trainStepIndices(self, data, idx, lr):
        """One Adam step on the pairs idx of `data` (gathered into a batch buffer)."""
        img = self.H * self.W
        b = len(idx)
        if getattr(self, "_bx", None) is None:
            self._bx = DeviceArray(self.max_batch * img)
            self._by = DeviceArray(self.max_batch * img)
        ii = np.ascontiguousarray(idx, dtype=np.int32)
        pi = ii.ctypes.data_as(C.POINTER(C.c_int))
        N.call("ob_dn_gather", data.x.ptr, data.npairs, self._bx.ptr, img, pi, b)
        N.call("ob_dn_gather", data.y.ptr, data.npairs, self._by.ptr, img, pi, b)
        N.call("ob_dn_train_step_dev", self._h, self._bx.ptr, self._by.ptr, b, float(lr))

    def lastLoss(self):
        v = np.zeros(1, dtype=np.float32)
        N.call("ob_dn_last_loss", self._h, N.fptr(v))
        return float(v[0])

    def train(self, data, epochs=1, batch=32, lr=1e-3, lr_final=None, log_every=0,
              log=print, npairs=None, shuffle=False, seed=0):
        """Adam over the pairs of `data`, `epochs` times, with the learning rate decayed
        geometrically to lr_final (default lr / 10); in order, or in a new random order
        every epoch (shuffle).  Returns (step, loss) every log_every-th step."""
        npairs = data.npairs if npairs is None else npairs
        steps_per_epoch = npairs // batch
        total = epochs * steps_per_epoch
        lr_final = lr / 10 if lr_final is None else lr_final
        rng = np.random.default_rng(seed)
        hist = []
        step = 0
        for ep in range(epochs):
            order = rng.permutation(npairs) if shuffle else None
            for s in range(steps_per_epoch):
                frac = step / max(1, total - 1)
                rate = lr * (lr_final / lr) ** frac
                if shuffle:
                    self.trainStepIndices(data, order[s * batch:(s + 1) * batch], rate)
                else:
                    self.trainStep(data, s * batch, batch, rate)
                step += 1
                if log_every and step % log_every == 0:
                    hist.append((step, self.lastLoss()))
                    if log:
                        log(f"step {step}/{total} loss {hist[-1][1]:.5f}")
        return hist
