"""CPU restatement of the learned charge-density denoiser (numpy, float64).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device kernels in
py-orbit_amd/csrc/ob_denoise.hip; never imported by the product.

The denoiser has no counterpart in the reference, so nothing here follows a reference
file: the network is the one the device library defines (a 3-level U-Net, 16/32/64
channels, 12 conv3x3 layers and a 1x1 head, parameter layout of ob_dn_set_params).

`bf16=True` rounds exactly where the device stores BF16 (weights, every activation and
activation gradient, the pooled and upsampled operands) with round-to-nearest-even, so the
device result differs from this one only by FP32 summation order.  `bf16=False` is exact
float64 arithmetic, the mode the finite-difference and adjoint checks pin.

Pinned by: conv3x3 against scipy.signal.correlate2d; the backward pass against central
finite differences of the loss (bf16=False); pool/upsample gradients by the adjoint
identity <A x, y> = <x, A^T y>; the pair generator's particles against closed-form
moments of each distribution family.
"""
import numpy as np

# (Cin, Cout, level) of the 12 conv3x3 layers c1a c1b c2a c2b c3a c3b c3c c3d c4a c4b c5a c5b
LAYERS = [(1, 16, 0), (16, 16, 0), (16, 32, 1), (32, 32, 1), (32, 64, 2), (64, 64, 2),
          (64, 64, 2), (64, 64, 2), (96, 32, 1), (32, 32, 1), (48, 16, 0), (16, 16, 0)]


def num_params():
    return sum(9 * ci * co + co for ci, co, _ in LAYERS) + 17


def split_params(p):
    """Parameter vector -> [(W[9][Cin][Cout], b[Cout])] per layer, head weights, head bias."""
    out, off = [], 0
    for ci, co, _ in LAYERS:
        w = p[off:off + 9 * ci * co].reshape(9, ci, co)
        off += 9 * ci * co
        b = p[off:off + co]
        off += co
        out.append((w, b))
    return out, p[off:off + 16], p[off + 16]


def bf16_round(a):
    """Round to the nearest BF16 value (ties to even), via float32; returned as float64."""
    f = np.ascontiguousarray(np.asarray(a, dtype=np.float64).astype(np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def conv3x3(x, w, b=None):
    """Same-size 3x3 correlation with zero padding: x [B,H,W,Cin], w [9][Cin][Cout]
    (tap 3*ky + kx reads pixel (y + ky - 1, x + kx - 1))."""
    B, H, W, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((B, H, W, w.shape[2]))
    for tap in range(9):
        ky, kx = divmod(tap, 3)
        out += xp[:, ky:ky + H, kx:kx + W, :] @ w[tap]
    if b is not None:
        out += b
    return out


def conv3x3_wgrad(x, gy):
    """dW[tap][ci][co] = sum_p x[p + off(tap)][ci] gy[p][co]."""
    B, H, W, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    dw = np.zeros((9, x.shape[3], gy.shape[3]))
    for tap in range(9):
        ky, kx = divmod(tap, 3)
        dw[tap] = np.einsum("bhwc,bhwd->cd", xp[:, ky:ky + H, kx:kx + W, :], gy)
    return dw


def conv3x3_dgrad(gy, w):
    """Gradient w.r.t. the conv input: correlation of gy with the tap-flipped, transposed
    kernel."""
    wt = np.transpose(w[::-1], (0, 2, 1))
    return conv3x3(gy, wt)


def pool(x):
    B, H, W, C = x.shape
    return x.reshape(B, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4))


def pool_grad(g):
    return 0.25 * np.repeat(np.repeat(g, 2, axis=1), 2, axis=2)


def up(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up_grad(g):
    B, H, W, C = g.shape
    return g.reshape(B, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4))


class Net:
    def __init__(self, params, bf16=True):
        self.layers, self.hw, self.hb = split_params(np.asarray(params, dtype=np.float64))
        self.bf16 = bf16
        self.r = bf16_round if bf16 else (lambda a: a)
        self.wq = [(self.r(w), b) for w, b in self.layers]

    def _input(self, l, A, x):
        r = self.r
        if l == 0:
            return x
        if l in (2, 4):
            return r(pool(A[l - 1]))
        if l == 8:
            return np.concatenate([up(A[7]), A[3]], axis=3)
        if l == 10:
            return np.concatenate([up(A[9]), A[1]], axis=3)
        return A[l - 1]

    def forward(self, x):
        """x: [B,H,W] -> (out [B,H,W], activations)."""
        x = np.asarray(x, dtype=np.float64)
        xin = self.r(x)[..., None]
        A = []
        for l in range(12):
            w, b = self.wq[l]
            A.append(self.r(np.maximum(conv3x3(self._input(l, A, xin), w, b), 0.0)))
        out = x + self.hb + A[11] @ self.hw
        self.cache = (x, xin, A)
        return out, A

    def loss_grad(self, x, y, lam=()):
        """Loss and the gradient vector (layout of the parameters).  Loss = mean e^2 +
        sum_s lam[s-1] mean over 2^s x 2^s blocks of (block mean of e)^2, e = out - y."""
        out, A = self.forward(x)
        x, xin, _ = self.cache
        r = self.r
        npix = out.size
        e = out - y
        loss = float(np.sum(e * e) / npix)
        es = e.copy()
        m = e
        for s, l in enumerate(lam, start=1):
            B, H, W = m.shape
            m = m.reshape(B, H // 2, 2, W // 2, 2).mean(axis=(2, 4))
            loss += l * float(np.mean(m * m))
            es += l * np.repeat(np.repeat(m, 2 ** s, axis=1), 2 ** s, axis=2)
        g = 2.0 * es / npix
        dhw = np.einsum("bhw,bhwc->c", g, A[11])
        dhb = g.sum()
        G = [None] * 12
        G[11] = r(g[..., None] * self.hw)
        grads = [None] * 12
        skip1 = skip2 = None
        for l in range(11, -1, -1):
            w, _ = self.wq[l]
            gy = G[l] * (A[l] > 0)
            X = self._input(l, A, xin)
            grads[l] = (conv3x3_wgrad(X, gy), gy.sum(axis=(0, 1, 2)))
            if l == 0:
                break
            dx = r(conv3x3_dgrad(gy, w))
            if l == 10:
                G[9] = r(up_grad(dx[..., :32]))
                skip1 = dx[..., 32:]
            elif l == 8:
                G[7] = r(up_grad(dx[..., :64]))
                skip2 = dx[..., 64:]
            elif l == 4:
                G[3] = r(pool_grad(dx) + skip2)
            elif l == 2:
                G[1] = r(pool_grad(dx) + skip1)
            else:
                G[l - 1] = dx
        vec = []
        for dw, db in grads:
            vec += [dw.ravel(), db]
        vec += [dhw, np.array([dhb])]
        return loss, np.concatenate(vec)


def he_init(seed):
    """He-normal conv weights, zero biases, zero head (the untrained net is the identity)."""
    rng = np.random.default_rng(seed)
    parts = []
    for ci, co, _ in LAYERS:
        parts.append(rng.normal(0.0, np.sqrt(2.0 / (9 * ci)), 9 * ci * co))
        parts.append(np.zeros(co))
    parts.append(np.zeros(17))
    return np.concatenate(parts)


# ---- training-pair generator ------------------------------------------------------------
M64 = (1 << 64) - 1


def _mix64(z):
    z = (z + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def pair_particles(spec, seed, k, which, n):
    """Particles 0..n-1 of pair k (which 0: low-N, 1: high-N), float64 restatement of the
    device generator (splitmix64 counters, Box-Muller / radial inverse CDFs)."""
    fam, sa, sb, cs, sn = spec
    xs, ys = np.empty(n), np.empty(n)
    key = _mix64(((k << 33) ^ (which << 32) ^ 0x5bd1e995) & M64)
    for i in range(n):
        h0 = _mix64(seed ^ key ^ ((i * 0x2545F4914F6CDD1D) & M64))
        h1 = _mix64(h0)
        u1 = ((h0 >> 40) + 0.5) / 16777216.0
        u2 = ((h1 >> 40) + 0.5) / 16777216.0
        if fam == 0:
            rr = np.sqrt(-2.0 * np.log(u1))
        elif fam == 1:
            rr = 2.0 * np.sqrt(u1)
        elif fam == 2:
            rr = 2.449489743 * np.sqrt(1.0 - np.sqrt(1.0 - u1))
        else:
            rr = 1.714985851 * np.sqrt(0.36 + 0.64 * u1)
        px = rr * np.cos(2 * np.pi * u2) * sa
        py = rr * np.sin(2 * np.pi * u2) * sb
        xs[i] = cs * px - sn * py
        ys[i] = sn * px + cs * py
    return xs, ys


def tsc_deposit(xs, ys, H, W, geo):
    """TSC deposit (Grid2D.cc:337-350 index clamp and weights) normalized to mean 1."""
    xmin, dx, ymin, dy = geo
    out = np.zeros((H, W))
    n = len(xs)
    for x, y in zip(xs, ys):
        fx, fy = (x - xmin) / dx, (y - ymin) / dy
        if not (0 <= fx <= H - 1 and 0 <= fy <= W - 1):
            continue
        ix = min(max(int(fx + 0.5), 1), H - 2)
        iy = min(max(int(fy + 0.5), 1), W - 2)
        xf, yf = fx - ix, fy - iy
        wx = [0.5 * (0.5 - xf) ** 2, 0.75 - xf * xf, 0.5 * (0.5 + xf) ** 2]
        wy = [0.5 * (0.5 - yf) ** 2, 0.75 - yf * yf, 0.5 * (0.5 + yf) ** 2]
        for a in range(3):
            for c in range(3):
                out[ix - 1 + a, iy - 1 + c] += wx[a] * wy[c]
    return out * (H * W / n)


def square_grid(xs, ys, H, W):
    """The generator's grid: the square box around the sample's extrema."""
    xc, yc = 0.5 * (xs.min() + xs.max()), 0.5 * (ys.min() + ys.max())
    half = 0.5 * max(xs.max() - xs.min(), ys.max() - ys.min())
    return (xc - half, 2 * half / (H - 1), yc - half, 2 * half / (W - 1))
