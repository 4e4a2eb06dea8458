"""Phase-space distributions (reference: py/orbit/bunch_generators/distribution_generators.py
:72-330).

Same classes and emittance conventions as the reference (the TwissContainer emittance is
the rms emittance; KV / waterbag scale it to their 100% emittance internally).  Sampling
uses a seeded numpy Generator instead of Python's `random`, and every class has a
vectorised sample(n) returning arrays; getCoordinates() draws one particle.
"""
import math

import numpy as np


class TwissContainer:
    def __init__(self, alpha, beta, emittance):
        self.alpha = alpha
        self.beta = beta
        self.emittance = emittance
        self._update()

    def _update(self):
        self.gamma = (1.0 + self.alpha ** 2) / self.beta
        self.u_max = math.sqrt(self.beta * self.emittance)
        self.up_coeff = math.sqrt(self.emittance / self.beta)
        self.up_max = math.sqrt(self.gamma * self.emittance)

    def setEmittance(self, emittance):
        self.emittance = emittance
        self._update()

    def getNormalizedH(self, u, up):
        return (u ** 2 + (self.alpha * u + self.beta * up) ** 2) / (self.beta * self.emittance)

    def getU_Max(self):
        return self.u_max

    def getUP_Max(self):
        return self.up_max

    def getU_UP(self, u_norm, up_norm):
        """u = sqrt(beta eps) u_n ; u' = sqrt(eps/beta) (u'_n - alpha u_n)."""
        return self.u_max * u_norm, self.up_coeff * (up_norm - self.alpha * u_norm)

    def getAlphaBetaGammaEmitt(self):
        return (self.alpha, self.beta, self.gamma, self.emittance)

    def getAlphaBetaEmitt(self):
        return (self.alpha, self.beta, self.emittance)


def _rng(rng):
    return rng if rng is not None else np.random.default_rng()


class _Dist:
    def getCoordinates(self):
        return tuple(float(v[0]) for v in self.sample(1, self._r))


class KVDist2D(_Dist):
    def __init__(self, twissX=TwissContainer(0., 1., 1.), twissY=TwissContainer(0., 1., 1.), rng=None):
        ax, bx, ex = twissX.getAlphaBetaEmitt()
        ay, by, ey = twissY.getAlphaBetaEmitt()
        self.twissX = TwissContainer(ax, bx, 4 * ex)
        self.twissY = TwissContainer(ay, by, 4 * ey)
        self._r = _rng(rng)

    def _norm(self, n, r):
        phi = 2 * np.pi * (r.random(n) - 0.5)
        rho = np.sqrt(r.random(n))
        xn, yn = rho * np.cos(phi), rho * np.sin(phi)
        p0 = np.sqrt(np.abs(1.0 - rho ** 2))
        phi = 2 * np.pi * (r.random(n) - 0.5)
        return xn, p0 * np.cos(phi), yn, p0 * np.sin(phi)

    def sample(self, n, rng=None):
        r = _rng(rng) if rng is not None else self._r
        xn, xpn, yn, ypn = self._norm(n, r)
        x, xp = self.twissX.getU_UP(xn, xpn)
        y, yp = self.twissY.getU_UP(yn, ypn)
        return x, xp, y, yp

    def getTwissContainers(self):
        return (self.twissX, self.twissY)


class WaterBagDist2D(_Dist):
    def __init__(self, twissX=TwissContainer(0., 1., 1.), twissY=TwissContainer(0., 1., 1.), rng=None):
        ax, bx, ex = twissX.getAlphaBetaEmitt()
        ay, by, ey = twissY.getAlphaBetaEmitt()
        self.kv = KVDist2D(TwissContainer(ax, bx, 6.0 * ex / 4.0), TwissContainer(ay, by, 6.0 * ey / 4.0))
        self._r = _rng(rng)

    def sample(self, n, rng=None):
        r = _rng(rng) if rng is not None else self._r
        x, xp, y, yp = self.kv.sample(n, r)
        g = np.sqrt(np.sqrt(r.random(n)))
        return g * x, g * xp, g * y, g * yp

    def getTwissContainers(self):
        return self.kv.getTwissContainers()


class GaussDist1D(_Dist):
    def __init__(self, twiss=TwissContainer(0., 1., 1.), cut_off=-1., rng=None):
        self.twiss = twiss
        self.cut_off = cut_off
        self._r = _rng(rng)

    def sample(self, n, rng=None):
        r = _rng(rng) if rng is not None else self._r
        xn, xpn = r.standard_normal(n), r.standard_normal(n)
        if self.cut_off > 0.:
            bad = xn ** 2 + xpn ** 2 > self.cut_off ** 2
            while bad.any():
                k = int(bad.sum())
                xn[bad], xpn[bad] = r.standard_normal(k), r.standard_normal(k)
                bad = xn ** 2 + xpn ** 2 > self.cut_off ** 2
        return self.twiss.getU_UP(xn, xpn)

    def getTwissContainers(self):
        return self.twiss


class GaussDist2D(_Dist):
    def __init__(self, twissX=TwissContainer(0., 1., 1.), twissY=TwissContainer(0., 1., 1.),
                 cut_off=-1., rng=None):
        self._r = _rng(rng)
        self.gx, self.gy = GaussDist1D(twissX, cut_off), GaussDist1D(twissY, cut_off)

    def sample(self, n, rng=None):
        r = _rng(rng) if rng is not None else self._r
        return (*self.gx.sample(n, r), *self.gy.sample(n, r))


class GaussDist3D(_Dist):
    def __init__(self, twissX=TwissContainer(0., 1., 1.), twissY=TwissContainer(0., 1., 1.),
                 twissZ=TwissContainer(0., 1., 1.), cut_off=-1., rng=None):
        self._r = _rng(rng)
        self.g = [GaussDist1D(t, cut_off) for t in (twissX, twissY, twissZ)]

    def sample(self, n, rng=None):
        r = _rng(rng) if rng is not None else self._r
        out = []
        for g in self.g:
            out += list(g.sample(n, r))
        return tuple(out)


class KVDist3D(_Dist):
    def __init__(self, twissX=TwissContainer(0., 1., 1.), twissY=TwissContainer(0., 1., 1.),
                 twissZ=TwissContainer(0., 1., 1.), rng=None):
        self.tw = [TwissContainer(t.alpha, t.beta, 6 * t.emittance) for t in (twissX, twissY, twissZ)]
        self._r = _rng(rng)

    def sample(self, n, rng=None):
        r = _rng(rng) if rng is not None else self._r
        q = np.empty((4, 0))
        while q.shape[1] < n:  # uniform in the unit 4-ball of (x, y, z, z') by rejection
            c = 2 * (r.random((4, 2 * (n - q.shape[1]) + 16)) - 0.5)
            c = c[:, (c ** 2).sum(axis=0) < 1.0]
            q = np.concatenate([q, c], axis=1)
        q = q[:, :n]
        pxy = np.sqrt(1.0 - (q ** 2).sum(axis=0))
        phi = 2 * np.pi * (r.random(n) - 0.5)
        x, xp = self.tw[0].getU_UP(q[0], pxy * np.cos(phi))
        y, yp = self.tw[1].getU_UP(q[1], pxy * np.sin(phi))
        z, zp = self.tw[2].getU_UP(q[2], q[3])
        return x, xp, y, yp, z, zp


class WaterBagDist3D(_Dist):
    def __init__(self, twissX=TwissContainer(0., 1., 1.), twissY=TwissContainer(0., 1., 1.),
                 twissZ=TwissContainer(0., 1., 1.), rng=None):
        tw = [TwissContainer(t.alpha, t.beta, 8 * t.emittance / 6.0) for t in (twissX, twissY, twissZ)]
        self.kv = KVDist3D(*tw)
        self._r = _rng(rng)

    def sample(self, n, rng=None):
        r = _rng(rng) if rng is not None else self._r
        c = self.kv.sample(n, r)
        g = r.random(n) ** (1.0 / 6.0)
        return tuple(g * v for v in c)


__all__ = ["TwissContainer", "KVDist2D", "WaterBagDist2D", "GaussDist1D", "GaussDist2D",
           "GaussDist3D", "KVDist3D", "WaterBagDist3D"]
