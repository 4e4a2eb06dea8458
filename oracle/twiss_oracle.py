"""CPU restatement of BunchTwissAnalysis (numpy, float64).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of ob_bunch_moments and of the
Python BunchTwissAnalysis built on it; never imported by the product.

Follows src/orbit/BunchDiagnostics/BunchTwissAnalysis.cc:
  analyze      :49-142  (weighted sums, normalized by the total macrosize when nonzero)
  correlation  :200-203, emittance :228-246, alpha/beta/gamma :257-306,
  dispersion   :309-327, effective emittance/alpha/beta/gamma :330-420.

The reference's analyzeBunch needs its Bunch class (Bunch.cc and the Python-2 wrappers it
links against), so it is not compiled here.  Pinned instead by closed forms: a particle
set built from a zero-mean, unit, uncorrelated base set through the Courant-Snyder
normalization with known (alpha, beta, emittance) and a known dispersion term must return
exactly those values (tests/test_twiss_oracle.py), and the weighted sums must equal
replicated-particle sums for integer weights.
"""
import sys

import numpy as np


def sums(c, w=None):
    """The 28 sums of ob_bunch_moments: [sum w, sum w c_k (6), sum w c_a c_b (a <= b, row
    by row)] for coordinates c [6, n] and weights w (None: 1)."""
    c = np.asarray(c, dtype=np.float64)
    w = np.ones(c.shape[1]) if w is None else np.asarray(w, dtype=np.float64)
    out = [w.sum()]
    out += list((c * w).sum(axis=1))
    for a in range(6):
        for b in range(a, 6):
            out.append(float(np.sum(w * c[a] * c[b])))
    return np.array(out)


class Twiss:
    """Results of analyzeBunch for coordinates c [6, n], weights w, and the synchronous
    particle's momentum p [GeV/c], beta and gamma."""

    def __init__(self, c, w=None, p=1.0, beta=1.0, gamma=1.0):
        s = sums(c, w)
        self.count = np.asarray(c).shape[1]
        self.total = s[0]
        raw = np.zeros((6, 6))
        k = 7
        for a in range(6):
            for b in range(a, 6):
                raw[b, a] = s[k]
                k += 1
        if abs(self.total) > 0:
            self.avg = s[1:7] / self.total
            low = raw / self.total
            self.corr = low + np.tril(low, -1).T
        else:
            self.avg = s[1:7]
            self.corr = raw
        self.p, self.beta, self.gamma = p, beta, gamma

    def correlation(self, i, j):
        return self.corr[i, j] - self.avg[i] * self.avg[j]

    def _plane(self, ic):
        g = self.correlation
        return (abs(g(2 * ic, 2 * ic)), abs(g(2 * ic + 1, 2 * ic + 1)), g(2 * ic, 2 * ic + 1),
                g(2 * ic, 5), g(2 * ic + 1, 5), abs(g(5, 5)))

    def emittance(self, ic):
        x2, xp2, xxp, xdE, xpdE, dE2 = self._plane(ic)
        if ic == 2 or dE2 == 0:
            return np.sqrt(abs(x2 * xp2 - xxp ** 2))
        return np.sqrt(abs((x2 - xdE ** 2 / dE2) * (xp2 - xpdE ** 2 / dE2)
                           - (xxp - xdE * xpdE / dE2) ** 2))

    def alpha(self, ic):
        x2, xp2, xxp, xdE, xpdE, dE2 = self._plane(ic)
        num = xxp if ic == 2 else xxp - xdE * xpdE / dE2
        return -num / self.emittance(ic)

    def beta_fn(self, ic):
        x2, xp2, xxp, xdE, xpdE, dE2 = self._plane(ic)
        num = x2 if ic == 2 else x2 - xdE ** 2 / dE2
        return num / self.emittance(ic)

    def gamma_fn(self, ic):
        x2, xp2, xxp, xdE, xpdE, dE2 = self._plane(ic)
        num = xp2 if ic == 2 else xp2 - xpdE ** 2 / dE2
        return num / self.emittance(ic)

    def dispersion(self, ic):
        return self.correlation(2 * ic, 5) / abs(self.correlation(5, 5)) * self.p * self.beta

    def dispersion_prime(self, ic):
        return self.correlation(2 * ic + 1, 5) / abs(self.correlation(5, 5)) * self.p * self.beta

    def effective(self, ic):
        """(emittance, alpha, beta, gamma) including the dispersive part."""
        x2, xp2, xxp = self._plane(ic)[:3]
        e2 = x2 * xp2 - xxp ** 2
        if e2 <= 0:
            return np.sqrt(abs(e2)), 0.0, 0.0, sys.float_info.max
        e = np.sqrt(e2)
        return e, -xxp / e, x2 / e, xp2 / e
