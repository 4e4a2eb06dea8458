"""Seeded synthetic bunches and lattices for the tracking workloads used by the tests, the
benchmark and the smoke check.

  fodo_2p5drb : protons at 1.0 GeV, 6D Gaussian (sigma_x = sigma_y = 3 mm,
                sigma_xp = sigma_yp = 0.5 mrad, z uniform in [-100, 100] m, dE sigma 1 MeV);
                FODO cell of two 0.5 m TEAPOT quads (k1 = +-0.5 m^-2) and two 2.5 m drifts,
                an SC2p5Drb node every 0.5 m (12 per cell) on a 128x128 grid with 64 z bins.
  ring_3d     : 4D waterbag (radii 6 mm and 1 mrad), z Gaussian sigma 0.3 m, dE sigma 1 MeV;
                24-cell ring of the FODO cell plus two 0.2 m sextupoles (k2 = +-2 m^-3),
                SpaceChargeCalc3D on 256x256x128, an SC node every ~0.5 m.

Intensity 1e13 protons per bunch (macroSize = 1e13 / N); pipe radius 0.1 m.
"""
import numpy as np

from bunch import Bunch
from orbit.teapot import DriftTEAPOT, MultipoleTEAPOT, QuadTEAPOT, TEAPOT_Lattice, TEAPOT_Ring

SEED = 20261202
INTENSITY = 1.0e13
PIPE_RADIUS = 0.1


def gaussian_coords(n, seed=SEED, sx=3e-3, sxp=0.5e-3, zhalf=100.0, sdE=1e-3):
    r = np.random.default_rng(seed)
    c = np.empty((6, n))
    c[0] = r.normal(0.0, sx, n)
    c[1] = r.normal(0.0, sxp, n)
    c[2] = r.normal(0.0, sx, n)
    c[3] = r.normal(0.0, sxp, n)
    c[4] = r.uniform(-zhalf, zhalf, n)
    c[5] = r.normal(0.0, sdE, n)
    return c


def waterbag_coords(n, seed=SEED, rx=6e-3, rxp=1e-3, sz=0.3, sdE=1e-3):
    """Uniform in the 4D ellipsoid (x/rx)^2+(xp/rxp)^2+(y/rx)^2+(yp/rxp)^2 <= 1."""
    r = np.random.default_rng(seed)
    g = r.standard_normal((4, n))
    g /= np.sqrt((g ** 2).sum(axis=0))
    g *= r.random(n) ** 0.25
    c = np.empty((6, n))
    c[0], c[1], c[2], c[3] = rx * g[0], rxp * g[1], rx * g[2], rxp * g[3]
    c[4] = r.normal(0.0, sz, n)
    c[5] = r.normal(0.0, sdE, n)
    return c


def kv_coords(n, seed=SEED, rx=6e-3, rxp=1e-3, zhalf=100.0, sdE=1e-3):
    """Kapchinskij-Vladimirskij: uniform on the 4D ellipsoid surface
    (x/rx)^2+(xp/rxp)^2+(y/rx)^2+(yp/rxp)^2 = 1 (uniform x-y projection)."""
    r = np.random.default_rng(seed)
    g = r.standard_normal((4, n))
    g /= np.sqrt((g ** 2).sum(axis=0))
    c = np.empty((6, n))
    c[0], c[1], c[2], c[3] = rx * g[0], rxp * g[1], rx * g[2], rxp * g[3]
    c[4] = r.uniform(-zhalf, zhalf, n)
    c[5] = r.normal(0.0, sdE, n)
    return c


BATCH_BETA = 6.0  # m, beta function used to turn an emittance into beam sizes


def batch_specs(nbunch=64, seed=SEED, nmin=50_000, nmax=2_000_000):
    """Bunches of the batched workload: sizes uniform in [nmin, nmax], Gaussian or KV,
    rms emittance 1-20 mm mrad, intensity 1e12-1e14, 2.5D grids 64^2, 128^2 or 256^2 with
    64 z bins (all seeded)."""
    r = np.random.default_rng(seed + 7)
    specs = []
    for k in range(nbunch):
        n = int(r.integers(nmin, nmax + 1))
        eps = float(r.uniform(1e-6, 20e-6))
        kind = "gauss" if r.random() < 0.5 else "kv"
        g = int(r.choice([64, 128, 256]))
        specs.append({"n": n, "eps": eps, "kind": kind, "grid": (g, g, 64),
                      "intensity": float(10 ** r.uniform(12, 14)), "seed": seed + 1000 + k})
    return specs


def batch_coords(spec):
    s, e = np.sqrt(spec["eps"] * BATCH_BETA), np.sqrt(spec["eps"] / BATCH_BETA)
    if spec["kind"] == "gauss":
        return gaussian_coords(spec["n"], seed=spec["seed"], sx=s, sxp=e)
    return kv_coords(spec["n"], seed=spec["seed"], rx=2 * s, rxp=2 * e)  # rms = edge / 2


def make_bunch(coords, kin_energy=1.0, intensity=INTENSITY):
    b = Bunch()
    b.getSyncParticle().kinEnergy(kin_energy)
    b.macroSize(intensity / coords.shape[1])
    b.setCoordinates(coords)
    return b


def fodo_cell(prefix="", sextupoles=False):
    """QF(0.5 m, nParts 2) [S1] D(2.5 m, nParts 5) QD [S2] D; k1 = +-0.5, k2 = +-2."""
    nodes = []
    qf = QuadTEAPOT(prefix + "QF")
    qf.setLength(0.5)
    qf.setParam("kq", 0.5)
    nodes.append(qf)
    if sextupoles:
        s1 = MultipoleTEAPOT(prefix + "S1")
        s1.setLength(0.2)
        s1.setParam("poles", [2])
        s1.setParam("kls", [2.0 * 0.2])
        s1.setParam("skews", [0])
        nodes.append(s1)
    d1 = DriftTEAPOT(prefix + "D1")
    d1.setLength(2.5)
    d1.setnParts(5)
    nodes.append(d1)
    qd = QuadTEAPOT(prefix + "QD")
    qd.setLength(0.5)
    qd.setParam("kq", -0.5)
    nodes.append(qd)
    if sextupoles:
        s2 = MultipoleTEAPOT(prefix + "S2")
        s2.setLength(0.2)
        s2.setParam("poles", [2])
        s2.setParam("kls", [-2.0 * 0.2])
        s2.setParam("skews", [0])
        nodes.append(s2)
    d2 = DriftTEAPOT(prefix + "D2")
    d2.setLength(2.5)
    d2.setnParts(5)
    nodes.append(d2)
    return nodes


def fodo_lattice(ncells=1):
    lat = TEAPOT_Lattice("FODO")
    for c in range(ncells):
        for n in fodo_cell("c%d:" % c):
            lat.addNode(n)
    lat.initialize()
    return lat


def ring_lattice(ncells=24):
    lat = TEAPOT_Ring("RING")
    for c in range(ncells):
        for n in fodo_cell("c%d:" % c, sextupoles=True):
            lat.addNode(n)
    lat._addChildren()
    lat.initialize()
    return lat
