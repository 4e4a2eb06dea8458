"""Pins of the oracle against what the mathematics fixes, for the parts no reference build
or golden vector covers here (TEAPOT maps, FFT, Poisson solves, the space-charge kicks).

Each check is independent of the oracle's own formula: numpy's FFT, brute-force
convolution, textbook transfer matrices, analytic fields of uniform beams, exact
polynomial fits, symplecticity of the maps' Jacobians."""
import math

import numpy as np
import pytest

import oracle as O
from helpers import rel_l2

# ---------------------------------------------------------------- kinematics


def test_sync_particle_kinematics():
    # protons at 1 GeV kinetic energy: gamma = 1 + T/m, beta = sqrt(1-1/gamma^2), p = m*beta*gamma
    s = O.Bunch(np.zeros((6, 1)), kin_energy=1.0).sync
    m = 0.938272013
    g = 1.0 + 1.0 / m
    assert s["gamma"] == pytest.approx(g, rel=1e-15)
    assert s["beta"] == pytest.approx(math.sqrt(1 - 1 / g ** 2), rel=1e-15)
    assert s["p"] == pytest.approx(m * g * math.sqrt(1 - 1 / g ** 2), rel=1e-15)


# ---------------------------------------------------------------- FFT

@pytest.mark.parametrize("n", [2, 8, 64, 512])
def test_fft_c2c_vs_numpy(n):
    r = np.random.default_rng(n)
    z = r.normal(size=n) + 1j * r.normal(size=n)
    assert rel_l2(O.fft_c2c(z), np.fft.fft(z)) < 1e-14
    assert rel_l2(O.fft_c2c(z, inverse=True), np.fft.ifft(z) * n) < 1e-14


def test_fft_r2c_2d_3d_vs_numpy():
    r = np.random.default_rng(1)
    a = r.normal(size=(16, 32))
    assert rel_l2(O.fft_r2c_2d(a), np.fft.rfft2(a)) < 1e-14
    b = r.normal(size=(8, 16, 32))
    assert rel_l2(O.fft_r2c_3d(b), np.fft.rfftn(b)) < 1e-14
    s = np.fft.rfft2(a)
    assert rel_l2(O.fft_c2r_2d(s, 32), a * a.size) < 1e-14


# ---------------------------------------------------------------- Poisson solves

def _brute2d(rho, dx, dy):
    nx, ny = rho.shape
    i = np.arange(nx)[:, None, None, None] - np.arange(nx)[None, None, :, None]
    j = np.arange(ny)[None, :, None, None] - np.arange(ny)[None, None, None, :]
    r2 = (i * dx) ** 2 + (j * dy) ** 2
    with np.errstate(divide="ignore"):
        G = np.where(r2 > 0, -0.5 * np.log(r2), 0.0)
    return np.einsum("abcd,cd->ab", G, rho)


@pytest.mark.parametrize("nx,ny,solver", [(8, 8, "same"), (12, 10, "same"), (16, 16, "scaled")])
def test_poisson2d_vs_brute_force(nx, ny, solver):
    """The zero-padded FFT convolution equals the direct sum of G(r) = -ln|r| (G(0) = 0);
    with the solver on another grid the log rescaling must make it exact too."""
    r = np.random.default_rng(nx + ny)
    rho = r.uniform(0, 1, size=(nx, ny))
    rl = np.array([-2e-3, 3e-3, -1e-3, 1e-3 * (5e-3 / (nx - 1)) / (2e-3 / (ny - 1)) * (2e-3 / 2e-3)])
    dx = (rl[1] - rl[0]) / (nx - 1)
    dy_target = dx * 0.7
    rl[3] = rl[2] + dy_target * (ny - 1)
    dy = (rl[3] - rl[2]) / (ny - 1)
    if solver == "same":
        sl = rl.copy()
    else:  # same aspect ratio, 37x larger: exercises the log(dx_solver/dx_rho) rescale
        sl = rl * 37.0
    phi = O.poisson2d(rho, sl, rl)
    assert rel_l2(phi, _brute2d(rho, dx, dy)) < 1e-13


def test_poisson3d_vs_brute_force():
    nx, ny, nz = 6, 5, 4
    r = np.random.default_rng(3)
    rho = r.uniform(0, 1, size=(nz, nx, ny))
    rl = np.array([-1.0, 1.5, -1.0, 1.0, -0.6, 1.0])
    dx, dy, dz = (rl[1] - rl[0]) / (nx - 1), (rl[3] - rl[2]) / (ny - 1), (rl[5] - rl[4]) / nz
    sl = rl * 3.0  # solver box of the same shape, 3x larger
    phi = O.poisson3d(rho, sl, rl)
    iz, ix, iy = np.meshgrid(np.arange(nz), np.arange(nx), np.arange(ny), indexing="ij")
    P = np.stack([ix.ravel() * dx, iy.ravel() * dy, iz.ravel() * dz], axis=1)
    d = np.sqrt(((P[:, None, :] - P[None, :, :]) ** 2).sum(-1))
    with np.errstate(divide="ignore"):
        G = np.where(d > 0, 1.0 / d, 0.0)
    brute = (G @ rho.ravel()).reshape(nz, nx, ny)
    assert rel_l2(phi, brute) < 1e-13


def test_poisson3d_external_bunch_images_vs_brute_force():
    """Green function with nb image bunches at spacing lam along z: the Hockney solve equals
    the direct sum of the charge of every cell and of its images at z + b*lam (b = +-1 ..
    +-nb/2), each contributing 1/distance; a cell's own images contribute
    2*(1 + 1/2 + ... + 1/(nb/2))/lam (harmonic sum) to its own potential."""
    nx, ny, nz, nb, lam = 5, 4, 6, 4, 0.37
    r = np.random.default_rng(9)
    rho = r.uniform(0, 1, size=(nz, nx, ny))
    rl = np.array([-1.0, 1.2, -0.5, 0.7, -0.4, 0.5])
    dx, dy, dz = (rl[1] - rl[0]) / (nx - 1), (rl[3] - rl[2]) / (ny - 1), (rl[5] - rl[4]) / nz
    phi = O.poisson3d_images(rho, rl, rl, nb, lam)
    iz, ix, iy = np.meshgrid(np.arange(nz), np.arange(nx), np.arange(ny), indexing="ij")
    P = np.stack([ix.ravel() * dx, iy.ravel() * dy, iz.ravel() * dz], axis=1)
    brute = np.zeros(P.shape[0])
    q = rho.ravel()
    for b in [b for b in range(-nb // 2, nb // 2 + 1)]:
        Q = P + np.array([0.0, 0.0, b * lam])
        d = np.sqrt(((P[:, None, :] - Q[None, :, :]) ** 2).sum(-1))
        with np.errstate(divide="ignore"):
            G = np.where(d > 0, 1.0 / d, 0.0)
        brute += G @ q
    assert rel_l2(phi.ravel(), brute) < 1e-13
    # a single unit charge: its own-image term is the harmonic sum
    one = np.zeros((nz, nx, ny))
    one[2, 2, 1] = 1.0
    p1 = O.poisson3d_images(one, rl, rl, nb, lam)
    p0 = O.poisson3d_images(one, rl, rl, 0, lam)
    harmonic = 2.0 * sum(1.0 / k for k in range(1, nb // 2 + 1)) / lam
    assert abs((p1 - p0)[2, 2, 1] - harmonic) < 1e-12 * harmonic


# ---------------------------------------------------------------- dlambda/dz

def test_long_derivative_exact_on_quadratic_density():
    """3-point least squares reproduces a quadratic exactly: dlambda/dz = B'(z)/dz^2 for a
    per-bin content B that is quadratic in z."""
    nz, zmin, zmax = 16, -2.0, 3.0
    dz = (zmax - zmin) / nz
    zc = zmin + (np.arange(nz) + 0.5) * dz
    B = 3.0 + 0.7 * zc - 0.2 * zc ** 2
    zd = O.long_derivative(B, zmin, zmax, 3)
    assert np.allclose(zd, (0.7 - 0.4 * zc) / dz, rtol=1e-10, atol=1e-10)


def test_long_derivative_matches_polyfit_windows():
    nz, zmin, zmax, navg = 12, 0.0, 1.0, 5
    dz = (zmax - zmin) / nz
    B = np.random.default_rng(4).uniform(1, 2, nz)
    zd = O.long_derivative(B, zmin, zmax, navg)
    for iz in range(nz):
        s = min(max(iz - navg // 2, 0), nz - navg)
        c = np.polyfit(np.arange(1, navg + 1, dtype=float), B[s:s + navg], 2)
        x = iz - s + 1
        assert zd[iz] == pytest.approx((2 * c[0] * x + c[1]) / dz / dz, rel=1e-9)


# ---------------------------------------------------------------- TEAPOT maps

def _one(c6, op, *args):
    ob = O.Bunch(np.array(c6, dtype=float).reshape(6, -1), kin_energy=1.0)
    ob.apply(op, *args)
    return ob


def test_quad1_is_thick_lens_matrix():
    x0 = np.array([2e-3, -1e-4, -1e-3, 3e-4, 0.0, 0.0])  # dE = 0: on momentum
    for kq in (0.5, -0.5):
        L = 0.37
        out = _one(x0, "quad1", L, kq).coords[:, 0]
        w = math.sqrt(abs(kq))
        foc = np.array([[math.cos(w * L), math.sin(w * L) / w], [-w * math.sin(w * L), math.cos(w * L)]])
        defo = np.array([[math.cosh(w * L), math.sinh(w * L) / w], [w * math.sinh(w * L), math.cosh(w * L)]])
        mx, my = (foc, defo) if kq > 0 else (defo, foc)
        assert np.allclose(out[:2], mx @ x0[:2], rtol=1e-14, atol=1e-18)
        assert np.allclose(out[2:4], my @ x0[2:4], rtol=1e-14, atol=1e-18)
        assert abs(np.linalg.det(mx) - 1) < 1e-14 and abs(np.linalg.det(my) - 1) < 1e-14


def test_multp_is_thin_multipole_kick():
    r = np.random.default_rng(5)
    c = r.normal(0, 1e-2, (6, 100))
    for pole, kl, skew in ((1, 0.3, 0), (2, -1.7, 0), (3, 0.9, 1)):
        ob = O.Bunch(c, kin_energy=1.0)
        ob.apply("multp", pole, kl, skew)
        zn = (c[0] + 1j * c[2]) ** pole * kl / math.factorial(pole)
        if skew:
            ex, ey = zn.imag, zn.real
        else:
            ex, ey = -zn.real, zn.imag
        # kicks ~1e-7 read back as differences of ~1e-2 coordinates: ~1e-11 resolution
        assert rel_l2(ob.coords[1] - c[1], ex) < 1e-9
        assert rel_l2(ob.coords[3] - c[3], ey) < 1e-9


def test_drift_composes():
    c = np.random.default_rng(6).normal(0, 1e-3, (6, 50))
    a = O.Bunch(c, kin_energy=1.0)
    a.apply("drift", 0.3)
    a.apply("drift", 0.45)
    b = O.Bunch(c, kin_energy=1.0)
    b.apply("drift", 0.75)
    assert np.allclose(a.coords, b.coords, rtol=1e-13, atol=1e-18)
    # on-axis, on-momentum particle does not move
    z0 = O.Bunch(np.zeros((6, 1)), kin_energy=1.0)
    z0.apply("drift", 5.0)
    assert not z0.coords.any()


MAPS = [("drift", (0.7,)), ("quad1", (0.25, 0.5)), ("quad1", (0.25, -0.5)), ("quad2", (0.25,)),
        ("quadfringeIN", (0.5,)), ("quadfringeOUT", (-0.5,)), ("multp", (2, 0.4, 0)),
        ("multpfringeIN", (2, 2.0, 0)), ("multpfringeOUT", (1, -0.7, 1)),
        ("rotatexy", (0.3,)), ("kick", (1e-4, 2e-4, 1e-5)), ("bend1", (0.5, 0.1)),
        ("bend1", (1.2, -0.3)), ("bend2", (0.5,)), ("bend3", (0.05,)), ("bend4", (0.05,)),
        ("soln", (0.5, 0.8)), ("soln", (0.3, -2.0)), ("RingRF", (100.0, 3, 1e-4, 0.0, 1))]


@pytest.mark.parametrize("op,args", MAPS, ids=[m[0] + str(i) for i, m in enumerate(MAPS)])
def test_maps_are_symplectic(op, args):
    """J^T Omega J = Omega with canonical pairs (x, x'), (y, y'), (z, dE/(p beta)).  A
    dropped term or a wrong sign in any map breaks this at the 1e-4 level; the central
    differences resolve ~1e-8."""
    s = O.Bunch(np.zeros((6, 1)), kin_energy=1.0).sync
    c = 1.0 / (s["p"] * s["beta"])
    Om = np.zeros((6, 6))
    Om[0, 1], Om[1, 0], Om[2, 3], Om[3, 2], Om[4, 5], Om[5, 4] = 1, -1, 1, -1, c, -c
    for x0 in (np.array([3e-3, 1e-3, -2e-3, 5e-4, 0.1, 2e-3]),
               np.array([-1e-2, -2e-3, 4e-3, 1e-3, -0.3, -3e-3])):
        J = np.zeros((6, 6))
        for j in range(6):
            h = 1e-7 * max(abs(x0[j]), 1e-3)
            pts = np.stack([x0, x0], axis=1)
            pts[j, 0] += h
            pts[j, 1] -= h
            ob = O.Bunch(pts, kin_energy=1.0)
            ob.apply(op, *args)
            J[:, j] = (ob.coords[:, 0] - ob.coords[:, 1]) / (2 * h)
        assert np.abs(J.T @ Om @ J - Om).max() < 1e-6


def _rand6(n, seed, scale=1.0):
    r = np.random.default_rng(seed)
    c = np.empty((6, n))
    c[0] = r.normal(0, 3e-3, n) * scale
    c[1] = r.normal(0, 1e-3, n) * scale
    c[2] = r.normal(0, 3e-3, n) * scale
    c[3] = r.normal(0, 1e-3, n) * scale
    c[4] = r.uniform(-1, 1, n)
    c[5] = r.normal(0, 2e-3, n) * scale
    return c


def _apply(c, *ops):
    ob = O.Bunch(c, kin_energy=1.0)
    for op in ops:
        ob.apply(op[0], *op[1:])
    return ob


@pytest.mark.parametrize("pair", [(("wedgerotate", 0.13, 0), ("wedgerotate", 0.13, 1)),
                                  (("bendfringeIN", 4.0), ("bendfringeOUT", 4.0))])
def test_edge_maps_undo_each_other(pair):
    """wedgerotate(e, 1) after wedgerotate(e, 0) is a rotation there and back, and the
    bend fringe OUT kick cancels the IN kick (the pair BendTEAPOT applies around the
    multipole fringes)."""
    c = _rand6(2000, 1)
    # dE goes through dp/p = p0 - 1: absolute rounding ~1e-16 on dE ~1e-3
    assert np.allclose(_apply(c, *pair).coords, c, rtol=1e-13, atol=1e-14)


def test_wedgedrift_lands_on_the_edge():
    """inout 0: each particle drifts (in its own straight line, on momentum) to the edge
    line s = x tan(e); the drift length recovered from y is x_new tan(e)."""
    c = _rand6(2000, 2)
    c[5] = 0.0
    e = 0.2
    out = _apply(c, ("wedgedrift", e, 0)).coords
    s = (out[2] - c[2]) / c[3]
    assert np.allclose(s, out[0] * np.tan(e), rtol=1e-9, atol=1e-15)


def test_wedgebend_is_stepped_drift_and_curvature_kick():
    """Straight infinite-radius steps reduce to one drift of length -x tan(e) (inout 0);
    with a radius the total x' change is -s/rho, since the drifts leave x' alone."""
    c = _rand6(2000, 3)
    e, n = 0.15, 6
    flat = _apply(c, ("wedgebend", e, 0, 1e300, n)).coords
    s = -c[0] * np.tan(e)
    ref = np.empty_like(c)
    for i in range(c.shape[1]):
        ob = O.Bunch(c[:, i:i + 1], kin_energy=1.0)
        ob.apply("drift", s[i])
        ref[:, i] = ob.coords[:, 0]
    assert np.allclose(flat, ref, rtol=1e-12, atol=1e-16)
    rho = 7.0
    bent = _apply(c, ("wedgebend", e, 0, rho, n)).coords
    assert np.allclose(bent[1] - c[1], -s / rho, rtol=1e-12, atol=1e-16)
    # wedgebendCF with no multipoles is wedgebend; with a zero strength too
    cf = _apply(c, ("wedgebendCF", e, 0, rho, 1, [2], [0.0], [0], n)).coords
    assert np.array_equal(cf, bent)


def test_bend1_composes():
    """The linear sector-bend map: two halves (same radius) equal the whole."""
    c = _rand6(2000, 4)
    half = _apply(c, ("bend1", 0.6, 0.05), ("bend1", 0.6, 0.05)).coords
    whole = _apply(c, ("bend1", 1.2, 0.1)).coords
    assert np.allclose(half, whole, rtol=1e-12, atol=1e-15)


def test_soln_composes_and_tends_to_a_drift():
    c = _rand6(2000, 5)
    two = _apply(c, ("soln", 0.2, 1.3), ("soln", 0.5, 1.3)).coords
    one = _apply(c, ("soln", 0.7, 1.3)).coords
    assert np.allclose(two, one, rtol=1e-11, atol=1e-15)
    # the departure from a drift shrinks linearly with the field
    drift = _apply(c, ("drift", 0.7)).coords
    d3 = np.abs(_apply(c, ("soln", 0.7, 1e-3)).coords - drift).max()
    d4 = np.abs(_apply(c, ("soln", 0.7, 1e-4)).coords - drift).max()
    assert d3 / d4 == pytest.approx(10.0, rel=0.05)
    # below 1e-100 the reference switches to drift exactly
    assert np.array_equal(_apply(c, ("soln", 0.7, 1e-101)).coords, drift)


def test_ringrf_closed_form():
    """dE += V sin(2 pi h z / L + phi); a non-zero phi raises the synchronous energy by
    V sin(phi), and x', y' scale by p_in / p_out."""
    c = _rand6(500, 6)
    L, h, V, phi = 248.0, 4, 2e-5, 0.3
    ob = O.Bunch(c, kin_energy=1.0)
    s0 = ob.sync
    ob.apply("RingRF", L, h, V, phi, 1)
    s1 = ob.sync
    m = s0["mass"]
    T1 = s0["T"] + V * math.sin(phi)
    p1 = math.sqrt(T1 * (T1 + 2 * m))
    assert s1["T"] == pytest.approx(T1, rel=1e-14)
    assert s1["p"] == pytest.approx(p1, rel=1e-14)
    assert np.allclose(ob.coords[5], c[5] + V * np.sin(2 * np.pi * h * c[4] / L + phi),
                       rtol=1e-13, atol=1e-19)
    assert np.allclose(ob.coords[1], c[1] * s0["p"] / p1, rtol=1e-14, atol=0)


# ---------------------------------------------------------------- space-charge kicks

def _uniform_disk(n, R, seed):
    r = np.random.default_rng(seed)
    rad = R * np.sqrt(r.random(n))
    th = 2 * np.pi * r.random(n)
    return rad * np.cos(th), rad * np.sin(th)


def test_sc2p5drb_transverse_kick_uniform_disk():
    """Uniform round beam of radius R, uniform in z: inside, x'' kick = 2 r0 lambda L x /
    (beta^2 gamma^3 R^2) (defocusing), lambda = particles per metre."""
    n, R, Lz, L = 400000, 5e-3, 20.0, 0.5
    x, y = _uniform_disk(n, R, 1)
    c = np.zeros((6, n))
    c[0], c[2] = x, y
    c[4] = np.random.default_rng(2).uniform(-Lz / 2, Lz / 2, n)
    ms = 1e13 / n
    ob = O.Bunch(c, macrosize=ms, kin_energy=1.0)
    calc = O.SC2p5Drb(128, 128, 16)
    calc.trackBunch(ob, L, 0.1)
    s = ob.sync
    lam = 1e13 / Lz
    K = 2 * 1.534698e-18 * lam * L / (s["beta"] ** 2 * s["gamma"] ** 3 * R ** 2)
    inside = np.hypot(x, y) < 0.7 * R
    dxp = ob.coords[1] - c[1]
    assert np.median(dxp[inside] / (K * x[inside])) == pytest.approx(1.0, abs=0.03)
    # a_bunch of a uniform disk is its radius: a^2 = 2 (sx^2 + sy^2) = R^2
    assert calc.moments()[2] == pytest.approx(R, rel=0.02)


def test_sc2p5drb_longitudinal_kick_linear_ramp():
    """Line density with a linear slope: dE kick = dlambda/dz * (-L r0 m / gamma^2) *
    (1 + 2 ln(b/a) - r^2/a^2) inside the beam (Baartman's g-factor)."""
    n, R, L, b = 400000, 5e-3, 0.5, 0.05
    rng = np.random.default_rng(3)
    x, y = _uniform_disk(n, R, 4)
    # density proportional to (1 + s z) on [-1, 1], s = 0.5, by inverse CDF
    s_ = 0.5
    u = rng.random(n)
    z = (-1 + np.sqrt(1 - 2 * s_ * (1 - s_ / 2) + 4 * s_ * u * 1.0 + 0 * u)) / s_
    z = np.clip(z, -1, 1)
    c = np.zeros((6, n))
    c[0], c[2], c[4] = x, y, z
    ob = O.Bunch(c, macrosize=1e13 / n, kin_energy=1.0)
    calc = O.SC2p5Drb(64, 64, 32)
    calc.trackBunch(ob, L, b)
    g = ob.sync["gamma"]
    dlam = 1e13 / 2.0 * s_  # d/dz of 1e13 (1 + s z)/2
    r2 = (x - calc.moments()[0]) ** 2 + (y - calc.moments()[1]) ** 2
    a = calc.moments()[2]
    expect = dlam * (-L * 1.534698e-18 * 0.938272013 / g ** 2) * (1 + 2 * np.log(b / a) - r2 / a ** 2)
    core = (np.hypot(x, y) < 0.7 * R) & (np.abs(z) < 0.6)
    ratio = (ob.coords[5] - c[5])[core] / expect[core]
    assert np.median(ratio) == pytest.approx(1.0, abs=0.05)


def test_sc2p5d_transverse_kick_uniform_disk():
    """SpaceChargeCalc2p5D: the same defocusing kick of a uniform round beam; dE is left
    alone and the total macrosize is the sum of the z line density (all particles)."""
    n, R, Lz, L = 400000, 5e-3, 20.0, 0.5
    x, y = _uniform_disk(n, R, 5)
    c = np.zeros((6, n))
    c[0], c[2] = x, y
    c[4] = np.random.default_rng(6).uniform(-Lz / 2, Lz / 2, n)
    c[5] = 1e-4
    ob = O.Bunch(c, macrosize=1e13 / n, kin_energy=1.0)
    calc = O.SC2p5D(128, 128, 16)
    calc.trackBunch(ob, L)
    s = ob.sync
    K = 2 * 1.534698e-18 * (1e13 / Lz) * L / (s["beta"] ** 2 * s["gamma"] ** 3 * R ** 2)
    inside = np.hypot(x, y) < 0.7 * R
    dxp = ob.coords[1] - c[1]
    dyp = ob.coords[3] - c[3]
    assert np.median(dxp[inside] / (K * x[inside])) == pytest.approx(1.0, abs=0.03)
    assert np.median(dyp[inside] / (K * y[inside])) == pytest.approx(1.0, abs=0.03)
    assert np.all(ob.coords[5] == c[5])
    assert calc.state()[2] == pytest.approx(1e13, rel=1e-12)


def test_boundary_potential_is_constant_on_a_conducting_pipe():
    """An off-centre line charge inside a circular pipe: after addBoundaryPotential the
    potential on the pipe is (nearly) constant, i.e. the pipe is an equipotential."""
    bnd = O.Boundary2D(128, 16, "Circle", 0.04)
    bx, by, lims = bnd.info()
    nx = ny = 129
    xs = np.linspace(lims[0], lims[1], nx)
    ys = np.linspace(lims[2], lims[3], ny)
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    phi = -0.5 * np.log((X - 0.006) ** 2 + (Y + 0.004) ** 2 + 1e-12)  # free space, off centre
    out = bnd.addBoundaryPotential(phi, lims)

    def on_pipe(a):
        return np.array([O.Grid2DValue(a, lims, x, y) for x, y in zip(bx, by)])
    before, after = on_pipe(phi), on_pipe(out)
    assert np.ptp(after) < 1e-2 * np.ptp(before)


def test_sc2p5d_centred_beam_in_round_pipe_keeps_its_field():
    """A centred round beam in a round pipe: the pipe only adds a constant potential, so
    the kick is the free-space uniform-disk kick 2 r0 lambda L x / (beta^2 gamma^3 R^2)."""
    n, R = 200000, 5e-3
    x, y = _uniform_disk(n, R, 9)
    c = np.zeros((6, n))
    c[0], c[2] = x, y
    c[4] = np.random.default_rng(10).uniform(-5, 5, n)
    bnd = O.Boundary2D(128, 12, "Circle", 0.02)
    a = O.Bunch(c, macrosize=1e9, kin_energy=1.0)
    O.SC2p5D(64, 64, 8).trackBunch(a, 0.5, bnd)
    s = a.sync
    K = 2 * 1.534698e-18 * (1e9 * n / 10.0) * 0.5 / (s["beta"] ** 2 * s["gamma"] ** 3 * R ** 2)
    inside = np.hypot(x, y) < 0.7 * R
    assert np.median((a.coords[1] - c[1])[inside] / (K * x[inside])) == pytest.approx(1.0, abs=0.04)


def test_slice2d_kick_uniform_disk_and_local_line_density():
    """SpaceChargeCalcSliceBySlice2D: a uniform round beam whose line density rises
    linearly in z is kicked like the uniform disk of the local density,
    2 r0 lambda(z) L x / (beta^2 gamma^3 R^2); dE is left alone."""
    n, R, L = 600000, 5e-3, 0.5
    x, y = _uniform_disk(n, R, 11)
    r = np.random.default_rng(12)
    s_ = 0.5  # density proportional to 1 + s z on [-1, 1]
    u = r.random(n)
    z = (-1 + np.sqrt((1 - s_) ** 2 + 4 * s_ * u)) / s_
    c = np.zeros((6, n))
    c[0], c[2], c[4] = x, y, z
    c[5] = 1e-4
    Q = 1e13
    ob = O.Bunch(c, macrosize=Q / n, kin_energy=1.0)
    calc = O.SliceBySlice2D(64, 64, 32)
    calc.trackBunch(ob, L)
    sy = ob.sync
    lam = Q / 2.0 * (1 + s_ * z)
    K = 2 * 1.534698e-18 * lam * L / (sy["beta"] ** 2 * sy["gamma"] ** 3 * R ** 2)
    core = (np.hypot(x, y) < 0.7 * R) & (np.abs(z) < 0.8)
    dxp = ob.coords[1] - c[1]
    assert np.median(dxp[core] / (K[core] * x[core])) == pytest.approx(1.0, abs=0.04)
    assert np.all(ob.coords[5] == c[5]) and np.all(ob.coords[4] == c[4])


def test_slice2d_ratio_rule_matches_2p5d():
    """The slice calculator sizes its transverse grid like SpaceChargeCalc2p5D: the same
    beams give the same box, ratio and rebuild count."""
    a, b = O.SliceBySlice2D(32, 32, 8, 1.0), O.SC2p5D(32, 32, 8, 1.0)
    for k, (ax, ay) in enumerate([(2e-3, 1.8e-3), (5e-3, 1.5e-3), (5e-3, 1.5e-3), (2e-3, 2e-3)]):
        c = _beam(5000, ax, ay, 50 + k)
        a.trackBunch(O.Bunch(c, macrosize=1e9, kin_energy=1.0), 0.1)
        b.trackBunch(O.Bunch(c, macrosize=1e9, kin_energy=1.0), 0.1)
        sa, sb = a.state(), b.state()
        assert sa[0] == sb[0] and sa[1] == sb[1] and sa[2] == sb[3]
        assert tuple(a.grid(0)[1][:4]) == tuple(b.grid(0)[1])


def _beam(n, ax, ay, seed):
    r = np.random.default_rng(seed)
    c = np.zeros((6, n))
    c[0] = r.normal(0, ax, n)
    c[1] = r.normal(0, 1e-3, n)
    c[2] = r.normal(0, ay, n)
    c[3] = r.normal(0, 1e-3, n)
    c[4] = r.uniform(-5, 5, n)
    c[5] = r.normal(0, 1e-3, n)
    return c


def test_sc2p5d_ratio_rule_and_green_rebuild():
    """Beam ratio within 25% of the calculator's: the box keeps the ratio and contains the
    beam, no rebuild.  Outside: the box is the beam's own, the ratio follows it and the
    solver is rebuilt on that box once; a repeat of the same beam rebuilds nothing."""
    calc = O.SC2p5D(64, 64, 8, 1.0)
    c = _beam(20000, 2e-3, 1.8e-3, 1)
    ob = O.Bunch(c, macrosize=1e9, kin_energy=1.0)
    calc.trackBunch(ob, 0.1)
    ratio, sol, total, rebuilds = calc.state()
    _, lims = calc.grid(0)
    e = c[[0, 0, 2, 2]].copy()
    assert rebuilds == 0 and sol == (-1.0, 1.0, -1.0, 1.0)
    assert ratio == pytest.approx(1.0, rel=1e-14)
    assert (lims[1] - lims[0]) / (lims[3] - lims[2]) == pytest.approx(1.0, rel=1e-14)
    assert lims[0] <= c[0].min() and lims[1] >= c[0].max()
    assert lims[2] <= c[2].min() and lims[3] >= c[2].max()
    c2 = _beam(20000, 4e-3, 1.5e-3, 2)
    ob2 = O.Bunch(c2, macrosize=1e9, kin_energy=1.0)
    calc.trackBunch(ob2, 0.1)
    ratio, sol, total, rebuilds = calc.state()
    _, lims = calc.grid(0)
    want = (c2[0].min(), c2[0].max(), c2[2].min(), c2[2].max())
    assert rebuilds == 1
    assert tuple(lims) == pytest.approx(want, rel=0, abs=0)
    assert sol == pytest.approx(want, rel=0, abs=0)
    assert ratio == (want[1] - want[0]) / (want[3] - want[2])
    calc.trackBunch(O.Bunch(c2, macrosize=1e9, kin_energy=1.0), 0.1)
    assert calc.state()[3] == 1


def test_sc2p5d_transverse_kick_equals_rb_on_the_same_box():
    """With the calculator's ratio set to the beam's, both calculators use the same box;
    the potential differs only by a constant (the solver box scale), so the transverse
    kicks agree to rounding (the rb kick is pinned by the uniform-disk test)."""
    c = _beam(30000, 3e-3, 2e-3, 3)
    ratio = (c[0].max() - c[0].min()) / (c[2].max() - c[2].min())
    a = O.Bunch(c, macrosize=1e9, kin_energy=1.0)
    b = O.Bunch(c, macrosize=1e9, kin_energy=1.0)
    O.SC2p5D(64, 64, 16, ratio).trackBunch(a, 0.3)
    rb = O.SC2p5Drb(64, 64, 16, ratio)
    rb.trackBunch(b, 0.3, 0.1)
    for k in (1, 3):
        d2, drb = a.coords[k] - c[k], b.coords[k] - c[k]
        assert rel_l2(d2, drb) < 1e-10
    assert np.all(a.coords[5] == c[5])


def test_sc3d_kick_uniform_sphere():
    """Uniform sphere of radius R in the rest frame (z semi-axis R/gamma in the lab): inside,
    x' kick = Q r0 L x / (R^3 beta^2 gamma^2), dE kick = Q r0 m L z_rest / R^3."""
    n, R, L = 300000, 4e-3, 0.5
    rng = np.random.default_rng(5)
    v = rng.normal(size=(3, n))
    v *= R * rng.random(n) ** (1 / 3) / np.sqrt((v ** 2).sum(0))
    s = O.Bunch(np.zeros((6, 1)), kin_energy=1.0).sync
    g, b = s["gamma"], s["beta"]
    c = np.zeros((6, n))
    c[0], c[2], c[4] = v[0], v[1], v[2] / g
    ob = O.Bunch(c, macrosize=1e11 / n, kin_energy=1.0)
    calc = O.SC3D(48, 48, 48)
    calc.trackBunch(ob, L)
    Q = 1e11
    inside = np.sqrt((v ** 2).sum(0)) < 0.6 * R
    kx = Q * 1.534698e-18 * L / (R ** 3 * b ** 2 * g ** 2)
    assert np.median((ob.coords[1] - c[1])[inside] / (kx * v[0][inside])) == pytest.approx(1.0, abs=0.05)
    kz = Q * 1.534698e-18 * 0.938272013 * L / R ** 3
    assert np.median((ob.coords[5] - c[5])[inside] / (kz * v[2][inside])) == pytest.approx(1.0, abs=0.05)


# ---------------------------------------------------------------- diagnostics

def test_extrema_and_emittance_vs_numpy():
    c = np.random.default_rng(9).normal(0, 1e-3, (6, 1000))
    c[5] *= 0.0  # no dispersion term
    ob = O.Bunch(c, kin_energy=1.0)
    e = ob.extrema()
    assert list(e) == [c[0].min(), c[0].max(), c[2].min(), c[2].max(), c[4].min(), c[4].max()]
    em = ob.emittances()
    for k in range(2):
        cov = np.cov(c[2 * k], c[2 * k + 1], bias=True)
        assert em[k] == pytest.approx(math.sqrt(np.linalg.det(cov)), rel=1e-9)


# ---------------------------------------------------------------- quirk pins

def test_sc2p5drb_moments_use_grid_x_for_y():
    """SpaceChargeCalc2p5Drb.cc:229-254 forms the y moments with getGridX(iy) (:238): node
    iy sits at xMin + iy dx for them.  The TSC weights of a particle at fraction f from
    its node reproduce its position (sum w k = f) and add a quarter cell to its variance
    (sum w k^2 = 1/4 + f^2), also at the clamped edge nodes, so for a beam inside the grid
        x_c = <x>,    y_c = xMin + (dx/dy) (<y> - yMin),
        a^2 = 2 (var x + dx^2/4 + (dx/dy)^2 (var y + dy^2/4)).
    An off-centre beam on a 64 x 40 grid (dx != dy, xMin != yMin) tells this apart from
    y_c = <y> by millimetres."""
    c = _beam(20000, 2e-3, 1e-3, 7)
    c[0] += 3e-3
    c[2] -= 4e-3
    calc = O.SC2p5Drb(64, 40, 8, 1.0)
    calc.trackBunch(O.Bunch(c, macrosize=1e9, kin_energy=1.0), 0.1, 0.1)
    _, (xmin, xmax, ymin, ymax) = calc.grid(0)
    dx, dy = (xmax - xmin) / 63, (ymax - ymin) / 39
    assert abs(dx / dy - 1.0) > 0.3 and abs(xmin - ymin) > 1e-3
    xm, ym = c[0].mean(), c[2].mean()
    x_c, y_c, a, total = calc.moments()
    assert total == pytest.approx(20000 * 1e9, rel=1e-13)
    assert x_c == pytest.approx(xm, rel=0, abs=1e-14 * (xmax - xmin))
    want_yc = xmin + dx / dy * (ym - ymin)
    assert y_c == pytest.approx(want_yc, rel=0, abs=1e-14 * (xmax - xmin))
    assert abs(y_c - ym) > 1e-3
    want_a2 = 2 * (c[0].var() + dx * dx / 4 + (dx / dy) ** 2 * (c[2].var() + dy * dy / 4))
    assert a * a == pytest.approx(want_a2, rel=1e-10)


def _box_beam(ax, ay, az_rest, seed, n=4000):
    """A beam whose extrema are exactly +-ax, +-ay and +-az_rest / gamma in lab z."""
    g = O.Bunch(np.zeros((6, 1)), kin_energy=1.0).sync["gamma"]
    r = np.random.default_rng(seed)
    c = np.zeros((6, n))
    c[0] = r.uniform(-ax, ax, n)
    c[2] = r.uniform(-ay, ay, n)
    c[4] = r.uniform(-az_rest / g, az_rest / g, n)
    c[1] = r.normal(0, 1e-3, n)
    c[3] = r.normal(0, 1e-3, n)
    c[5] = r.normal(0, 1e-3, n)
    for k, (i, a) in enumerate([(0, ax), (2, ay), (4, az_rest / g)]):
        c[i, 2 * k], c[i, 2 * k + 1] = -a, a
    return c


def test_sc3d_green_rebuild_sequence():
    """SpaceChargeCalc3D.cc:186-203 rebuilds the Green function when the beam's x/y or x/z
    (rest frame) extent ratio leaves [1/limit, limit] times the ratio of the last rebuild;
    the solver starts at ratio 1 with limit 1.2 (:51).  Beams with hand-chosen extents and
    the rebuild count each must give, derived by hand:
      (2, 2, 2) mm       xy 1, xz 1 against 1, 1: inside                        -> 0
      (2.2, 2, 2)        1.1, 1.1: inside                                        -> 0
      (2.6, 2, 2)        1.3 > 1.2: rebuild, ratios now 1.3, 1.3                 -> 1
      (2.6, 1.8, 2.3)    1.444/1.3 = 1.11, 1.130/1.3 = 0.87 > 1/1.2: inside      -> 1
      (2.6, 2, 2.9)      xz 0.897/1.3 = 0.69 < 0.833: rebuild (xz alone)         -> 2
      limit 2, (2.6, 1.2, 2.9)  xy 2.167/1.3 = 1.67 < 2: inside                 -> 2
      limit 2, (2.6, 0.6, 2.9)  xy 4.333/1.3 = 3.33 > 2: rebuild                -> 3
    Right after a rebuild the rho grid is the beam's own box; otherwise it keeps the
    solver's aspect and covers the beam."""
    g = O.Bunch(np.zeros((6, 1)), kin_energy=1.0).sync["gamma"]
    calc = O.SC3D(16, 16, 16)
    seq = [((2.0, 2.0, 2.0), 1.2, 0), ((2.2, 2.0, 2.0), 1.2, 0), ((2.6, 2.0, 2.0), 1.2, 1),
           ((2.6, 1.8, 2.3), 1.2, 1), ((2.6, 2.0, 2.9), 1.2, 2), ((2.6, 1.2, 2.9), 2.0, 2),
           ((2.6, 0.6, 2.9), 2.0, 3)]
    last_box = (2.0, 2.0, 2.0)  # the initial solver box has unit ratios
    for k, ((ax, ay, az), limit, want) in enumerate(seq):
        calc.setRatioLimit(limit)
        c = _box_beam(ax * 1e-3, ay * 1e-3, az * 1e-3, 60 + k)
        calc.trackBunch(O.Bunch(c, macrosize=1e9, kin_energy=1.0), 0.1)
        assert calc.green_rebuilds() == want, k
        _, L = calc.grid(0)
        wx, wy, wz = L[1] - L[0], L[3] - L[2], L[5] - L[4]
        rebuilt = k == 0 and want == 1 or k > 0 and want != seq[k - 1][2]
        if rebuilt:
            last_box = (ax, ay, az)
            assert (wx, wy, wz) == pytest.approx((2 * ax * 1e-3, 2 * ay * 1e-3, 2 * az * 1e-3),
                                                 rel=1e-12)
        else:
            assert wx / wy == pytest.approx(last_box[0] / last_box[1], rel=1e-12)
            assert wx / wz == pytest.approx(last_box[0] / last_box[2], rel=1e-12)
            assert wx >= 2 * ax * 1e-3 * (1 - 1e-12) and wy >= 2 * ay * 1e-3 * (1 - 1e-12)
            assert wz >= 2 * az * 1e-3 * (1 - 1e-12)
    assert g > 2
