"""The oracle's Grid1D/2D/3D restatement vs the reference's own Grid classes compiled
from its sources (oracle/_ref): bit-exact indices, fractions, deposits, values and
gradients, including particles exactly on grid edges and nodes, and outside the grid."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built here")


def _pts(n, lo, hi, seed, nodes=None):
    r = np.random.default_rng(seed)
    v = r.uniform(lo, hi, n)
    v[:4] = [lo, hi, lo - 1e-3 * (hi - lo), hi + 1e-3 * (hi - lo)]  # edges and outside
    if nodes is not None:
        v[4:4 + len(nodes)] = nodes
    return v


@pytest.mark.parametrize("nx,ny", [(3, 3), (16, 16), (33, 20), (128, 128)])
def test_grid2d(nx, ny):
    lims = np.array([-3e-3, 4e-3, -2e-3, 2.5e-3])
    dx = (lims[1] - lims[0]) / (nx - 1)
    n = 4000
    x = _pts(n, lims[0], lims[1], 1, nodes=lims[0] + dx * np.arange(min(nx, 50)))
    y = _pts(n, lims[2], lims[3], 2)
    m = np.random.default_rng(3).uniform(0.5, 1.5, n)
    ref, orc = O.ref_grid(), O.orc_grid()
    assert np.array_equal(ref.grid2d_bin(nx, ny, lims, x, y, m), orc.grid2d_bin(nx, ny, lims, x, y, m))
    for a, b in zip(ref.grid2d_indfrac(nx, ny, lims, x, y), orc.grid2d_indfrac(nx, ny, lims, x, y)):
        assert np.array_equal(a, b)
    arr = np.random.default_rng(4).normal(size=(nx, ny))
    for a, b in zip(ref.grid2d_grad(nx, ny, lims, arr, x, y), orc.grid2d_grad(nx, ny, lims, arr, x, y)):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("nz", [1, 2, 3, 8, 64])
def test_grid1d(nz):
    zmin, zmax = -100.0, 100.0
    dz = (zmax - zmin) / nz
    n = 3000
    z = _pts(n, zmin, zmax, 5, nodes=zmin + dz * (np.arange(nz) + 0.5))
    m = np.random.default_rng(6).uniform(0.5, 1.5, n)
    ref, orc = O.ref_grid(), O.orc_grid()
    assert np.array_equal(ref.grid1d_bin(nz, zmin, zmax, z, m), orc.grid1d_bin(nz, zmin, zmax, z, m))
    arr = np.random.default_rng(7).normal(size=nz)
    for a, b in zip(ref.grid1d_value(nz, zmin, zmax, arr, z), orc.grid1d_value(nz, zmin, zmax, arr, z)):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("nx,ny,nz", [(3, 3, 1), (8, 8, 2), (8, 9, 3), (16, 16, 16), (17, 24, 9)])
def test_grid3d(nx, ny, nz):
    lims = np.array([-3e-3, 3e-3, -2e-3, 2e-3, -0.5, 0.7])
    n = 3000
    x = _pts(n, lims[0], lims[1], 8)
    y = _pts(n, lims[2], lims[3], 9)
    z = _pts(n, lims[4], lims[5], 10)
    m = np.random.default_rng(11).uniform(0.5, 1.5, n)
    ref, orc = O.ref_grid(), O.orc_grid()
    assert np.array_equal(ref.grid3d_bin(nx, ny, nz, lims, x, y, z, m),
                          orc.grid3d_bin(nx, ny, nz, lims, x, y, z, m))
    for a, b in zip(ref.grid3d_indfrac(nx, ny, nz, lims, x, y, z), orc.grid3d_indfrac(nx, ny, nz, lims, x, y, z)):
        assert np.array_equal(a, b)
    arr = np.random.default_rng(12).normal(size=(nz, nx, ny))
    for a, b in zip(ref.grid3d_grad(nx, ny, nz, lims, arr, x, y, z),
                    orc.grid3d_grad(nx, ny, nz, lims, arr, x, y, z)):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("nz", [2, 3, 8, 24])
def test_grid3d_long_wrapping(nz):
    """longWrapping = 1 (the external-bunch mode of SpaceChargeCalc3D): z is wrapped into one
    bunch spacing before binning, the z index clamps to [0, nz-1] and the stencil is
    periodic.  Bit-exact vs the reference's Grid3D with setLongWrapping(1)."""
    nx, ny = 10, 9
    lam = 2.0
    lims = np.array([-3e-3, 3e-3, -2e-3, 2e-3, -1.0001, 1.0001])
    n = 3000
    r = np.random.default_rng(11)
    x = _pts(n, lims[0], lims[1], 5)
    y = _pts(n, lims[2], lims[3], 6)
    z = r.uniform(-5.0, 5.0, n)  # several spacings away: remainder brings them back
    z[:3] = [0.9999, -0.9999, 3.0]
    m = r.uniform(0.5, 1.5, n)
    ref, orc = O.ref_grid(), O.orc_grid()
    a = ref.grid3d_bin_wrapped(nx, ny, nz, lims, x, y, z, m, lam)
    b = orc.grid3d_bin_wrapped(nx, ny, nz, lims, x, y, z, m, lam)
    assert np.array_equal(a, b)
    arr = r.normal(size=(nz, nx, ny))
    zz = r.uniform(lims[4], lims[5], n)
    for u, v in zip(ref.grid3d_grad_wrapped(nx, ny, nz, lims, arr, x, y, zz),
                    orc.grid3d_grad_wrapped(nx, ny, nz, lims, arr, x, y, zz)):
        assert np.array_equal(u, v)


BOUNDARIES = [("Circle", 128, 10, 0.02, 0.02), ("Ellipse", 130, 12, 0.03, 0.018),
              ("Rectangle", 100, 8, 0.04, 0.025), (None, 64, 6, 0.0, 0.0)]


@pytest.mark.parametrize("shape,npts,nmodes,xd,yd", BOUNDARIES,
                         ids=[str(b[0]) for b in BOUNDARIES])
def test_boundary2d(shape, npts, nmodes, xd, yd):
    """Base/ShapedBoundary2D: boundary points, limits, isInside and addBoundaryPotential
    bit for bit against the reference build, for the three shapes and a user-defined
    polygon."""
    r = np.random.default_rng(7)
    pts = None
    if shape is None:  # a lopsided polygon of user points
        t = np.linspace(0, 2 * np.pi, npts, endpoint=False)
        rad = 0.015 * (1 + 0.2 * np.cos(3 * t))
        pts = (rad * np.cos(t) + 1e-3, rad * np.sin(t) - 2e-3)
    tx = r.uniform(-0.03, 0.03, 3000)
    ty = r.uniform(-0.03, 0.03, 3000)
    nx, ny = 40, 33
    rbx, rby, rlims, rin, _ = O.ref_boundary2d(shape, npts, nmodes, xd, yd, points=pts,
                                                tests=(tx, ty))
    glims = np.array([rlims[0], rlims[1], rlims[2], rlims[3]])
    phi = r.normal(size=(nx, ny))
    _, _, _, _, rphi = O.ref_boundary2d(shape, npts, nmodes, xd, yd, points=pts,
                                         tests=(tx, ty), phi=phi, glims=glims)
    if shape is None:
        b = O.Boundary2D(npts, nmodes)
        for i in range(npts):
            b.setBoundaryPoint(i, pts[0][i], pts[1][i])
        b.initialize()
    else:
        b = O.Boundary2D(npts, nmodes, shape, xd, yd)
    bx, by, lims = b.info()
    assert np.array_equal(bx, rbx) and np.array_equal(by, rby)
    assert np.array_equal(lims, rlims)
    assert [b.isInside(x, y) for x, y in zip(tx, ty)] == list(rin)
    assert np.array_equal(b.addBoundaryPotential(phi, glims), rphi)
