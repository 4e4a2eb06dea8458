"""SpaceChargeCalc2p5D (no boundary) on the device vs the oracle: grids, the x/y ratio
rule with its on-device Green function rebuilds, kicks, and lattice tracking.  Bars as for
the 2p5Drb calculator: rho, phi and the line density within 1e-12 relative L2, bin indices
exact on non-overlapping stencils, transverse kicks within 1e-10 relative L2, and x, y, z,
dE untouched bit for bit."""
import numpy as np
import pytest

import oracle as O
from helpers import rel_l2, trace_lattice, replay_oracle
from test_gpu_sc import _bunches, sparse_coords

pytestmark = pytest.mark.gpu


def _coords(n, ax, ay, seed, zhalf=50.0):
    r = np.random.default_rng(seed)
    c = np.empty((6, n))
    c[0] = r.normal(2e-4, ax, n)
    c[1] = r.normal(0, 1e-3, n)
    c[2] = r.normal(-1e-4, ay, n)
    c[3] = r.normal(0, 1e-3, n)
    c[4] = r.uniform(-zhalf, zhalf, n)
    c[5] = r.normal(0, 1e-3, n)
    return c


def _compare(calc, ocalc, b, ob, c):
    out = b.getCoordinates()
    rho, orho = calc.getRhoGrid(), ocalc.grid(0)
    assert [rho.getMinX(), rho.getMaxX(), rho.getMinY(), rho.getMaxY()] == list(orho[1])
    assert rel_l2(rho.array(), orho[0]) < 1e-12
    assert rel_l2(calc.getPhiGrid().array(), ocalc.grid(1)[0]) < 1e-12
    zg = calc.getLongGrid()
    assert rel_l2(zg.array(), ocalc.grid(2)[0]) < 1e-12
    assert [zg.getMinZ(), zg.getMaxZ()] == list(ocalc.grid(2)[1])
    ratio, sol, total, rebuilds = calc.getState()
    oratio, osol, ototal, orebuilds = ocalc.state()
    assert ratio == oratio and sol == osol and rebuilds == orebuilds
    assert total == pytest.approx(ototal, rel=1e-13)
    for k in (1, 3):
        assert rel_l2(out[k] - c[k], ob.coords[k] - c[k]) < 1e-10, k
    for k in (0, 2, 4, 5):
        assert np.array_equal(out[k], c[k]), k


@pytest.mark.parametrize("grid", [(16, 16, 8), (33, 20, 5), (128, 128, 64)])
def test_sc2p5d_step(grid):
    import spacecharge
    c = _coords(20000, 3e-3, 2.7e-3, seed=11)
    b, ob = _bunches(c)
    calc = spacecharge.SpaceChargeCalc2p5D(*grid)
    ocalc = O.SC2p5D(*grid)
    calc.trackBunch(b, 0.5)
    ocalc.trackBunch(ob, 0.5)
    _compare(calc, ocalc, b, ob, c)
    assert calc.getState()[3] == 0


def test_sc2p5d_ratio_rule_over_steps():
    """Beams whose x/y ratio changes: kept ratio (no rebuild), a wide beam (rebuild on its
    box), the same beam again (none), a round beam again (rebuild), a tall beam passed
    with a changed calculator ratio... every step against the oracle's own state."""
    import spacecharge
    calc = spacecharge.SpaceChargeCalc2p5D(64, 48, 16, 1.2)
    ocalc = O.SC2p5D(64, 48, 16, 1.2)
    shapes = [(3e-3, 2.7e-3), (6e-3, 2e-3), (6e-3, 2e-3), (2.5e-3, 2.5e-3), (1e-3, 4e-3),
              (1.1e-3, 4e-3)]
    expect_rebuilds = [0, 1, 1, 2, 3, 3]
    for k, (ax, ay) in enumerate(shapes):
        c = _coords(30000, ax, ay, seed=100 + k)
        b, ob = _bunches(c)
        calc.trackBunch(b, 0.4)
        ocalc.trackBunch(ob, 0.4)
        _compare(calc, ocalc, b, ob, c)
        assert calc.getState()[3] == expect_rebuilds[k], k


def test_sc2p5d_bin_indices_exact():
    import spacecharge
    c = sparse_coords()
    b, ob = _bunches(c)
    calc = spacecharge.SpaceChargeCalc2p5D(64, 64, 200)
    ocalc = O.SC2p5D(64, 64, 200)
    calc.trackBunch(b, 0.5)
    ocalc.trackBunch(ob, 0.5)
    assert np.array_equal(calc.getRhoGrid().array(), ocalc.grid(0)[0])
    assert np.array_equal(calc.getLongGrid().array(), ocalc.grid(2)[0])


def test_sc2p5d_macrosize_attribute():
    import spacecharge
    c = _coords(5000, 2e-3, 2e-3, seed=5)
    m = np.random.default_rng(2).uniform(0.5e8, 1.5e8, c.shape[1])
    b, _ = _bunches(c)
    b.setMacroSizes(m)
    ob = O.Bunch(c, macrosize=1e8, kin_energy=1.0, msize=m)
    calc = spacecharge.SpaceChargeCalc2p5D(32, 32, 16, 1.5)
    ocalc = O.SC2p5D(32, 32, 16, 1.5)
    calc.trackBunch(b, 0.3)
    ocalc.trackBunch(ob, 0.3)
    _compare(calc, ocalc, b, ob, c)
    assert calc.getState()[2] == pytest.approx(m.sum(), rel=1e-12)


def test_sc2p5d_zero_extent_raises_and_bad_boundary_refused():
    import spacecharge
    from _obnative import OrbitError
    c = np.zeros((6, 10))
    c[0] = np.linspace(-1e-3, 1e-3, 10)
    b, _ = _bunches(c)
    calc = spacecharge.SpaceChargeCalc2p5D(16, 16, 8)
    calc.trackBunch(b, 0.5)
    with pytest.raises(OrbitError, match="zero extent"):
        b.getCoordinates()
    with pytest.raises(TypeError):
        calc.trackBunch(b, 0.5, object())


def test_sc2p5d_calculator_refused_by_batches():
    import spacecharge
    from batch import BunchBatch
    from _obnative import OrbitError
    from workloads import gaussian_coords, make_bunch
    bunches = [make_bunch(gaussian_coords(1000, seed=s)) for s in (1, 2)]
    calcs = [spacecharge.SpaceChargeCalc2p5D(16, 16, 8) for _ in bunches]
    with pytest.raises((OrbitError, TypeError)):
        BunchBatch(bunches, calcs)


def _fodo_with_sc2p5d(calc, ncells):
    from orbit.space_charge.sc2p5d import setSC2p5DAccNodes
    from workloads import fodo_lattice
    lat = fodo_lattice(ncells)
    nodes = setSC2p5DAccNodes(lat, 0.49, calc)
    assert all(nd.getType() == "SC2p5D" and nd.getBoundary() is None for nd in nodes)
    return lat


def test_fodo_2p5d_tracking(monkeypatch):
    import spacecharge
    from workloads import gaussian_coords, make_bunch
    n, grid, cells = 50000, (64, 64, 32), 4
    calls = trace_lattice(monkeypatch, lambda calc: _fodo_with_sc2p5d(calc, cells), "sc2p5d")
    monkeypatch.undo()
    assert sum(1 for c in calls if c[0] == "sc2p5d") == 12 * cells
    c = gaussian_coords(n, seed=8)
    b = make_bunch(c)
    lat = _fodo_with_sc2p5d(spacecharge.SpaceChargeCalc2p5D(*grid), cells)
    lat.trackBunch(b)
    out = b.getCoordinates()
    ob = O.Bunch(c, macrosize=1e13 / n, kin_energy=1.0)
    replay_oracle(calls, ob, O.SC2p5D(*grid))
    for k in range(6):
        assert rel_l2(out[k], ob.coords[k]) < 1e-9, (k, rel_l2(out[k], ob.coords[k]))


BOUNDS = [("Circle", 128, 10, 0.02, 0.02), ("Ellipse", 128, 12, 0.03, 0.018),
          ("Rectangle", 120, 8, 0.04, 0.025)]


def _both_boundaries(shape, npts, nm, xd, yd):
    import spacecharge
    return (spacecharge.Boundary2D(npts, nm, shape, xd, yd), O.Boundary2D(npts, nm, shape, xd, yd))


@pytest.mark.parametrize("shape,npts,nm,xd,yd", BOUNDS, ids=[b[0] for b in BOUNDS])
def test_boundary2d_matches_oracle(shape, npts, nm, xd, yd):
    """Points, limits and isInside exactly; addBoundaryPotential (on the device) to 1e-14
    of the potential's scale."""
    import spacecharge
    b, ob = _both_boundaries(shape, npts, nm, xd, yd)
    bx, by, lims = ob.info()
    assert b.getNumberOfPoints() == len(bx) and b.getNumberOfModes() == nm
    assert b.getShapeName() == shape and b.isInitialized() == 1
    assert [b.getMinX(), b.getMaxX(), b.getMinY(), b.getMaxY()] == list(lims)
    assert all(b.getBoundaryPoint(i) == (bx[i], by[i]) for i in range(len(bx)))
    r = np.random.default_rng(3)
    for x, y in zip(r.uniform(-0.03, 0.03, 300), r.uniform(-0.03, 0.03, 300)):
        assert b.isInside(x, y) == ob.isInside(x, y)
    phi = r.normal(size=(40, 33))
    g = spacecharge.Grid2D(phi, *lims)
    rho = spacecharge.Grid2D(np.zeros_like(phi), *lims)
    b.addBoundaryPotential(rho, g)
    want = ob.addBoundaryPotential(phi, lims)
    assert np.abs(g.array() - want).max() <= 1e-14 * np.abs(want).max()


def test_boundary2d_point_set_and_errors():
    import spacecharge
    from _obnative import OrbitError
    with pytest.raises(OrbitError):
        spacecharge.Boundary2D(10, 5)  # 2*nModes+1 >= nPoints
    with pytest.raises(OrbitError):
        spacecharge.Boundary2D(64, 4, "Circle", 0.02, 0.03)
    with pytest.raises(OrbitError):
        spacecharge.Boundary2D(64, 4, "Hexagon", 0.02)
    shaped = spacecharge.Boundary2D(64, 4, "Circle", 0.02)
    with pytest.raises(OrbitError, match="shape has been defined"):
        shaped.setBoundaryPoint(0, 0.0, 0.0)
    b = spacecharge.Boundary2D(48, 5)
    ob = O.Boundary2D(48, 5)
    t = np.linspace(0, 2 * np.pi, 48, endpoint=False)
    for i in range(48):
        x, y = 0.012 * np.cos(t[i]) + 1e-3, 0.008 * np.sin(t[i]) * (1 + 0.1 * np.cos(t[i]))
        b.setBoundaryPoint(i, x, y)
        ob.setBoundaryPoint(i, x, y)
    assert b.isInitialized() == 0
    g = spacecharge.Grid2D(np.ones((16, 16)), -0.02, 0.02, -0.02, 0.02)
    with pytest.raises(OrbitError, match="not initialized"):
        b.addBoundaryPotential(g, g)
    b.initialize()
    ob.initialize()
    _, _, lims = ob.info()
    assert [b.getMinX(), b.getMaxX(), b.getMinY(), b.getMaxY()] == list(lims)
    phi = np.random.default_rng(4).normal(size=(24, 20))
    gg = spacecharge.Grid2D(phi, *lims)
    b.addBoundaryPotential(gg, gg)
    want = ob.addBoundaryPotential(phi, lims)
    assert np.abs(gg.array() - want).max() <= 1e-14 * np.abs(want).max()


@pytest.mark.parametrize("shape,npts,nm,xd,yd", BOUNDS, ids=[b[0] for b in BOUNDS])
def test_sc2p5d_with_boundary(shape, npts, nm, xd, yd):
    """trackBunch with a boundary: the grid is the boundary's box, phi carries the
    boundary's correction, particles outside the boundary keep their angles exactly."""
    import spacecharge
    c = _coords(30000, 4e-3, 3e-3, seed=40)
    c[0, :50] = 0.018  # some particles outside every boundary
    c[2, :50] = 0.013
    b, ob = _bunches(c)
    bnd, obnd = _both_boundaries(shape, npts, nm, xd, yd)
    calc = spacecharge.SpaceChargeCalc2p5D(64, 64, 16)
    ocalc = O.SC2p5D(64, 64, 16)
    for step in range(2):
        c0 = b.getCoordinates()
        calc.trackBunch(b, 0.5, bnd)
        ocalc.trackBunch(ob, 0.5, obnd)
        out = b.getCoordinates()
        rho, orho = calc.getRhoGrid(), ocalc.grid(0)
        assert [rho.getMinX(), rho.getMaxX(), rho.getMinY(), rho.getMaxY()] == list(orho[1])
        assert rel_l2(rho.array(), orho[0]) < 1e-12
        assert rel_l2(calc.getPhiGrid().array(), ocalc.grid(1)[0]) < 1e-12
        st, ost = calc.getState(), ocalc.state()
        assert st[0] == ost[0] and st[1] == ost[1] and st[3] == ost[3]
        for k in (1, 3):
            assert rel_l2(out[k] - c0[k], ob.coords[k] - c0[k]) < 1e-10, (step, k)
        outside = np.array([obnd.isInside(x, y) != 1 for x, y in zip(c0[0], c0[2])])
        assert outside.sum() >= 50
        for k in (1, 3):
            assert np.array_equal(out[k][outside], c0[k][outside])
    # the same calculator without the boundary afterwards: ratio rule from the boundary ratio
    c2 = _coords(20000, 3e-3, 3e-3, seed=41)
    b2, ob2 = _bunches(c2)
    calc.trackBunch(b2, 0.5)
    ocalc.trackBunch(ob2, 0.5)
    _compare(calc, ocalc, b2, ob2, c2)
