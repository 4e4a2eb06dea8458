"""SpaceChargeCalcSliceBySlice2D on the device vs the oracle: the 3D rho, the per-slice
potentials, the grid/ratio rule state, the transverse kicks, with and without a
Boundary2D, and lattice tracking.  Bars: rho within 1e-12 relative L2 (bin indices exact
on non-overlapping stencils), phi within 1e-12, transverse kicks within 1e-10, and x, y, z,
dE untouched bit for bit."""
import numpy as np
import pytest

import oracle as O
from helpers import rel_l2, trace_lattice, replay_oracle
from test_gpu_sc import _bunches

pytestmark = pytest.mark.gpu


def _coords(n, ax, ay, seed, zhalf=20.0):
    r = np.random.default_rng(seed)
    c = np.empty((6, n))
    c[0] = r.normal(2e-4, ax, n)
    c[1] = r.normal(0, 1e-3, n)
    c[2] = r.normal(-1e-4, ay, n)
    c[3] = r.normal(0, 1e-3, n)
    c[4] = r.normal(0, zhalf / 3, n)
    c[5] = r.normal(0, 1e-3, n)
    return c


def _compare(calc, ocalc, b, ob, c):
    out = b.getCoordinates()
    rho, orho = calc.getRhoGrid(), ocalc.grid(0)
    lims = [rho.getMinX(), rho.getMaxX(), rho.getMinY(), rho.getMaxY(), rho.getMinZ(),
            rho.getMaxZ()]
    assert lims == list(orho[1])
    assert rel_l2(rho.array(), orho[0]) < 1e-12
    assert rel_l2(calc.getPhiGrid().array(), ocalc.grid(1)[0]) < 1e-12
    st, ost = calc.getState(), ocalc.state()
    assert st == ost
    for k in (1, 3):
        assert rel_l2(out[k] - c[k], ob.coords[k] - c[k]) < 1e-10, k
    for k in (0, 2, 4, 5):
        assert np.array_equal(out[k], c[k]), k


@pytest.mark.parametrize("grid", [(16, 16, 8), (33, 20, 5), (64, 64, 32), (16, 16, 2),
                                  (128, 128, 64)])
def test_slice2d_step(grid):
    import spacecharge
    c = _coords(30000, 3e-3, 2.7e-3, seed=21)
    b, ob = _bunches(c)
    calc = spacecharge.SpaceChargeCalcSliceBySlice2D(*grid)
    ocalc = O.SliceBySlice2D(*grid)
    calc.trackBunch(b, 0.5)
    ocalc.trackBunch(ob, 0.5)
    _compare(calc, ocalc, b, ob, c)


def test_slice2d_ratio_rule_over_steps():
    import spacecharge
    calc = spacecharge.SpaceChargeCalcSliceBySlice2D(48, 40, 12, 1.2)
    ocalc = O.SliceBySlice2D(48, 40, 12, 1.2)
    for k, (ax, ay) in enumerate([(3e-3, 2.7e-3), (6e-3, 2e-3), (6e-3, 2e-3), (1e-3, 4e-3)]):
        c = _coords(20000, ax, ay, seed=200 + k)
        b, ob = _bunches(c)
        calc.trackBunch(b, 0.4)
        ocalc.trackBunch(ob, 0.4)
        _compare(calc, ocalc, b, ob, c)
    assert calc.getState()[2] == 2


def test_slice2d_with_boundary_and_macrosizes():
    import spacecharge
    c = _coords(20000, 3e-3, 2.5e-3, seed=33)
    c[0, :40] = 0.018
    c[2, :40] = 0.013
    m = np.random.default_rng(5).uniform(0.5e8, 1.5e8, c.shape[1])
    b, _ = _bunches(c)
    b.setMacroSizes(m)
    ob = O.Bunch(c, macrosize=1e8, kin_energy=1.0, msize=m)
    bnd = spacecharge.Boundary2D(128, 10, "Ellipse", 0.03, 0.02)
    obnd = O.Boundary2D(128, 10, "Ellipse", 0.03, 0.02)
    calc = spacecharge.SpaceChargeCalcSliceBySlice2D(40, 40, 16)
    ocalc = O.SliceBySlice2D(40, 40, 16)
    calc.trackBunch(b, 0.5, bnd)
    ocalc.trackBunch(ob, 0.5, obnd)
    _compare(calc, ocalc, b, ob, c)
    outside = np.array([obnd.isInside(x, y) != 1 for x, y in zip(c[0], c[2])])
    out = b.getCoordinates()
    assert outside.sum() >= 40
    assert np.array_equal(out[1][outside], c[1][outside])


def test_slice2d_zero_extent_raises():
    import spacecharge
    from _obnative import OrbitError
    c = np.zeros((6, 10))
    c[0] = np.linspace(-1e-3, 1e-3, 10)
    b, _ = _bunches(c)
    calc = spacecharge.SpaceChargeCalcSliceBySlice2D(16, 16, 8)
    with pytest.raises(OrbitError, match="zero extent"):
        calc.trackBunch(b, 0.5)


def _fodo_with_slices(calc, ncells):
    from orbit.space_charge.sc2dslicebyslice import setSC2DSliceBySliceAccNodes
    from workloads import fodo_lattice
    lat = fodo_lattice(ncells)
    nodes = setSC2DSliceBySliceAccNodes(lat, 0.49, calc)
    assert all(nd.getType() == "SC2DSliceBySlice" for nd in nodes)
    return lat


def test_fodo_slice2d_tracking(monkeypatch):
    import spacecharge
    from workloads import gaussian_coords, make_bunch
    n, grid, cells = 40000, (48, 48, 16), 2
    calls = trace_lattice(monkeypatch, lambda calc: _fodo_with_slices(calc, cells), "slice2d")
    monkeypatch.undo()
    assert sum(1 for c in calls if c[0] == "slice2d") == 12 * cells
    c = gaussian_coords(n, seed=9)
    b = make_bunch(c)
    lat = _fodo_with_slices(spacecharge.SpaceChargeCalcSliceBySlice2D(*grid), cells)
    lat.trackBunch(b)
    out = b.getCoordinates()
    ob = O.Bunch(c, macrosize=1e13 / n, kin_energy=1.0)
    replay_oracle(calls, ob, O.SliceBySlice2D(*grid))
    for k in range(6):
        assert rel_l2(out[k], ob.coords[k]) < 1e-9, (k, rel_l2(out[k], ob.coords[k]))
