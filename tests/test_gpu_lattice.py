"""Lattice tracking with space charge on the device vs the oracle replaying the same
map/kick sequence.  Bar: coordinates within 1e-9 relative L2 after the tracked cells."""
import numpy as np
import pytest

import oracle as O
from helpers import rel_l2, trace_lattice, replay_oracle

pytestmark = pytest.mark.gpu


def _fodo_with_sc(calc, ncells):
    from orbit.space_charge.sc2p5d import setSC2p5DrbAccNodes
    from workloads import fodo_lattice, PIPE_RADIUS
    lat = fodo_lattice(ncells)
    setSC2p5DrbAccNodes(lat, 0.49, calc, PIPE_RADIUS)
    return lat


def _ring_with_sc(calc, ncells):
    from orbit.space_charge.sc3d import setSC3DAccNodes
    from workloads import ring_lattice
    lat = ring_lattice(ncells)
    setSC3DAccNodes(lat, 0.49, calc)
    return lat


@pytest.mark.parametrize("n,grid,cells", [(20000, (32, 32, 16), 10), (100000, (128, 128, 64), 100)])
def test_fodo_2p5drb_tracking(monkeypatch, n, grid, cells):
    """(100000, 128 x 128 x 64, 100 cells) is the reference's CPU workload in full: 1,200
    SC steps, coordinates within 1e-9 of the oracle at the end."""
    import spacecharge
    from workloads import gaussian_coords, make_bunch
    O.set_threads(16)
    calls = trace_lattice(monkeypatch, lambda calc: _fodo_with_sc(calc, cells), "sc2p5drb")
    monkeypatch.undo()
    assert sum(1 for c in calls if c[0] == "sc2p5drb") == 12 * cells
    c = gaussian_coords(n, seed=3)
    b = make_bunch(c)
    lat = _fodo_with_sc(spacecharge.SpaceChargeCalc2p5Drb(*grid), cells)
    lat.trackBunch(b)
    out = b.getCoordinates()
    ob = O.Bunch(c, macrosize=1e13 / n, kin_energy=1.0)
    replay_oracle(calls, ob, O.SC2p5Drb(*grid))
    errs = [rel_l2(out[k], ob.coords[k]) for k in range(6)]
    print(f"FODO {n} particles {grid} {cells} cells: rel L2 per coordinate "
          + " ".join(f"{e:.2e}" for e in errs), flush=True)
    for k in range(6):
        assert errs[k] < 1e-9, (k, errs[k])


def test_ring_3d_tracking_stepwise(monkeypatch):
    """3D ring, two cells, 25 SC steps, each checked on identical inputs.

    The reference's 3D grid limits are center +- width (SpaceChargeCalc3D.cc:207-235); that
    rounding can fall just inside the extreme particle, which is then left out of rho.  A
    last-bit difference in the previous kick flips that decision, so two FP64
    implementations of the reference drift apart over many steps by whole-particle rho
    changes.  The bar is therefore per step: before every SC call the oracle bunch is
    reset to the device bunch, and rho/phi (1e-12) and the kick (1e-10) are compared."""
    import spacecharge
    import teapot_base as TPB
    from workloads import waterbag_coords, make_bunch
    calls = trace_lattice(monkeypatch, lambda calc: _ring_with_sc(calc, 2), "sc3d")
    monkeypatch.undo()
    n, grid = 20000, (16, 16, 16)
    c = waterbag_coords(n, seed=5)
    b = make_bunch(c)
    ob = O.Bunch(c, macrosize=1e13 / n, kin_energy=1.0)
    calc, ocalc = spacecharge.SpaceChargeCalc3D(*grid), O.SC3D(*grid)
    nsc = 0
    for cl in calls:
        if cl[0] != "sc3d":
            getattr(TPB, cl[0])(b, *cl[1:])
            replay_oracle([cl], ob, None)
            continue
        before = b.getCoordinates()
        assert np.array_equal(before, ob.coords)  # maps are bit-exact
        calc.trackBunch(b, *cl[1:])
        ocalc.trackBunch(ob, *cl[1:])
        after = b.getCoordinates()
        assert rel_l2(calc.getRhoGrid().array(), ocalc.grid(0)[0]) < 1e-12
        assert rel_l2(calc.getPhiGrid().array(), ocalc.grid(1)[0]) < 1e-12
        for k in (1, 3, 5):
            assert rel_l2(after[k] - before[k], ob.coords[k] - before[k]) < 1e-10, (nsc, k)
        ob.coords[:] = after  # identical input for the next step
        nsc += 1
    assert nsc == 25 and calc.getGreenRebuilds() == ocalc.green_rebuilds()
