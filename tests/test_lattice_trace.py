"""Order of the teapot_base and space-charge calls the Python lattice issues, against
sequences derived by hand from the reference's traversal rules (AccNode.py:302-350,
teapot.py:548-565, 754-816, 843-970, 1322-1377, scLatticeModifications.py:10-61).
No GPU: the maps and the calculator are replaced by recorders."""
import pytest

from helpers import trace_lattice


def _fodo(calc, ncells=1):
    from orbit.space_charge.sc2p5d import setSC2p5DrbAccNodes
    from workloads import fodo_lattice
    lat = fodo_lattice(ncells)
    setSC2p5DrbAccNodes(lat, 0.49, calc, 0.1)
    return lat


def test_fodo_cell_call_sequence(monkeypatch):
    """QF (0.5 m, 2 parts): fringe in, [SC], quad1(L/2); part 1: quad2(L/2) x2 (no
    multipoles), quad1(L/2); fringe out.  Drifts of 5 parts of 0.5 m get an SC node before
    each part once 0.49 m of path has accumulated: the first SC node sits before QF part 0."""
    calls = trace_lattice(monkeypatch, _fodo, "sc2p5drb")
    SC = ("sc2p5drb", 0.5, 0.1)
    quad = lambda kq, sc_first: [("quadfringeIN", kq, 1)] + ([SC] if sc_first else []) + [
        ("quad1", 0.25, kq, 1), ("quad2", 0.25), ("quad2", 0.25), ("quad1", 0.25, kq, 1),
        ("quadfringeOUT", kq, 1)]
    drift5 = [SC, ("drift", 0.5)] * 5
    expected = quad(0.5, True) + drift5 + [("quadfringeIN", -0.5, 1), SC] + \
        quad(-0.5, False)[1:] + drift5
    assert calls == expected


def test_sc_nodes_cover_the_lattice(monkeypatch):
    """The SC lengths add up to the lattice length (the last node takes the rest)."""
    calls = trace_lattice(monkeypatch, lambda c: _fodo(c, 3), "sc2p5drb")
    sc = [c for c in calls if c[0] == "sc2p5drb"]
    assert len(sc) == 36
    assert sum(c[1] for c in sc) == pytest.approx(18.0, abs=1e-12)


def test_ring_wraps_and_multipole_parts(monkeypatch):
    from orbit.space_charge.sc3d import setSC3DAccNodes
    from workloads import ring_lattice

    def build(calc):
        lat = ring_lattice(1)
        setSC3DAccNodes(lat, 0.49, calc)
        return lat
    calls = trace_lattice(monkeypatch, build, "sc3d")
    names = [c[0] for c in calls]
    # every first-level node with length gets a wrap at its exit; 6 nodes per cell
    assert names.count("wrapbunch") == 6
    assert all(c[1] == pytest.approx(6.4) for c in calls if c[0] == "wrapbunch")
    # sextupole S1: fringe in with kl/L, drift L/2, multp kl/(nParts-1), drift L/2, fringe out
    i = names.index("multpfringeIN")
    seq = [c for c in calls[i:i + 6] if c[0] != "sc3d"]
    assert seq == [("multpfringeIN", 2, 2.0, 0, 1), ("drift", 0.1), ("multp", 2, 0.4, 0, 1),
                   ("drift", 0.1), ("multpfringeOUT", 2, 2.0, 0, 1)]


def test_kick_node_parts(monkeypatch):
    import orbit.teapot.teapot as tt
    from helpers import TraceTPB
    from orbit.teapot import KickTEAPOT, TEAPOT_Lattice
    tpb = TraceTPB()
    monkeypatch.setattr(tt, "TPB", tpb)
    k = KickTEAPOT("K")
    k.setLength(1.0)
    k.setnParts(3)
    k.setParam("kx", 2e-4)
    lat = TEAPOT_Lattice()
    lat.addNode(k)
    lat.initialize()

    class B:
        pass
    lat.trackBunch(B())
    assert tpb.calls == [("drift", 0.25), ("kick", 1e-4, 0.0, 0.0, 1), ("drift", 0.5),
                         ("kick", 1e-4, 0.0, 0.0, 1), ("drift", 0.25)]


def test_orbit_finalize_raises():
    from orbit.teapot import QuadTEAPOT
    from orbit.utils import OrbitFinalizeError
    q = QuadTEAPOT("Q")
    with pytest.raises(OrbitFinalizeError):
        q.setnParts(1)


def _trace_nodes(monkeypatch, *nodes):
    import orbit.teapot.teapot as tt
    from helpers import TraceTPB
    from orbit.teapot import TEAPOT_Lattice
    tpb = TraceTPB()
    monkeypatch.setattr(tt, "TPB", tpb)
    lat = TEAPOT_Lattice()
    for nd in nodes:
        lat.addNode(nd)
    lat.initialize()

    class B:
        pass
    lat.trackBunch(B())
    return tpb.calls


def test_bend_with_edges_and_sextupole(monkeypatch):
    """BendTEAPOT, 3 parts, wedge angles on both ends, one sextupole component
    (teapot.py:999-1150): wedge drift, rotate, bend fringe, multipole fringe with kl/L,
    rotate back, wedge bend with nParts-1 steps; parts 0 / 1 / 2 of the bend1..4 splitting
    with theta/(nParts-1) and kl/(nParts-1); the exit mirrored with -e."""
    from orbit.teapot import BendTEAPOT
    b = BendTEAPOT("B")
    b.setLength(2.0)
    b.setnParts(3)
    b.setParam("theta", 0.2)
    b.setParam("ea1", 0.05)
    b.setParam("ea2", 0.07)
    b.setParam("poles", [2])
    b.setParam("kls", [0.5])
    b.setParam("skews", [0])
    calls = _trace_nodes(monkeypatch, b)
    rho = 2.0 / 0.2
    th = 0.1  # theta / (nParts - 1)
    L0, L1 = 0.5, 1.0  # part lengths: half steps at the ends
    expected = [("wedgedrift", 0.05, 0), ("wedgerotate", 0.05, 0), ("bendfringeIN", rho),
                ("multpfringeIN", 2, 0.25, 0, 1), ("wedgerotate", 0.05, 1),
                ("wedgebendCF", 0.05, 0, rho, 1, [2], [0.5], [0], 2, 1),
                ("bend1", L0, th / 2),
                ("bend2", L1 / 2), ("bend3", th / 2), ("bend4", th / 2), ("multp", 2, 0.25, 0, 1),
                ("bend4", th / 2), ("bend3", th / 2), ("bend2", L1 / 2), ("bend1", L1, th),
                ("bend2", L0), ("bend3", th / 2), ("bend4", th / 2), ("multp", 2, 0.25, 0, 1),
                ("bend4", th / 2), ("bend3", th / 2), ("bend2", L0), ("bend1", L0, th / 2),
                ("wedgebendCF", 0.07, 1, rho, 1, [2], [0.5], [0], 2, 1),
                ("wedgerotate", -0.07, 0), ("bendfringeOUT", rho),
                ("multpfringeOUT", 2, 0.25, 0, 1), ("wedgerotate", -0.07, 1),
                ("wedgedrift", 0.07, 1)]
    assert calls == expected


def test_bend_without_edges_solenoid_rf_monitor(monkeypatch):
    from orbit.teapot import BendTEAPOT, SolenoidTEAPOT, RingRFTEAPOT
    b = BendTEAPOT("B")
    b.setLength(1.0)
    b.setParam("theta", 0.1)
    s = SolenoidTEAPOT("S")
    s.setLength(0.4)
    s.setParam("B", 0.8)
    rf = RingRFTEAPOT("RF")
    rf.setParam("ring_length", 248.0)
    rf.addRF(1, 1e-5, 0.0)
    rf.addRF(2, -5e-6, 0.1)
    b.setUsageFringeFieldOUT(False)
    calls = _trace_nodes(monkeypatch, b, s, rf)
    rho = 1.0 / 0.1
    assert calls == [("bendfringeIN", rho), ("bend1", 0.5, 0.05),
                     ("bend2", 0.5), ("bend3", 0.05), ("bend4", 0.05), ("bend4", 0.05),
                     ("bend3", 0.05), ("bend2", 0.5), ("bend1", 0.5, 0.05),
                     ("soln", 0.4, 0.8, 1),
                     ("RingRF", 248.0, 1, 1e-5, 0.0, 1), ("RingRF", 248.0, 2, -5e-6, 0.1, 1)]


def test_bend_and_rf_part_rules():
    from orbit.teapot import BendTEAPOT, RingRFTEAPOT
    from orbit.utils import OrbitFinalizeError
    b = BendTEAPOT("B")
    with pytest.raises(OrbitFinalizeError):
        b.setnParts(1)
    rf = RingRFTEAPOT("RF")
    with pytest.raises(OrbitFinalizeError):
        rf.setnParts(2)
