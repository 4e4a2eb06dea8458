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
