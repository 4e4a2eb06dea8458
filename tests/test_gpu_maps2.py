"""Bends, wedges, bend fringes, solenoid and ring RF on the device vs the oracle.

Maps made of +, -, *, / only are bit-exact (the host computes the same cos/sin/log
constants with the same C library as the oracle).  soln and RingRF call cos/sin per
particle on the device, whose last bit may differ from the host C library: they are held
to 1e-14 of each coordinate's scale, and the synchronous particle (computed on the host)
is exact."""
import numpy as np
import pytest

import oracle as O
from test_gpu_maps import coords, product_bunch, call_product

pytestmark = pytest.mark.gpu

EXACT = [
    ("bend1", 0.5, 0.1), ("bend1", 1.2, -0.3), ("bend2", 0.4), ("bend3", 0.05), ("bend4", -0.07),
    ("bendfringeIN", 4.0), ("bendfringeOUT", -6.0), ("wedgerotate", 0.13, 0),
    ("wedgerotate", -0.2, 1), ("wedgedrift", 0.1, 0), ("wedgedrift", 0.1, 1),
    ("wedgebend", 0.12, 0, 5.0, 4), ("wedgebend", 0.12, 1, 5.0, 7),
    ("wedgebendCF", 0.1, 0, 6.0, 2, [1, 2], [0.3, 1.5], [0, 1], 6),
    ("wedgebendCF", -0.1, 1, 6.0, 3, [0, 3, 5], [1e-3, -4.0, 20.0], [1, 0, 0], 3),
]
TRANS = [("soln", 0.5, 0.8), ("soln", 0.3, -2.0), ("soln", 0.7, 1e-101),
         ("RingRF", 248.0, 4, 2e-5, 0.0, 1), ("RingRF", 248.0, 1, -3e-5, 0.4, 1)]


def _oracle(c, ops):
    ob = O.Bunch(c, kin_energy=1.0)
    for op in ops:
        name, args = op[0], op[1:]
        if name == "soln":
            args = args[:2]
        ob.apply(name, *args)
    return ob


def _sync(b):
    sp = b.getSyncParticle()
    return sp.time(), sp.momentum(), sp.kinEnergy()


@pytest.mark.parametrize("op", EXACT, ids=[f"{o[0]}{i}" for i, o in enumerate(EXACT)])
def test_new_map_bit_exact(op):
    c = coords()
    ob = _oracle(c, [op])
    b = product_bunch(c)
    call_product(b, op)
    out = b.getCoordinates()
    assert np.array_equal(out, ob.coords), np.abs(out - ob.coords).max()
    assert _sync(b) == (ob.sync["time"], ob.sync["p"], ob.sync["T"])


@pytest.mark.parametrize("op", TRANS, ids=[f"{o[0]}{i}" for i, o in enumerate(TRANS)])
def test_transcendental_maps(op):
    c = coords()
    ob = _oracle(c, [op])
    b = product_bunch(c)
    call_product(b, op)
    out = b.getCoordinates()
    for k in range(6):
        scale = np.abs(ob.coords[k]).max()
        assert np.abs(out[k] - ob.coords[k]).max() <= 1e-14 * scale, k
    assert _sync(b) == (ob.sync["time"], ob.sync["p"], ob.sync["T"])


def _bend_element(e, rho, th, L, n):
    """The map list of a BendTEAPOT with wedge edges and a sextupole component
    (teapot.py:999-1150), as the lattice layer emits it for 3 parts."""
    ops = [("wedgedrift", e, 0), ("wedgerotate", e, 0), ("bendfringeIN", rho),
           ("multpfringeIN", 2, 0.5 / L, 0), ("wedgerotate", e, 1),
           ("wedgebendCF", e, 0, rho, 1, [2], [0.5], [0], n - 1), ("bend1", L / 4, th / 4)]
    ops += [("bend2", L / 4), ("bend3", th / 4), ("bend4", th / 4), ("multp", 2, 0.25, 0),
            ("bend4", th / 4), ("bend3", th / 4), ("bend2", L / 4), ("bend1", L / 2, th / 2)]
    ops += [("bend2", L / 4), ("bend3", th / 4), ("bend4", th / 4), ("multp", 2, 0.25, 0),
            ("bend4", th / 4), ("bend3", th / 4), ("bend2", L / 4), ("bend1", L / 2, th / 4)]
    ops += [("wedgebendCF", e, 1, rho, 1, [2], [0.5], [0], n - 1), ("wedgerotate", -e, 0),
            ("bendfringeOUT", rho), ("multpfringeOUT", 2, 0.5 / L, 0), ("wedgerotate", -e, 1),
            ("wedgedrift", e, 1)]
    return ops


def test_bend_element_sequence_bit_exact():
    """A whole bend element through ob_track_sequence (one fused launch) and through the
    single entry points (queued, fused at the read): the oracle's bits both ways."""
    import teapot_base as TPB
    ops = _bend_element(0.05, 10.0, 0.2, 2.0, 3) * 2
    c = coords(30000, 9)
    ob = _oracle(c, ops)
    b1 = product_bunch(c)
    TPB.trackSequence(b1, ops)
    b2 = product_bunch(c)
    for op in ops:
        call_product(b2, op)
    for b in (b1, b2):
        out = b.getCoordinates()
        assert np.array_equal(out, ob.coords), np.abs(out - ob.coords).max()
        assert _sync(b) == (ob.sync["time"], ob.sync["p"], ob.sync["T"])


def test_drifti_and_quad3():
    import teapot_base as TPB
    c = coords(100, 2)
    b = product_bunch(c)
    TPB.quad3(b, 0.5, 0.3)
    TPB.drift(b, 0.2)   # queued: drifti must see it applied
    TPB.drifti(b, 17, 0.9)
    ob = _oracle(c, [("drift", 0.2)])
    one = O.Bunch(ob.coords[:, 17:18].copy(), kin_energy=1.0)
    one.apply("drift", 0.9)
    ob.coords[:, 17] = one.coords[:, 0]
    assert np.array_equal(b.getCoordinates(), ob.coords)


def test_batch_ringrf_moves_every_bunch():
    """RingRF with a synchronous phase on a batch: the momentum change reaches every bunch
    and the maps after it use the new momentum."""
    from batch import BunchBatch
    import teapot_base as TPB
    cs = [coords(3000, 20), coords(5000, 21)]
    bunches = [product_bunch(c) for c in cs]
    bb = BunchBatch(bunches)
    ops = [("RingRF", 248.0, 2, 4e-5, 0.3, 1), ("drift", 0.7), ("bend1", 0.5, 0.1)]
    for op in ops:
        getattr(TPB, op[0])(bb.bunch(), *op[1:])
    for b, c in zip(bunches, cs):
        ob = _oracle(c, ops)
        out = b.getCoordinates()
        for k in range(6):
            assert np.abs(out[k] - ob.coords[k]).max() <= 1e-14 * np.abs(ob.coords[k]).max()
        assert _sync(b) == (ob.sync["time"], ob.sync["p"], ob.sync["T"])


def _ring_with_bends():
    from orbit.teapot import (TEAPOT_Ring, BendTEAPOT, QuadTEAPOT, DriftTEAPOT,
                              SolenoidTEAPOT, RingRFTEAPOT)
    lat = TEAPOT_Ring("ring")
    for k in range(4):
        q = QuadTEAPOT(f"Q{k}")
        q.setLength(0.5)
        q.setParam("kq", 0.6 if k % 2 == 0 else -0.6)
        b = BendTEAPOT(f"B{k}")
        b.setLength(2.0)
        b.setnParts(4)
        b.setParam("theta", 2 * 3.141592653589793 / 4)
        b.setParam("ea1", 0.1)
        b.setParam("ea2", 0.1)
        b.setParam("poles", [2])
        b.setParam("kls", [0.05])
        b.setParam("skews", [0])
        d1, d2 = DriftTEAPOT(f"D{k}"), DriftTEAPOT(f"E{k}")
        d1.setLength(1.0)
        d2.setLength(1.0)
        for nd in (q, d1, b, d2):
            lat.addNode(nd)
    s = SolenoidTEAPOT("S")
    s.setLength(0.3)
    s.setParam("B", 0.2)
    lat.addNode(s)
    rf = RingRFTEAPOT("RF")
    rf.setParam("ring_length", 20.3)
    rf.addRF(1, 3e-6, 0.0)
    lat.addNode(rf)
    lat.initialize()
    return lat


def test_ring_with_bends_solenoid_rf_vs_oracle(monkeypatch):
    """Two turns of a ring of quads, combined-function bends with edges, a solenoid and an
    RF cavity: the device against the oracle replaying the recorded map calls."""
    from helpers import TraceTPB, replay_oracle
    import orbit.teapot.teapot as tt
    tpb = TraceTPB()
    monkeypatch.setattr(tt, "TPB", tpb)

    class FakeBunch:
        def hasBunchAttrInt(self, name):
            return 0
    lat = _ring_with_bends()
    for _ in range(2):
        lat.trackBunch(FakeBunch())
    monkeypatch.undo()
    names = {c[0] for c in tpb.calls}
    assert {"wedgebendCF", "bend1", "bendfringeIN", "soln", "RingRF"} <= names
    c = coords(20000, 31)
    b = product_bunch(c)
    lat = _ring_with_bends()
    for _ in range(2):
        lat.trackBunch(b)
    out = b.getCoordinates()
    ob = O.Bunch(c, kin_energy=1.0)
    replay_oracle(tpb.calls, ob, None)
    for k in range(6):
        scale = np.abs(ob.coords[k]).max()
        assert np.abs(out[k] - ob.coords[k]).max() <= 1e-12 * scale, k
