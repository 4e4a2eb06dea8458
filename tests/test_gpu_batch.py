"""Batches of bunches (uneven sizes and grids) tracked together vs each bunch tracked on
its own, through the lattice layer; hipGraph capture and replay; tuning; errors."""
import numpy as np
import pytest

from helpers import rel_l2

pytestmark = pytest.mark.gpu

SIZES = [3000, 17000, 40000, 9000]
GRIDS = [(16, 16, 8), (32, 32, 16), (64, 64, 32), (24, 40, 9)]


def _coords(k, n):
    from workloads import gaussian_coords
    c = gaussian_coords(n, seed=100 + k)
    c[0] *= 1.0 + 0.3 * k  # different emittances
    return c


def _setup(sizes=SIZES, grids=GRIDS):
    import spacecharge
    from workloads import make_bunch
    bunches, calcs = [], []
    for k, (n, g) in enumerate(zip(sizes, grids)):
        bunches.append(make_bunch(_coords(k, n), intensity=1e12 * (1 + 10 * k)))
        calcs.append(spacecharge.SpaceChargeCalc2p5Drb(*g))
    return bunches, calcs


def _lattice(calc, cells):
    from orbit.space_charge.sc2p5d import setSC2p5DrbAccNodes
    from workloads import fodo_lattice, PIPE_RADIUS
    lat = fodo_lattice(cells)
    setSC2p5DrbAccNodes(lat, 0.49, calc, PIPE_RADIUS)
    return lat


def _single(cells, sizes=SIZES, grids=GRIDS):
    bunches, calcs = _setup(sizes, grids)
    for b, c in zip(bunches, calcs):
        _lattice(c, cells).trackBunch(b)
    return [b.getCoordinates() for b in bunches], bunches


def test_batch_matches_single_bunches():
    from batch import BunchBatch
    ref, refb = _single(2)
    bunches, calcs = _setup()
    bb = BunchBatch(bunches, calcs)
    _lattice(bb.calculator(), 2).trackBunch(bb.bunch())
    bb.synchronize()
    for k, b in enumerate(bunches):
        out = b.getCoordinates()
        for j in range(6):
            assert rel_l2(out[j], ref[k][j]) < 1e-9, (k, j)
        assert b.getSyncParticle().time() == refb[k].getSyncParticle().time()
    assert bb.tuning()["tuned"]


def test_batch_graph_replay():
    """One captured cell replayed twice equals three cells tracked directly."""
    from batch import BunchBatch
    ref, refb = _single(3)
    bunches, calcs = _setup()
    bb = BunchBatch(bunches, calcs)
    lat = _lattice(bb.calculator(), 1)
    lat.trackBunch(bb.bunch())            # first cell directly (tunes, builds records)
    bb.capture(lambda: lat.trackBunch(bb.bunch()))
    t_mid = bunches[0].getSyncParticle().time()
    bb.replay(2)
    bb.synchronize()
    for k, b in enumerate(bunches):
        out = b.getCoordinates()
        for j in range(6):
            assert rel_l2(out[j], ref[k][j]) < 1e-9, (k, j)
        assert abs(b.getSyncParticle().time() - refb[k].getSyncParticle().time()) <= \
            1e-12 * abs(refb[k].getSyncParticle().time())
    assert bunches[0].getSyncParticle().time() > t_mid


def test_batch_results_do_not_depend_on_tuning():
    from batch import BunchBatch
    outs = []
    for chunk, item in [(4096, 256), (65536, 8192)]:
        bunches, calcs = _setup()
        bb = BunchBatch(bunches, calcs)
        bb.setTuning(chunk, item)
        _lattice(bb.calculator(), 1).trackBunch(bb.bunch())
        assert bb.tuning() == {"chunk": chunk, "deposit_item": item, "tuned": True}
        outs.append([b.getCoordinates() for b in bunches])
    for k in range(len(SIZES)):
        for j in range(6):
            assert rel_l2(outs[0][k][j], outs[1][k][j]) < 1e-11


def test_batch_maps_are_bit_exact_and_reads_apply_pending_work():
    """Maps only (no SC): the batch gives the bits of per-bunch tracking; reading a bunch
    applies the batch's queued maps first."""
    from batch import BunchBatch
    import teapot_base as TPB
    ops = [("quadfringeIN", 0.5), ("quad1", 0.25, 0.5), ("quad2", 0.25), ("multp", 2, 0.4, 0),
           ("drift", 0.7)]
    single, _ = _setup()
    for b in single:
        TPB.trackSequence(b, ops)
    bunches, _ = _setup()
    bb = BunchBatch(bunches)
    bb.trackSequence(ops)
    for b, s in zip(bunches, single):
        assert np.array_equal(b.getCoordinates(), s.getCoordinates())


def test_batch_rejects_mixed_species_and_zero_extent():
    from batch import BunchBatch
    from _obnative import OrbitError
    import spacecharge
    bunches, calcs = _setup(SIZES[:2], GRIDS[:2])
    bunches[1].getSyncParticle().kinEnergy(2.0)
    with pytest.raises(OrbitError, match="synchronous momentum"):
        BunchBatch(bunches, calcs)
    bunches, calcs = _setup(SIZES[:2], GRIDS[:2])
    c = bunches[1].getCoordinates()
    c[2] = 1e-3  # y has zero extent
    bunches[1].setCoordinates(c)
    bb = BunchBatch(bunches, calcs)
    bb.calculator().trackBunch(bb.bunch(), 0.5, 0.1)
    with pytest.raises(OrbitError, match="zero extent"):
        bb.synchronize()


def _destroy_now(calc, fn):
    """Free a calculator's device state while work that reads it may still be queued."""
    import _obnative as N
    h, calc._h = calc._h, None
    N.call(fn, h)


def test_destroying_calculators_applies_pending_kicks():
    """The last SC kick of a lattice pass is deferred to the next push; freeing the
    calculator first must apply it (batch and single bunch), not drop it or read freed
    buffers."""
    from batch import BunchBatch
    ref, _ = _single(1)
    bunches, calcs = _setup()
    bb = BunchBatch(bunches, calcs)
    _lattice(bb.calculator(), 1).trackBunch(bb.bunch())
    for c in calcs:
        _destroy_now(c, "ob_sc2p5drb_destroy")
    for k, b in enumerate(bunches):
        out = b.getCoordinates()
        for j in range(6):
            assert rel_l2(out[j], ref[k][j]) < 1e-9, (k, j)
    with pytest.raises(Exception):
        bb.calculator().trackBunch(bb.bunch(), 0.1, 0.1)   # no calculators left
    del bb

    import spacecharge
    from workloads import make_bunch
    for k in range(2):
        b1 = make_bunch(_coords(k, 5000), intensity=1e12)
        b2 = make_bunch(_coords(k, 5000), intensity=1e12)
        mk = (lambda: spacecharge.SpaceChargeCalc2p5Drb(32, 32, 8)) if k == 0 else \
             (lambda: spacecharge.SpaceChargeCalc3D(16, 16, 16))
        args = (0.3, 0.05) if k == 0 else (0.3,)
        c1, c2 = mk(), mk()
        c1.trackBunch(b1, *args)
        ref1 = b1.getCoordinates()
        c2.trackBunch(b2, *args)
        _destroy_now(c2, "ob_sc2p5drb_destroy" if k == 0 else "ob_sc3d_destroy")
        out = b2.getCoordinates()
        for j in range(6):
            assert rel_l2(out[j], ref1[j]) < 1e-9, (k, j)
