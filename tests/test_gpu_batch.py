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


def test_batch_tuning_cached_by_shape():
    """The autotuner times its candidates once per batch shape (bunch sizes, grids) and
    process: a second batch of the same shape takes the first one's choice, and tracks
    like single bunches with it."""
    from batch import BunchBatch
    sizes, grids = [5000, 11000], [(16, 16, 8), (32, 32, 16)]
    ref, _ = _single(1, sizes, grids)
    tunings = []
    for _ in range(2):
        bunches, calcs = _setup(sizes, grids)
        bb = BunchBatch(bunches, calcs)
        _lattice(bb.calculator(), 1).trackBunch(bb.bunch())
        bb.synchronize()
        tunings.append(bb.tuning())
        for k, b in enumerate(bunches):
            out = b.getCoordinates()
            for j in range(6):
                assert rel_l2(out[j], ref[k][j]) < 1e-9, (k, j)
    assert tunings[0]["tuned"] and tunings[1] == tunings[0]


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


def _replay_each(bunch_coords, specs, cells, calls):
    """Track each bunch alone on the oracle through the recorded map/kick sequence."""
    import oracle as O
    from helpers import replay_oracle
    outs = []
    for c, sp in zip(bunch_coords, specs):
        ob = O.Bunch(c, macrosize=sp["intensity"] / c.shape[1], kin_energy=1.0)
        replay_oracle(calls, ob, O.SC2p5Drb(*sp["grid"]))
        outs.append(ob.coords)
    return outs


def _c3_specs():
    """Bunches of the 64-bunch workload's kinds: Gaussian and KV, 50k to 2M particles,
    grids 64^2, 128^2 and 256^2 with 64 z bins, intensities 1e12 to 1e14."""
    from workloads import batch_specs
    specs = batch_specs(64)
    pick = []
    for g in (64, 128, 256):
        for kind in ("gauss", "kv"):
            pick.append(next(s for s in specs if s["grid"][0] == g and s["kind"] == kind
                             and all(s is not p for p in pick)))
    sizes = [50_000, 2_000_000, 300_000, 120_000, 800_000, 1_500_000]
    for s, n in zip(pick, sizes):
        s["n"] = n
    return pick


def test_batch_c3_shapes_match_oracle(monkeypatch):
    """Six bunches spanning every grid size and the size range of the 64-bunch workload,
    batched, tracked two FODO cells (24 SC steps, hipGraph replay of the second) and
    compared with the oracle tracking each bunch alone: 1e-9 relative L2."""
    import oracle as O
    import spacecharge
    from batch import BunchBatch
    from helpers import trace_lattice
    from workloads import batch_coords, make_bunch
    O.set_threads(16)
    specs = _c3_specs()
    calls = trace_lattice(monkeypatch, lambda calc: _lattice(calc, 1), "sc2p5drb")
    monkeypatch.undo()
    coords = [batch_coords(s) for s in specs]
    bunches = [make_bunch(c, intensity=s["intensity"]) for c, s in zip(coords, specs)]
    calcs = [spacecharge.SpaceChargeCalc2p5Drb(*s["grid"]) for s in specs]
    bb = BunchBatch(bunches, calcs)
    lat = _lattice(bb.calculator(), 1)
    lat.trackBunch(bb.bunch())
    bb.capture(lambda: lat.trackBunch(bb.bunch()))
    bb.replay(1)
    bb.synchronize()
    ref = _replay_each(coords, specs, 2, calls + calls)
    for k, (b, s) in enumerate(zip(bunches, specs)):
        out = b.getCoordinates()
        errs = [rel_l2(out[j], ref[k][j]) for j in range(6)]
        print(f"bunch {k} {s['kind']} n={s['n']} grid={s['grid']}: "
              + " ".join(f"{e:.1e}" for e in errs), flush=True)
        for j in range(6):
            assert errs[j] < 1e-9, (k, j, errs[j])


def _dn_params(seed):
    """Parameters with non-zero biases and head: the net changes rho visibly."""
    import denoise_oracle as D
    r = np.random.default_rng(seed)
    p = D.he_init(seed)
    off = 0
    for ci, co, _ in D.LAYERS:
        off += 9 * ci * co
        p[off:off + co] = r.normal(0, 0.05, co)
        off += co
    p[-17:-1] = r.normal(0, 0.3, 16)
    p[-1] = 0.1
    return p.astype(np.float32)


def test_batch_with_denoiser_equals_bunches_alone():
    """Batched + denoised tracking (one denoiser in every SC node of every bunch, 256 x 256
    x 64 grids) equals each bunch tracked alone with the same denoiser; the denoiser does
    change the result (it differs from plain tracking)."""
    import spacecharge
    from batch import BunchBatch
    from denoiser import Denoiser
    from workloads import gaussian_coords, kv_coords, make_bunch
    sizes = [200_000, 500_000, 120_000]
    mk = [lambda n, s: gaussian_coords(n, seed=s), lambda n, s: kv_coords(n, seed=s),
          lambda n, s: gaussian_coords(n, seed=s, sx=2e-3)]
    coords = [f(n, 300 + k) for k, (f, n) in enumerate(zip(mk, sizes))]
    dn = Denoiser(256, 256, max_batch=4)
    dn.setParams(_dn_params(9))

    def run(batched, denoise):
        bunches = [make_bunch(c, intensity=1e13) for c in coords]
        calcs = [spacecharge.SpaceChargeCalc2p5Drb(256, 256, 64) for _ in coords]
        if denoise:
            for c in calcs:
                c.setDenoiser(dn)
        if batched:
            bb = BunchBatch(bunches, calcs)
            _lattice(bb.calculator(), 1).trackBunch(bb.bunch())
            bb.synchronize()
        else:
            for b, c in zip(bunches, calcs):
                _lattice(c, 1).trackBunch(b)
        return [b.getCoordinates() for b in bunches]

    alone, batched, plain = run(False, True), run(True, True), run(True, False)
    for k in range(len(coords)):
        for j in range(6):
            assert rel_l2(batched[k][j], alone[k][j]) < 1e-9, (k, j)
        # the net's correction is visible in the tracked coordinates
        dk = rel_l2(batched[k][1], plain[k][1])
        assert dk > 1e-7, (k, dk)
