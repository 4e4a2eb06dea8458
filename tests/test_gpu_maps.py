"""TEAPOT maps on the device vs the oracle: bit-exact (integer-free FP64 arithmetic done in
the same order, no FMA contraction)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

OPS = [
    ("drift", 0.7),
    ("drift", -0.3),
    ("quad1", 0.25, 0.5),
    ("quad1", 0.25, -0.5),
    ("quad1", 0.3, 0.0),
    ("quad2", 0.25),
    ("quadfringeIN", 0.5),
    ("quadfringeOUT", -0.5),
    ("multp", 2, 0.4, 0),
    ("multp", 3, -1.2, 1),
    ("multp", 0, 1e-3, 0),
    ("multpfringeIN", 2, 2.0, 0),
    ("multpfringeOUT", 2, 2.0, 1),
    ("multpfringeIN", 1, -0.7, 1),
    ("wrapbunch", 50.0),
    ("rotatexy", 0.3),
    ("kick", 1e-4, -2e-4, 1e-5),
]


def coords(n=5000, seed=7):
    r = np.random.default_rng(seed)
    c = np.empty((6, n))
    c[0] = r.normal(0, 2e-2, n)
    c[1] = r.normal(0, 2e-3, n)
    c[2] = r.normal(0, 2e-2, n)
    c[3] = r.normal(0, 2e-3, n)
    c[4] = r.uniform(-100, 100, n)
    c[5] = r.normal(0, 3e-3, n)
    return c


def product_bunch(c):
    from workloads import make_bunch
    return make_bunch(c)


def call_product(b, op):
    import teapot_base as TPB
    name, args = op[0], op[1:]
    getattr(TPB, name)(b, *args)


@pytest.mark.parametrize("op", OPS, ids=[f"{o[0]}{i}" for i, o in enumerate(OPS)])
def test_map_bit_exact(op):
    c = coords()
    ob = O.Bunch(c, kin_energy=1.0)
    ob.apply(op[0], *op[1:])
    b = product_bunch(c)
    call_product(b, op)
    out = b.getCoordinates()
    assert np.array_equal(out, ob.coords), np.abs(out - ob.coords).max()
    assert b.getSyncParticle().time() == ob.sync["time"]


def test_fused_program_bit_exact():
    """All maps queued back to back run as one fused kernel; same bits as one by one."""
    c = coords(20000, 3)
    ob = O.Bunch(c, kin_energy=1.0)
    b = product_bunch(c)
    for op in OPS * 3:  # 51 ops: more than one launch's worth
        ob.apply(op[0], *op[1:])
        call_product(b, op)
    out = b.getCoordinates()
    assert np.array_equal(out, ob.coords), np.abs(out - ob.coords).max()


def test_track_sequence_entry_point_bit_exact():
    """ob_track_sequence (one call for a whole map list) gives the bits of the oracle applying
    the same maps one at a time."""
    import teapot_base as TPB
    c = coords(20000, 4)
    ob = O.Bunch(c, kin_energy=1.0)
    b = product_bunch(c)
    for op in OPS * 2:
        ob.apply(op[0], *op[1:])
    TPB.trackSequence(b, OPS * 2)
    out = b.getCoordinates()
    assert np.array_equal(out, ob.coords), np.abs(out - ob.coords).max()
    assert b.getSyncParticle().time() == ob.sync["time"]


def test_per_particle_accessors_and_empty_bunch():
    from bunch import Bunch
    import teapot_base as TPB
    b = Bunch()
    b.getSyncParticle().kinEnergy(1.0)
    TPB.drift(b, 1.0)  # empty bunch: no-op but advances the synchronous time
    assert b.getSize() == 0
    b.addParticle(1e-3, 1e-4, -2e-3, 0.0, 0.5, 1e-4)
    b.addParticle(0.0, 0.0, 0.0, 0.0, 0.0, 0.0)
    TPB.drift(b, 2.0)
    ob = O.Bunch(np.array([[1e-3, 0], [1e-4, 0], [-2e-3, 0], [0, 0], [0.5, 0], [1e-4, 0]]))
    ob.apply("drift", 2.0)
    assert b.x(0) == ob.coords[0, 0] and b.z(0) == ob.coords[4, 0]
    b.x(1, 5.0)
    TPB.drift(b, 1.0)
    assert b.x(1) == 5.0
    b.deleteParticle(0)
    assert b.getSize() == 1 and b.x(0) == 5.0
