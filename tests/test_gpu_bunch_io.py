"""dumpBunch / readBunch text files and the binary checkpoint, through the device bunch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bunch(n=257, seed=3, msize=True):
    from bunch import Bunch
    r = np.random.default_rng(seed)
    b = Bunch()
    b.macroSize(1.2345678e8)
    b.getSyncParticle().kinEnergy(1.0)
    b.getSyncParticle().time(1.25e-6)
    c = r.normal(size=(6, n)) * np.array([[3e-3], [5e-4], [3e-3], [5e-4], [0.3], [1e-3]])
    b.setCoordinates(c)
    if msize:
        b.setMacroSizes(r.uniform(0.5, 1.5, n) * 1e8)
    return b, c


def _g(v, p):
    return np.array([float(("%." + str(p) + "g") % x) for x in np.ravel(v)]).reshape(np.shape(v))


def test_dump_layout(tmp_path):
    b, c = _bunch(n=3)
    f = tmp_path / "b.dat"
    b.dumpBunch(str(f))
    lines = f.read_text().split("\n")
    assert lines[0] == "% PARTICLE_ATTRIBUTES_CONTROLLERS_NAMES macrosize "
    assert [ln.split()[2] for ln in lines[1:5]] == ["charge", "classical_radius", "macro_size",
                                                     "mass"]
    assert lines[4] == "% BUNCH_ATTRIBUTE_DOUBLE mass   0.938272 "
    assert lines[5].startswith("%  SYNC_PART_COORDS 0 0 0  x, y, z positions")
    assert lines[12].startswith("%  SYNC_PART_TIME 1.25e-06 ")
    assert lines[13] == "% x[m] px[rad] y[m] py[rad] z[m]  (pz or dE [GeV]) macro_size "
    row = [float(t) for t in lines[14].split()]
    assert np.array_equal(row[:6], _g(c[:, 0], 8)) and lines[14].endswith(" ")


def test_dump_read_round_trip(tmp_path):
    from bunch import Bunch
    b, c = _bunch()
    f = tmp_path / "b.dat"
    b.dumpBunch(str(f))
    b2 = Bunch()
    b2.addPartAttr("macrosize")
    assert b2.readBunch(str(f)) == c.shape[1]
    assert np.array_equal(b2.getCoordinates(), _g(c, 8))
    assert np.array_equal(b2._msize, _g(b._msize, 8))
    assert b2.mass() == 0.938272 and b2.macroSize() == float("%.6g" % b.macroSize())
    sp, sp2 = b.getSyncParticle(), b2.getSyncParticle()
    assert sp2.momentum() == float("%.14g" % sp.momentum())
    assert sp2.time() == 1.25e-6
    # appends, and honours nParts
    assert b2.readBunch(str(f), 10) == c.shape[1] + 10


def test_read_rejects_wrong_column_count(tmp_path):
    from bunch import Bunch
    from _obnative import OrbitError
    b, _ = _bunch(n=5)
    f = tmp_path / "b.dat"
    b.dumpBunch(str(f))
    with pytest.raises(OrbitError, match="data structure"):
        Bunch().readBunch(str(f))  # 7 columns, target bunch has no macrosize attribute


def test_binary_checkpoint_is_exact(tmp_path):
    from bunch import Bunch
    b, c = _bunch()
    b.bunchAttrDouble("ring_length", 248.0)
    f = tmp_path / "b.npz"
    b.saveBunchBinary(str(f))
    b2 = Bunch()
    assert b2.readBunchBinary(str(f)) == c.shape[1]
    assert np.array_equal(b2.getCoordinates(), c)
    assert np.array_equal(b2._msize, b._msize)
    for k in ("mass", "charge", "classicalRadius", "macroSize"):
        assert getattr(b2, k)() == getattr(b, k)()
    assert b2.getSyncParticle().momentum() == b.getSyncParticle().momentum()
    assert b2.getSyncParticle().kinEnergy() == b.getSyncParticle().kinEnergy()
    assert b2.bunchAttrDouble("ring_length") == 248.0


def test_add_particles_to_keeps_reference_attribute_semantics():
    """addParticlesTo (Bunch.cc:339-402): coordinates are appended; the appended particles'
    macrosizes start at 0 and the source's value i lands on the target's particle i; an
    empty target therefore gets a plain copy.  Different attribute sets are refused."""
    from bunch import Bunch
    from _obnative import OrbitError
    src, cs = _bunch(n=5, seed=1)
    dst, cd = _bunch(n=7, seed=2)
    ms, md = np.array(src.getMacroSizes()), np.array(dst.getMacroSizes())
    src.addParticlesTo(dst)
    assert np.array_equal(dst.getCoordinates(), np.concatenate([cd, cs], axis=1))
    want = np.concatenate([md, np.zeros(5)])
    want[:5] = ms
    assert np.array_equal(dst.getMacroSizes(), want)
    empty = Bunch()
    empty.addPartAttr("macrosize")
    src.addParticlesTo(empty)
    assert np.array_equal(empty.getCoordinates(), cs)
    assert np.array_equal(empty.getMacroSizes(), ms)
    plain, _ = _bunch(n=4, seed=3, msize=False)
    with pytest.raises(OrbitError, match="different structures"):
        src.addParticlesTo(plain)
    n0 = plain.getSize()
    plain2, c2 = _bunch(n=3, seed=4, msize=False)
    plain2.addParticlesTo(plain)
    assert plain.getSize() == n0 + 3 and not plain.hasPartAttr("macrosize")
