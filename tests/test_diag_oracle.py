"""Oracle pins for the diagnostics: computeBunchMoments on particle sets whose moments are
known in closed form (including the reference's macrosize powers, doubled terms and the
stray [2][order] entry), the six extrema, Grid1D binning along any axis against the
reference's own Grid1D, the host side of spacecharge.Grid1D, and where
addTeapotDiagnosticsNode puts a node."""
import numpy as np
import pytest

import oracle as O


def _square(a, b, n_rep=1):
    """Particles at (+-a, +-b): <x^i y^j> = a^i b^j for even i, j, else 0."""
    x = np.tile([a, -a, a, -a], n_rep)
    y = np.tile([b, b, -b, -b], n_rep)
    c = np.zeros((6, x.size))
    c[0], c[2] = x, y
    c[5] = np.tile([1e-3, -1e-3, 2e-3, -2e-3], n_rep)
    return c


def _expected(order, a, b, xb, yb, m=None, npart=4):
    e = np.zeros((order + 1, order + 1))
    for j in range(order):
        for i in range(order + 1 - j):
            if i % 2 or j % 2:
                continue
            if m is None:
                e[i, j] = a ** i * b ** j / (xb ** i * yb ** j)
            else:  # every factor carries m, the term is added twice, the sum divided by sum m
                e[i, j] = m ** (i + j) * a ** i * b ** j * (1 / (xb ** i * yb ** j) + 1) / m
    if m is None and order >= 2 and order % 2 == 0:
        e[2, order] = npart * a * a * b ** order  # undivided, one term per particle
    e[0, 0], e[1, 0], e[0, 1] = 1.0, 0.0, 0.0
    return e


@pytest.mark.parametrize("order", [1, 2, 3, 4, 6])
@pytest.mark.parametrize("xb,yb", [(1.0, 1.0), (2.0, 0.5)])
def test_moments_closed_form(order, xb, yb):
    a, b = 3e-3, 2e-3
    ob = O.Bunch(_square(a, b), macrosize=1.0)
    got = ob.compute_moments(order, -1, 0.0, 0.0, 0.0, 1.0, 1.0, xb, yb)
    want = _expected(order, a, b, xb, yb)
    if order == 1:
        want[2:, :] = 0  # no stray entry: the reference writes outside its array there
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=0)


@pytest.mark.parametrize("order", [2, 4, 5])
def test_moments_with_macrosizes_closed_form(order):
    a, b, m = 3e-3, 2e-3, 2.5
    c = _square(a, b, n_rep=3)
    ob = O.Bunch(c, macrosize=1.0, msize=np.full(c.shape[1], m))
    got = ob.compute_moments(order, -1, 0.0, 0.0, 0.0, 1.0, 1.0, 1.5, 0.7)
    want = _expected(order, a, b, 1.5, 0.7, m=m)
    # odd entries cancel to the rounding of the doubled terms
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-14 * np.abs(want[1:, :]).max())


def test_moments_subtract_the_dispersive_part_and_the_averages():
    """x = k dE + pattern + x0 and y = pattern + y0: with dispersion flag 1 and
    D = k (E + M) beta^2 the moments are the pattern's; with flag -1 they are not."""
    a, b, k, x0, y0 = 3e-3, 2e-3, 1.7, 4e-4, -2e-4
    c = _square(a, b, n_rep=2)
    em, b2 = 1.938, 0.76
    c0 = c.copy()
    c[0] += k * c[5] + x0
    c[2] += y0
    D = k * em * b2
    got = O.Bunch(c, macrosize=1.0).compute_moments(4, 1, x0, y0, D, em, b2, 1.0, 1.0)
    want = O.Bunch(c0, macrosize=1.0).compute_moments(4, -1, 0.0, 0.0, 0.0, em, b2, 1.0, 1.0)
    want[1, 0], want[0, 1] = x0, y0
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-22)
    off = O.Bunch(c, macrosize=1.0).compute_moments(4, -1, x0, y0, D, em, b2, 1.0, 1.0)
    assert abs(off[2, 0] - want[2, 0]) > 0.1 * want[2, 0]


def test_extrema6_and_empty():
    r = np.random.default_rng(1)
    c = r.normal(size=(6, 1001))
    e = O.Bunch(c).extrema6()
    assert np.array_equal(e.reshape(6, 2), np.stack([c.min(axis=1), c.max(axis=1)], axis=1))
    e0 = O.Bunch(np.zeros((6, 0))).extrema6()
    assert np.all(e0[0::2] == np.finfo(float).max) and np.all(e0[1::2] == -np.finfo(float).max)


def test_extrema_r_closed_form():
    """Points on a circle of radius R around (a, b) plus points inside it: the largest
    distance from the mean is R (the mean is the centre for the symmetric set)."""
    R, a, b = 3e-3, 1e-3, -2e-3
    t = np.arange(16) * (2 * np.pi / 16)
    x = np.concatenate([a + R * np.cos(t), a + 0.5 * R * np.cos(t + 0.1)])
    y = np.concatenate([b + R * np.sin(t), b + 0.5 * R * np.sin(t + 0.1)])
    c = np.zeros((6, x.size))
    c[0], c[2] = x, y
    assert O.Bunch(c).extrema_r() == pytest.approx(R, rel=1e-13)
    assert O.Bunch(np.zeros((6, 0))).extrema_r() == 0.0


@pytest.mark.parametrize("axis", range(6))
def test_bin_axis_is_grid1d_bin_of_that_column(axis):
    r = np.random.default_rng(axis)
    c = r.normal(size=(6, 3000))
    m = r.uniform(0.5, 1.5, 3000)
    ob = O.Bunch(c, macrosize=2.0, msize=m)
    orc = O.orc_grid()
    lo, hi = -2.0, 2.5
    assert np.array_equal(ob.bin_axis(37, lo, hi, axis), orc.grid1d_bin(37, lo, hi, c[axis], m))
    assert np.array_equal(ob.bin_axis(37, lo, hi, axis, by_particle=True),
                          orc.grid1d_bin(37, lo, hi, c[axis], np.ones(3000)))
    ob2 = O.Bunch(c, macrosize=2.0)
    assert np.array_equal(ob2.bin_axis(5, lo, hi, axis),
                          orc.grid1d_bin(5, lo, hi, c[axis], np.full(3000, 2.0)))
    if O.ref_available():
        assert np.array_equal(ob.bin_axis(37, lo, hi, axis),
                              O.ref_grid().grid1d_bin(37, lo, hi, c[axis], m))


@pytest.mark.parametrize("nz", [1, 2, 7, 64])
def test_grid1d_host_methods_match_oracle(nz):
    import spacecharge
    zmin, zmax = -3.0, 5.0
    r = np.random.default_rng(nz)
    z = r.uniform(-3.5, 5.5, 500)
    z[:4] = [zmin, zmax, zmin + (zmax - zmin) / nz * 0.5, zmax - 1e-12]
    v = r.uniform(0.5, 1.5, 500)
    g = spacecharge.Grid1D(nz, zmin, zmax)
    for zz, vv in zip(z, v):
        g.binValue(vv, zz)
    orc = O.orc_grid()
    assert np.array_equal(g.array(), orc.grid1d_bin(nz, zmin, zmax, z, v))
    val, grad = orc.grid1d_value(nz, zmin, zmax, g.array(), z)
    assert np.array_equal(np.array([g.getValue(zz) for zz in z]), val)
    assert np.array_equal(np.array([g.calcGradient(zz) for zz in z]), grad)
    assert [g.isInside(zz) for zz in z] == [int(zmin <= zz <= zmax) for zz in z]
    assert g.getGridZ(0) == zmin + 0.5 * (zmax - zmin) / nz and g.getSizeZ() == nz


def test_grid1d_constructors():
    import spacecharge
    assert (spacecharge.Grid1D(10).getMinZ(), spacecharge.Grid1D(10).getMaxZ()) == (-1.0, 1.0)
    assert (spacecharge.Grid1D(10, 4.0).getMinZ(), spacecharge.Grid1D(10, 4.0).getMaxZ()) == \
        (0.0, 4.0)
    g = spacecharge.Grid1D(4, -2.0, 6.0)
    assert g.getStepZ() == 2.0
    g.setValue(3.0, 2)
    g.multiply(2.0)
    assert g.getValueOnGrid(2) == 6.0 and g.getSum() == 6.0
    g.setZero()
    assert g.getSum() == 0.0


def _drift_lattice(lengths):
    from orbit.teapot import TEAPOT_Lattice, DriftTEAPOT, QuadTEAPOT
    lat = TEAPOT_Lattice()
    for k, L in enumerate(lengths):
        if L is None:
            q = QuadTEAPOT("q")
            q.setLength(0.5)
            q.addParam("kq", 0.1)
            lat.addNode(q)
        else:
            d = DriftTEAPOT("d%d" % k)
            d.setLength(L)
            lat.addNode(d)
    lat.initialize()
    return lat


def test_add_diagnostics_node_splits_the_drift(tmp_path):
    from orbit.diagnostics import TeapotStatLatsNode, addTeapotDiagnosticsNode
    lat = _drift_lattice([2.0, None, 3.0])
    node = TeapotStatLatsNode(str(tmp_path / "s.dat"))
    addTeapotDiagnosticsNode(lat, 3.7, node)
    names = [(n.getName(), n.getLength()) for n in lat.getNodes()]
    assert [n for n, _ in names] == ["d0", "q", "d2", "statlats no name", "d2"]
    assert names[2][1] == pytest.approx(1.2) and names[4][1] == pytest.approx(1.8)
    assert node.position == 3.7 and node.lattlength == pytest.approx(5.5)
    # at the start of a drift: no leftover piece before the node
    lat2 = _drift_lattice([2.0, None, 3.0])
    addTeapotDiagnosticsNode(lat2, 2.5, TeapotStatLatsNode(str(tmp_path / "t.dat")))
    assert [n.getName() for n in lat2.getNodes()] == ["d0", "q", "statlats no name", "d2"]
    from orbit.utils import OrbitFinalizeError
    with pytest.raises(OrbitFinalizeError):
        addTeapotDiagnosticsNode(_drift_lattice([2.0, None, 3.0]), 2.2,
                                 TeapotStatLatsNode(str(tmp_path / "u.dat")))


def test_node_sets_attach_one_member_per_node(tmp_path):
    from orbit.diagnostics import addTeapotStatLatsNodeSet, addTeapotMomentsNodeSet
    lat = _drift_lattice([1.0, None, 2.0])
    ctl = addTeapotStatLatsNodeSet(lat, str(tmp_path / "set.dat"))
    assert len(ctl._nodelist) == 3
    assert [m.position for m in ctl._nodelist] == [0.0, 1.0, 1.5]
    ctl.deactivate()
    assert not any(m.active for m in ctl._nodelist)
    ctl.activate()
    ctl.resetFile(str(tmp_path / "set2.dat"))
    assert all(m.file is ctl._file for m in ctl._nodelist)
    mctl = addTeapotMomentsNodeSet(lat, str(tmp_path / "mom.dat"), 3)
    assert all(m.moments.dispterm == 1 and m.moments.emitnormterm == -1
               for m in mctl._nodelist)
