"""Space-charge calculators on the device vs the oracle.

Bars: bin indices and TSC weights bit-exact (sparse bunches whose stencils do not overlap,
so the FP64 atomic order cannot matter); rho, phi, line density and its derivative within
1e-12 relative L2; kicks within 1e-10 relative L2 (the kick multiplies the field by
O(1e-8) factors; the bar is on the kick itself)."""
import numpy as np
import pytest

import oracle as O
from helpers import rel_l2

pytestmark = pytest.mark.gpu


def _bunches(c, macrosize=1e8):
    from workloads import make_bunch
    b = make_bunch(c, intensity=macrosize * c.shape[1])
    ob = O.Bunch(c, macrosize=macrosize, kin_energy=1.0)
    return b, ob


def sparse_coords(n_side=6, spacing_cells=5, nx=64, seed=1):
    """Particles on a jittered lattice: 3x3(x3) stencils never overlap."""
    r = np.random.default_rng(seed)
    pts = []
    step = spacing_cells / (nx - 1)
    for i in range(n_side):
        for j in range(n_side):
            pts.append((i * step + r.uniform(-0.2, 0.2) * step, j * step + r.uniform(-0.2, 0.2) * step))
    p = np.array(pts).T
    n = p.shape[1]
    c = np.zeros((6, n))
    c[0] = p[0] * 1e-2
    c[2] = p[1] * 1e-2
    c[1] = r.normal(0, 1e-3, n)
    c[3] = r.normal(0, 1e-3, n)
    c[4] = np.linspace(-1.0, 1.0, n) + r.uniform(-1e-3, 1e-3, n)
    c[5] = r.normal(0, 1e-3, n)
    return c


@pytest.mark.parametrize("grid,n", [((16, 16, 8), 20000), ((32, 32, 16), 20000),
                                    ((33, 20, 5), 20000), ((128, 128, 64), 20000),
                                    ((256, 256, 64), 1_000_000), ((256, 128, 64), 500_000)])
def test_sc2p5drb_step(grid, n):
    """One SpaceChargeCalc2p5Drb step; (256, 256, 64) pads the solve to 512 x 512, the
    largest grid of the batched and denoised workloads."""
    import spacecharge
    from workloads import gaussian_coords
    nx, ny, nz = grid
    c = gaussian_coords(n, seed=11)
    b, ob = _bunches(c)
    calc = spacecharge.SpaceChargeCalc2p5Drb(nx, ny, nz)
    ocalc = O.SC2p5Drb(nx, ny, nz)
    calc.trackBunch(b, 0.5, 0.1)
    ocalc.trackBunch(ob, 0.5, 0.1)
    out = b.getCoordinates()
    rho, orho = calc.getRhoGrid(), ocalc.grid(0)
    assert np.allclose([rho.getMinX(), rho.getMaxX(), rho.getMinY(), rho.getMaxY()], orho[1], rtol=0, atol=0)
    assert rel_l2(rho.array(), orho[0]) < 1e-12
    assert rel_l2(calc.getPhiGrid().array(), ocalc.grid(1)[0]) < 1e-12
    zg = calc.getLongGrid()
    assert rel_l2(zg.array(), ocalc.grid(2)[0]) < 1e-12
    # dlambda/dz: the reference's 3-point least-squares formula cancels terms ~1e4 times
    # larger than its result, so the 1e-16 difference between an atomic and a sequential
    # sum of the line density grows to ~1e-11.  Given the same line density the device
    # derivative is bit-identical to the oracle's.
    zd = calc.getLongDerivativeGrid().array()
    assert np.array_equal(zd, O.long_derivative(zg.array(), zg.getMinZ(), zg.getMaxZ()))
    assert rel_l2(zd, ocalc.grid(3)[0]) < 1e-9
    m, om = calc.getMoments(), ocalc.moments()
    # x_c, y_c are sums of terms of the grid's size: the rounding scales with the span,
    # not with the (near-zero) centre of a centred beam
    span = orho[1][1] - orho[1][0]
    assert np.allclose(m[:2], om[:2], rtol=0, atol=1e-13 * span)
    assert np.allclose(m[2:], om[2:], rtol=1e-12, atol=0)
    for k in (1, 3):  # transverse kicks: from phi, bar 1e-10
        assert rel_l2(out[k] - c[k], ob.coords[k] - c[k]) < 1e-10, k
    # longitudinal kick is proportional to dlambda/dz (conditioning above)
    assert rel_l2(out[5] - c[5], ob.coords[5] - c[5]) < 1e-8
    for k in (0, 2, 4):
        assert np.array_equal(out[k], ob.coords[k])


def test_sc2p5drb_bin_indices_exact():
    """Non-overlapping stencils: every rho cell gets at most one contribution, so rho and
    the line density must equal the oracle bit for bit (indices and weights exact)."""
    import spacecharge
    c = sparse_coords()
    b, ob = _bunches(c)
    calc = spacecharge.SpaceChargeCalc2p5Drb(64, 64, 200)
    ocalc = O.SC2p5Drb(64, 64, 200)
    calc.trackBunch(b, 0.5, 0.1)
    ocalc.trackBunch(ob, 0.5, 0.1)
    assert np.array_equal(calc.getRhoGrid().array(), ocalc.grid(0)[0])
    assert np.array_equal(calc.getLongGrid().array(), ocalc.grid(2)[0])


def test_sc2p5drb_macrosize_attribute_and_long_avg():
    import spacecharge
    from workloads import gaussian_coords
    c = gaussian_coords(5000, seed=5)
    m = np.random.default_rng(2).uniform(0.5e8, 1.5e8, c.shape[1])
    b, _ = _bunches(c)
    b.setMacroSizes(m)
    ob = O.Bunch(c, macrosize=1e8, kin_energy=1.0, msize=m)
    calc = spacecharge.SpaceChargeCalc2p5Drb(32, 32, 16, 1.5)
    calc.setLongAveragingPointsN(5)
    ocalc = O.SC2p5Drb(32, 32, 16, 1.5)
    ocalc.setLongAveragingPointsN(5)
    calc.trackBunch(b, 0.3, 0.05)
    ocalc.trackBunch(ob, 0.3, 0.05)
    out = b.getCoordinates()
    assert rel_l2(calc.getRhoGrid().array(), ocalc.grid(0)[0]) < 1e-12
    zg = calc.getLongGrid()
    assert np.array_equal(calc.getLongDerivativeGrid().array(),
                          O.long_derivative(zg.array(), zg.getMinZ(), zg.getMaxZ(), 5))
    for k in (1, 3):
        assert rel_l2(out[k] - c[k], ob.coords[k] - c[k]) < 1e-10
    assert rel_l2(out[5] - c[5], ob.coords[5] - c[5]) < 1e-8


def test_sc2p5drb_zero_extent_raises():
    import spacecharge
    from _obnative import OrbitError
    c = np.zeros((6, 10))
    c[0] = np.linspace(-1e-3, 1e-3, 10)  # y and z all equal: zero extent
    b, _ = _bunches(c)
    calc = spacecharge.SpaceChargeCalc2p5Drb(16, 16, 8)
    calc.trackBunch(b, 0.5, 0.1)
    with pytest.raises(OrbitError, match="zero extent"):
        b.getCoordinates()


# (24, 20, 64) and (16, 16, 256): z transforms of 128 and 512 points, whose LDS exchanges
# are padded against bank conflicts
@pytest.mark.parametrize("grid", [(16, 16, 16), (32, 32, 16), (17, 24, 9), (24, 20, 64),
                                  (16, 16, 256)])
def test_sc3d_step(grid):
    import spacecharge
    from workloads import waterbag_coords
    nx, ny, nz = grid
    c = waterbag_coords(20000, seed=13)
    b, ob = _bunches(c)
    calc = spacecharge.SpaceChargeCalc3D(nx, ny, nz)
    ocalc = O.SC3D(nx, ny, nz)
    for step in range(2):
        calc.trackBunch(b, 0.5)
        ocalc.trackBunch(ob, 0.5)
        out = b.getCoordinates()
        rho, orho = calc.getRhoGrid(), ocalc.grid(0)
        assert rel_l2(rho.array(), orho[0]) < 1e-12
        assert rel_l2(calc.getPhiGrid().array(), ocalc.grid(1)[0]) < 1e-12
        assert calc.getGreenRebuilds() == ocalc.green_rebuilds()
        for k in (1, 3, 5):
            assert rel_l2(out[k] - c[k], ob.coords[k] - c[k]) < 1e-10
        c = out.copy()
        ob.coords[:] = out


def test_sc3d_full_size_256x256x128():
    """SpaceChargeCalc3D at its largest workload: 4M waterbag particles on 256 x 256 x 128
    (the solve pads to 512 x 512 x 256, the Green function's real quadrant is
    257 x 257 x 129).  Three steps, each from identical inputs on both sides: the first
    builds the Green function, the second (after a short drift) keeps it, the third (x widened
    1.5 times: the xy ratio leaves the 1.2 limit) rebuilds it."""
    import spacecharge
    import teapot_base as TPB
    from workloads import waterbag_coords
    O.set_threads(16)
    n = 4_000_000
    c = waterbag_coords(n, seed=23)
    b, ob = _bunches(c, macrosize=1e13 / n)
    calc = spacecharge.SpaceChargeCalc3D(256, 256, 128)
    ocalc = O.SC3D(256, 256, 128)
    for step in range(3):
        if step == 1:
            TPB.drift(b, 0.05)  # x and y grow ~3%: both ratios stay within 1.2
        if step == 2:
            cc = b.getCoordinates()
            cc[0] *= 1.5
            b.setCoordinates(cc)
        before = b.getCoordinates()
        ob.coords[:] = before
        calc.trackBunch(b, 0.5)
        ocalc.trackBunch(ob, 0.5)
        after = b.getCoordinates()
        rho, orho = calc.getRhoGrid(), ocalc.grid(0)
        lims = [rho.getMinX(), rho.getMaxX(), rho.getMinY(), rho.getMaxY(), rho.getMinZ(),
                rho.getMaxZ()]
        assert np.array_equal(lims, orho[1]), step
        assert rel_l2(rho.array(), orho[0]) < 1e-12, step
        assert rel_l2(calc.getPhiGrid().array(), ocalc.grid(1)[0]) < 1e-12, step
        assert calc.getGreenRebuilds() == ocalc.green_rebuilds() == [1, 1, 2][step]
        for k in (1, 3, 5):
            assert rel_l2(after[k] - before[k], ob.coords[k] - before[k]) < 1e-10, (step, k)
        for k in (0, 2, 4):
            assert np.array_equal(after[k], ob.coords[k]), (step, k)
        print(f"3D 256x256x128 step {step}: rebuilds {calc.getGreenRebuilds()}", flush=True)


def test_sc3d_many_green_rebuilds():
    """Fourteen Green-function rebuilds in a row (the beam's x size alternates by 1.6), each
    step against the oracle from identical inputs.  H2 = 65 kz lines is odd: the last kz
    pair of the axis transforms has a padding partner, which must stay zero across
    rebuilds (rounding noise kept there grows by ~(2n)^2 per rebuild and reaches the
    potential within a few)."""
    import spacecharge
    from workloads import waterbag_coords
    c = waterbag_coords(20000, seed=29)
    b, ob = _bunches(c)
    calc = spacecharge.SpaceChargeCalc3D(64, 64, 64)
    ocalc = O.SC3D(64, 64, 64)
    for step in range(14):
        cc = b.getCoordinates()
        cc[0] *= 1.6 if step % 2 == 0 else 1 / 1.6
        b.setCoordinates(cc)
        ob.coords[:] = cc
        calc.trackBunch(b, 0.5)
        ocalc.trackBunch(ob, 0.5)
        after = b.getCoordinates()
        assert rel_l2(calc.getRhoGrid().array(), ocalc.grid(0)[0]) < 1e-12, step
        assert rel_l2(calc.getPhiGrid().array(), ocalc.grid(1)[0]) < 1e-12, step
        for k in (1, 3, 5):
            assert rel_l2(after[k] - cc[k], ob.coords[k] - cc[k]) < 1e-10, (step, k)
    assert calc.getGreenRebuilds() == ocalc.green_rebuilds() == 14


@pytest.mark.parametrize("zscale,nb", [(0.05, 4), (1.0, 3)])
def test_sc3d_external_bunches(zscale, nb):
    """numExtBunches > 0 (wrappedBunchAnalysis): z wrapped to one bunch spacing, image bunches
    in the Green function (rebuilt every call), periodic kick; then back to 0 external
    bunches, where the potential grid keeps its periodic z stencil as in the reference."""
    import spacecharge
    from workloads import waterbag_coords
    c = waterbag_coords(20000, seed=17)
    c[4] *= zscale  # 0.05: inside half a spacing (0.33 m); 1.0: the extension branch
    b, ob = _bunches(c)
    calc = spacecharge.SpaceChargeCalc3D(16, 16, 32)
    ocalc = O.SC3D(16, 16, 32)
    assert calc.numExtBunches(nb) == ocalc.numExtBunches(nb) == nb + nb % 2
    assert calc.freqOfBunches() == ocalc.freqOfBunches() == 402.5e6
    for step in range(3):
        if step == 2:
            calc.numExtBunches(0)
            ocalc.numExtBunches(0)
        calc.trackBunch(b, 0.5)
        ocalc.trackBunch(ob, 0.5)
        out = b.getCoordinates()
        assert rel_l2(calc.getRhoGrid().array(), ocalc.grid(0)[0]) < 1e-12
        assert rel_l2(calc.getPhiGrid().array(), ocalc.grid(1)[0]) < 1e-12
        assert calc.getGreenRebuilds() == ocalc.green_rebuilds()
        for k in (1, 3, 5):
            assert rel_l2(out[k] - c[k], ob.coords[k] - c[k]) < 1e-10
        c = out.copy()
        ob.coords[:] = out


def test_sc3d_bin_indices_exact():
    import spacecharge
    c = sparse_coords(nx=32)
    b, ob = _bunches(c)
    calc = spacecharge.SpaceChargeCalc3D(32, 32, 64)
    ocalc = O.SC3D(32, 32, 64)
    calc.trackBunch(b, 0.5)
    ocalc.trackBunch(ob, 0.5)
    assert np.array_equal(calc.getRhoGrid().array(), ocalc.grid(0)[0])


@pytest.mark.parametrize("budget", [110 * 1024, 20 * 1024])
def test_sc3d_tiled_matches_particle_order_deposit(budget):
    """Tile-sorted LDS deposit vs particle-order atomics: same rho to summation order, and
    the user-visible particle order is unchanged by the device sort."""
    import spacecharge
    from workloads import waterbag_coords
    c = waterbag_coords(50000, seed=21)
    b1, ob = _bunches(c)
    b2, _ = _bunches(c)
    c1 = spacecharge.SpaceChargeCalc3D(32, 32, 32)
    c2 = spacecharge.SpaceChargeCalc3D(32, 32, 32)
    c1.setDepositMode(True, budget)
    c2.setDepositMode(False)
    ocalc = O.SC3D(32, 32, 32)
    for _ in range(3):
        before = b1.getCoordinates()
        ob.coords[:] = before  # the oracle steps from the tiled path's input every time
        c1.trackBunch(b1, 0.5)
        c2.trackBunch(b2, 0.5)
        ocalc.trackBunch(ob, 0.5)
        assert rel_l2(c1.getRhoGrid().array(), c2.getRhoGrid().array()) < 1e-13
        o1, o2 = b1.getCoordinates(), b2.getCoordinates()
        for k in range(6):
            assert rel_l2(o1[k], o2[k]) < 1e-12
        assert rel_l2(c1.getRhoGrid().array(), ocalc.grid(0)[0]) < 1e-12
        assert rel_l2(c1.getPhiGrid().array(), ocalc.grid(1)[0]) < 1e-12
        for k in (1, 3, 5):
            assert rel_l2(o1[k] - before[k], ob.coords[k] - before[k]) < 1e-10, k
        for k in (0, 2, 4):
            assert np.array_equal(o1[k], ob.coords[k])


def test_sc3d_deposit_budget_range():
    """The tile deposit stages half a work item beside its LDS block: budgets above 116 KiB
    are refused, 116 KiB is accepted and deposits like the default."""
    import spacecharge
    from _obnative import OrbitError
    from workloads import waterbag_coords
    c = spacecharge.SpaceChargeCalc3D(32, 32, 32)
    with pytest.raises(OrbitError, match="LDS budget"):
        c.setDepositMode(True, 117 * 1024)
    c.setDepositMode(True, 116 * 1024)
    coords = waterbag_coords(20000, seed=23)
    b1, _ = _bunches(coords)
    b2, _ = _bunches(coords)
    c.trackBunch(b1, 0.5)
    c2 = spacecharge.SpaceChargeCalc3D(32, 32, 32)
    c2.trackBunch(b2, 0.5)
    assert rel_l2(c.getRhoGrid().array(), c2.getRhoGrid().array()) < 1e-13


# padded transform lengths (rows x columns): 128 x 64, 1024 x 512, 2048 x 2048, 4096 x 128
@pytest.mark.parametrize("nx,ny", [(40, 24), (300, 200), (600, 520), (1100, 40)])
def test_poisson_solver_fft2d_find_potential(nx, ny):
    """PoissonSolverFFT2D.findPotential on a user grid (solver on its own box, rho on a
    box of the same step ratio) against the oracle's PoissonSolverFFT2D."""
    import spacecharge
    r = np.random.default_rng(8)
    solver_lims = (-2.0, 2.0, -1.0, 1.0)
    # the rho grid's steps keep dx/dy of the solver:
    # (xMax-xMin)/(yMax-yMin) = 4 (ny-1) / (2 (nx-1))
    hx = 0.05
    hy = hx * (2.0 / (ny - 1)) / (4.0 / (nx - 1))
    rho_lims = (0.3, 0.3 + hx * (nx - 1), -0.7, -0.7 + hy * (ny - 1))
    ps = spacecharge.PoissonSolverFFT2D(nx, ny, *solver_lims)
    arr = r.normal(size=(nx, ny))
    rho = spacecharge.Grid2D(arr, *rho_lims)
    phi = spacecharge.Grid2D(nx, ny, *rho_lims)
    ps.findPotential(rho, phi)
    assert rel_l2(phi.array(), O.poisson2d(arr, solver_lims, rho_lims)) < 1e-12
    bad = spacecharge.Grid2D(nx, ny, 0.0, 1.0, 0.0, 1.0)  # step ratio differs
    with pytest.raises(Exception):
        ps.findPotential(bad, spacecharge.Grid2D(nx, ny, 0.0, 1.0, 0.0, 1.0))


def test_sc2p5drb_with_identity_denoiser_equals_plain():
    """An untrained denoiser (zero head: output = input) changes rho only by its FP32
    normalization; the kicks match the plain calculator to that level and the total
    charge is kept exactly."""
    import spacecharge
    from denoiser import Denoiser
    from workloads import gaussian_coords
    c = gaussian_coords(20000, seed=12)
    b1, _ = _bunches(c)
    b2, _ = _bunches(c)
    plain = spacecharge.SpaceChargeCalc2p5Drb(64, 64, 16)
    calc = spacecharge.SpaceChargeCalc2p5Drb(64, 64, 16)
    dn = Denoiser(64, 64, max_batch=1)
    dn.setParams(Denoiser.initParams(1))
    calc.setDenoiser(dn)
    plain.trackBunch(b1, 0.5, 0.1)
    calc.trackBunch(b2, 0.5, 0.1)
    o1, o2 = b1.getCoordinates(), b2.getCoordinates()
    r1, r2 = plain.getRhoGrid().array(), calc.getRhoGrid().array()
    assert rel_l2(r2, r1) < 1e-6
    assert r2.sum() == pytest.approx(r1.sum(), rel=1e-12)
    for k in (1, 3):
        assert rel_l2(o2[k] - c[k], o1[k] - c[k]) < 1e-5, k
    calc.setDenoiser(None)
    assert calc.getDenoiser() is None
