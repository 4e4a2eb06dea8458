"""The oracle against golden vectors written by the reference's own Grid classes
(tests/golden/make_golden.py).  Bit-exact; runs without the reference tree."""
import os

import numpy as np

import oracle as O

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "grid_golden.npz"))


def test_grid2d_golden():
    orc = O.orc_grid()
    l, x, y, m = G["g2_lims"], G["g2_x"], G["g2_y"], G["g2_m"]
    assert np.array_equal(orc.grid2d_bin(16, 12, l, x, y, m), G["g2_rho"])
    ix, fx, iy, fy = orc.grid2d_indfrac(16, 12, l, x, y)
    assert np.array_equal(ix, G["g2_ix"]) and np.array_equal(iy, G["g2_iy"])
    assert np.array_equal(fx, G["g2_fx"]) and np.array_equal(fy, G["g2_fy"])
    ex, ey, val = orc.grid2d_grad(16, 12, l, G["g2_arr"], x, y)
    assert np.array_equal(ex, G["g2_ex"]) and np.array_equal(ey, G["g2_ey"])
    assert np.array_equal(val, G["g2_val"])


def test_grid1d_golden():
    orc = O.orc_grid()
    z, m = G["g1_z"], G["g1_m"]
    assert np.array_equal(orc.grid1d_bin(8, -100.0, 100.0, z, m), G["g1_bin"])
    v, g = orc.grid1d_value(8, -100.0, 100.0, G["g1_arr"], z)
    assert np.array_equal(v, G["g1_val"]) and np.array_equal(g, G["g1_grad"])


def test_grid3d_golden():
    orc = O.orc_grid()
    l, x, y, z, m = G["g3_lims"], G["g3_x"], G["g3_y"], G["g3_z"], G["g2_m"]
    assert np.array_equal(orc.grid3d_bin(10, 9, 7, l, x, y, z, m), G["g3_rho"])
    assert np.array_equal(orc.grid3d_bin(10, 9, 2, l, x, y, z, m), G["g3n2_rho"])
    ind, frac = orc.grid3d_indfrac(10, 9, 7, l, x, y, z)
    assert np.array_equal(ind, G["g3_ind"]) and np.array_equal(frac, G["g3_frac"])
    gx, gy, gz = orc.grid3d_grad(10, 9, 7, l, G["g3_arr"], x, y, z)
    assert np.array_equal(gx, G["g3_gx"]) and np.array_equal(gy, G["g3_gy"])
    assert np.array_equal(gz, G["g3_gz"])


def test_golden_shapes():
    # the fixture carries edge, node and outside particles
    x, l = G["g2_x"], G["g2_lims"]
    assert x[0] == l[0] and x[1] == l[1] and x[2] < l[0] and x[3] > l[1]
