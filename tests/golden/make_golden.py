"""Write the golden fixtures of the Grid deposition/interpolation functions from the
reference's own Grid1D/Grid2D/Grid3D (oracle/_ref, compiled from the reference sources).

    python tests/golden/make_golden.py      # needs oracle/_ref/libref_grid.so

Output: tests/golden/grid_golden.npz -- inputs (seeded, including particles on edges,
on nodes and outside the grid) and the reference outputs.  tests/test_golden.py checks the
oracle against it bit for bit, wherever the reference tree is absent.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as O  # noqa: E402


def pts(n, lo, hi, seed):
    r = np.random.default_rng(seed)
    v = r.uniform(lo, hi, n)
    v[:4] = [lo, hi, lo - 1e-3 * (hi - lo), hi + 1e-3 * (hi - lo)]
    return v


def main():
    ref = O.ref_grid()
    out = {}
    n = 500
    # 2D: 16 x 12 grid
    l2 = np.array([-3e-3, 4e-3, -2e-3, 2.5e-3])
    x, y = pts(n, l2[0], l2[1], 1), pts(n, l2[2], l2[3], 2)
    x[4:10] = l2[0] + (l2[1] - l2[0]) / 15 * np.arange(6)  # on nodes
    m = np.random.default_rng(3).uniform(0.5, 1.5, n)
    arr2 = np.random.default_rng(4).normal(size=(16, 12))
    out.update(g2_lims=l2, g2_x=x, g2_y=y, g2_m=m, g2_arr=arr2,
               g2_rho=ref.grid2d_bin(16, 12, l2, x, y, m))
    ix, fx, iy, fy = ref.grid2d_indfrac(16, 12, l2, x, y)
    ex, ey, val = ref.grid2d_grad(16, 12, l2, arr2, x, y)
    out.update(g2_ix=ix, g2_fx=fx, g2_iy=iy, g2_fy=fy, g2_ex=ex, g2_ey=ey, g2_val=val)
    # 1D: 8 cells, periodic
    z = pts(n, -100.0, 100.0, 5)
    arr1 = np.random.default_rng(7).normal(size=8)
    v1, gr1 = ref.grid1d_value(8, -100.0, 100.0, arr1, z)
    out.update(g1_z=z, g1_m=m, g1_arr=arr1, g1_bin=ref.grid1d_bin(8, -100.0, 100.0, z, m),
               g1_val=v1, g1_grad=gr1)
    # 3D: 10 x 9 x 7 (and nz = 2)
    l3 = np.array([-3e-3, 3e-3, -2e-3, 2e-3, -0.5, 0.7])
    x3, y3, z3 = pts(n, l3[0], l3[1], 8), pts(n, l3[2], l3[3], 9), pts(n, l3[4], l3[5], 10)
    arr3 = np.random.default_rng(12).normal(size=(7, 10, 9))
    ind, frac = ref.grid3d_indfrac(10, 9, 7, l3, x3, y3, z3)
    gx, gy, gz = ref.grid3d_grad(10, 9, 7, l3, arr3, x3, y3, z3)
    out.update(g3_lims=l3, g3_x=x3, g3_y=y3, g3_z=z3, g3_arr=arr3,
               g3_rho=ref.grid3d_bin(10, 9, 7, l3, x3, y3, z3, m), g3_ind=ind, g3_frac=frac,
               g3_gx=gx, g3_gy=gy, g3_gz=gz,
               g3n2_rho=ref.grid3d_bin(10, 9, 2, l3, x3, y3, z3, m))
    path = os.path.join(HERE, "grid_golden.npz")
    np.savez_compressed(path, **out)
    print(path)


if __name__ == "__main__":
    main()
