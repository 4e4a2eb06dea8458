// ob_sc_device.hpp -- device-side grid index/weight rules and the space-charge kicks.
//
// The index and fraction rules reproduce Grid2D.cc:337-350 (node-centred x,y with the
// [1, n-2] clamp), Grid1D.cc:568-626 (cell-centred, periodic at both ends) and
// Grid3D.cc:226-264 (cell-centred z clamped to [1, nz-2]).  With -ffp-contract=off the
// device computes the same bits as the reference on x86-64: bin indices are exact.
#pragma once

#include <cfloat>

#include "ob_internal.hpp"

namespace ob {

template <typename T>
__device__ __forceinline__ T op_ptr(double d) {
  return reinterpret_cast<T>(__double_as_longlong(d));
}

// Every buffer of the library is global memory, but a pointer read from a struct in device
// memory is generic to the compiler: its accesses become flat instructions, which count on
// the vector-memory and the LDS/scalar counters alike, so each wait for one of them drains
// both.  Accesses through a gptr are global instructions; a uniform load through one that
// no store of the kernel can precede becomes a scalar load.
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ gptr<T> to_global(T* p) {
  return (gptr<T>)p;
}
// A 16-byte pair for loads and stores through a gptr (HIP's double2 class cannot be
// assigned in another address space).
typedef double dpair __attribute__((ext_vector_type(2)));
// atomicAdd (relaxed, device scope) through a gptr
template <typename T>
__device__ __forceinline__ void global_add(gptr<T> p, T v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Grid2D::getIndAndFracX/Y (Grid2D.cc:337-350); Grid3D.cc:212-224 is identical.
__device__ __forceinline__ void node_ind_frac(double v, double vmin, double d, int n, int& ind,
                                              double& frac) {
  int i = (int)((v - vmin) / d + 0.5);
  if (i < 1) i = 1;
  if (i > (n - 2)) i = n - 2;
  ind = i;
  frac = (v - (vmin + i * d)) / d;
}

// Grid1D::getIndAndWZ (Grid1D.cc:568-626), quirk: periodic wrap at both ends.
__device__ __forceinline__ void grid1d_ind_wz(double z, double zMin, double dz, int nz, int& i0,
                                              int& ip, double& w0, double& wp) {
  if (nz > 1) {
    const double ind_dbl = ((z - zMin) / dz) - 0.5;
    int iZ0 = (int)ind_dbl;
    // In range for every z in [zMin, zMax]; the clamp only keeps a NaN coordinate (which
    // the reference aborts on, Grid1D.cc:578-580) from addressing outside the grid.
    iZ0 = iZ0 < 0 ? 0 : (iZ0 > nz - 1 ? nz - 1 : iZ0);
    const double zFrac = (z - (zMin + (iZ0 + 0.5) * dz)) / dz;
    if (iZ0 == 0 && zFrac < 0.) {
      i0 = nz - 1; ip = 0; w0 = -zFrac; wp = 1.0 + zFrac;
      return;
    }
    if (iZ0 == nz - 1) {
      if (zFrac > 0.) {
        i0 = iZ0; ip = 0; w0 = 1.0 - zFrac; wp = zFrac;
      } else {
        ip = iZ0; i0 = nz - 2; w0 = -zFrac; wp = 1.0 + zFrac;
      }
      return;
    }
    i0 = iZ0; ip = iZ0 + 1; w0 = 1.0 - zFrac; wp = zFrac;
  } else {
    i0 = 0; ip = 0; w0 = 1.0; wp = 0.0;
  }
}

// Grid1D::getValue (Grid1D.cc:191-210): zero outside [zMin, zMax].
template <typename P>
__device__ __forceinline__ double grid1d_value(P arr, double z, double zMin, double zMax,
                                               double dz, int nz) {
  if (z < zMin || z > zMax) return 0.;
  int i0, ip;
  double w0, wp;
  grid1d_ind_wz(z, zMin, dz, nz, i0, ip, w0, wp);
  if (nz > 1) return w0 * arr[i0] + wp * arr[ip];
  return arr[0];
}

// Grid2D::calcGradient(x, y) (Grid2D.cc:212-256) on a row-major [nx][ny] array, in two
// stages: the 3 x 3 values around the particle's node, then the weighted sums.
struct Grad2DLoads {
  double xF, yF;
  double a00, a01, a02, a10, a11, a12, a20, a21, a22;
};
template <typename P>
__device__ __forceinline__ void grid2d_gradient_loads(P a, int nx, int ny, double xMin, double dx,
                                                      double yMin, double dy, double x, double y,
                                                      Grad2DLoads& g) {
  int iX, iY;
  node_ind_frac(x, xMin, dx, nx, iX, g.xF);
  node_ind_frac(y, yMin, dy, ny, iY, g.yF);
  const auto r0 = a + (int64_t)(iX - 1) * ny + iY;
  const auto r1 = a + (int64_t)iX * ny + iY;
  const auto r2 = a + (int64_t)(iX + 1) * ny + iY;
  g.a00 = r0[-1]; g.a01 = r0[0]; g.a02 = r0[1];
  g.a10 = r1[-1]; g.a11 = r1[0]; g.a12 = r1[1];
  g.a20 = r2[-1]; g.a21 = r2[0]; g.a22 = r2[1];
}
__device__ __forceinline__ void grid2d_gradient_sums(const Grad2DLoads& g, double dx, double dy,
                                                     double& ex, double& ey) {
  const double xF = g.xF, yF = g.yF;
  const double a00 = g.a00, a01 = g.a01, a02 = g.a02, a10 = g.a10, a11 = g.a11, a12 = g.a12;
  const double a20 = g.a20, a21 = g.a21, a22 = g.a22;
  const double xF2 = xF * xF, yF2 = yF * yF;
  const double Wxm = 0.5 * (0.25 - xF + xF2), Wx0 = 0.75 - xF2, Wxp = 0.5 * (0.25 + xF + xF2);
  const double Wym = 0.5 * (0.25 - yF + yF2), Wy0 = 0.75 - yF2, Wyp = 0.5 * (0.25 + yF + yF2);
  const double dWxm = (-1.0) * (0.5 - xF), dWx0 = (-1.0) * (+2.) * xF,
               dWxp = (-1.0) * (-(0.5 + xF));
  const double dWym = (-1.0) * (0.5 - yF), dWy0 = (-1.0) * (+2.) * yF,
               dWyp = (-1.0) * (-(0.5 + yF));
  double e = dWxm * Wym * a00 + dWxm * Wy0 * a01 + dWxm * Wyp * a02 + dWx0 * Wym * a10 +
             dWx0 * Wy0 * a11 + dWx0 * Wyp * a12 + dWxp * Wym * a20 + dWxp * Wy0 * a21 +
             dWxp * Wyp * a22;
  ex = e / dx;
  e = Wxm * dWym * a00 + Wxm * dWy0 * a01 + Wxm * dWyp * a02 + Wx0 * dWym * a10 +
      Wx0 * dWy0 * a11 + Wx0 * dWyp * a12 + Wxp * dWym * a20 + Wxp * dWy0 * a21 +
      Wxp * dWyp * a22;
  ey = e / dy;
}
template <typename P>
__device__ __forceinline__ void grid2d_gradient(P a, int nx, int ny, double xMin, double dx,
                                                double yMin, double dy, double x, double y,
                                                double& ex, double& ey) {
  Grad2DLoads g;
  grid2d_gradient_loads(a, nx, ny, xMin, dx, yMin, dy, x, y, g);
  grid2d_gradient_sums(g, dx, dy, ex, ey);
}

// Grid3D::getGridIndAndFrac z part (Grid3D.cc:236-263); wrap = longWrapping.
__device__ __forceinline__ void grid3d_z_ind_frac(double z, double zMin, double dz, int nz,
                                                  int& zi, double& zf, bool wrap = false) {
  if (nz > 1) {
    zi = (int)((z - zMin) / dz);
    zf = 0.0;
    if (nz > 2) {
      if (!wrap) {
        if (zi <= 0) zi = 1;
        if (zi >= nz - 1) zi = nz - 2;
      } else {
        if (zi <= 0) zi = 0;
        if (zi >= nz - 1) zi = nz - 1;
      }
      zf = (z - (zMin + (zi + 0.5) * dz)) / dz;
    }
    if (nz == 2) {
      zf = (z - (zMin + 0.5 * dz)) / dz;
      if (zi < 0) { zi = 0; zf = 0.0; }
      else if (zi > 1) { zi = 1; zf = 1.0; }
    }
  } else {
    zi = 0;
    zf = 0.0;
  }
}

// A window of a z-fastest 3D array: node (iX, iY, iZ) is at
// a[(iX - x0)*sx + (iY - y0)*sy + (iZ - z0)].  The whole grid is the window with origin 0,
// sx = ny*nz, sy = nz; an LDS copy of a tile is a window with the tile's origin.
struct Grid3View {
  const double* a;
  int64_t sx, sy;
  int x0, y0, z0;
  __device__ __forceinline__ const double* at(int iX, int iY, int iZ) const {
    return a + (int64_t)(iX - x0) * sx + (int64_t)(iY - y0) * sy + (iZ - z0);
  }
};

// Grid3D::calcSheetGradient (Grid3D.cc:718-733).  The device keeps 3D grids z-fastest,
// a[(ix*ny + iy)*nz + iz] (the layout of the reference's FFT buffer, PoissonSolverFFT3D.cc:
// 305-309); the reference's Arr3D[iz][ix][iy] order is restored only when copied out.
__device__ __forceinline__ double sheet_grad(const Grid3View& v, int iZ, int iX, int iY,
                                             double xm, double x0, double xp, double ym,
                                             double y0, double yp) {
  const int64_t sx = v.sx, sy = v.sy;
  const double* s = v.at(iX, iY, iZ);
  return xm * ym * s[-sx - sy] + x0 * ym * s[-sy] + xp * ym * s[sx - sy] + xm * y0 * s[-sx] +
         x0 * y0 * s[0] + xp * y0 * s[sx] + xm * yp * s[-sx + sy] + x0 * yp * s[sy] +
         xp * yp * s[sx + sy];
}

// Grid3D::calcGradient (Grid3D.cc:438-532): TSC weights and their derivatives at the
// particle's node (iX, iY, iZ), combined over the 27 surrounding nodes.
__device__ __forceinline__ void grid3d_gradient_at(const Grid3View& a, int nz, int iX, int iY,
                                                   int iZ, double xF, double yF, double zF,
                                                   double dx, double dy, double dz, double& gX,
                                                   double& gY, double& gZ) {
  const double Wxm = 0.5 * (0.5 - xF) * (0.5 - xF), Wx0 = 0.75 - xF * xF,
               Wxp = 0.5 * (0.5 + xF) * (0.5 + xF);
  const double Wym = 0.5 * (0.5 - yF) * (0.5 - yF), Wy0 = 0.75 - yF * yF,
               Wyp = 0.5 * (0.5 + yF) * (0.5 + yF);
  double Wzm = 0.0, Wz0 = 0.0, Wzp = 0.0;
  if (nz >= 3) {
    Wzm = 0.5 * (0.5 - zF) * (0.5 - zF);
    Wz0 = 0.75 - zF * zF;
    Wzp = 0.5 * (0.5 + zF) * (0.5 + zF);
  } else if (nz == 2) {
    Wzm = 1.0 - zF; Wz0 = 0.0; Wzp = zF;
  }
  const double dWxm = (-1.0) * (0.5 - xF) / dx, dWx0 = (-1.0) * 2. * xF / dx,
               dWxp = (-1.0) * -(0.5 + xF) / dx;
  const double dWym = (-1.0) * (0.5 - yF) / dy, dWy0 = (-1.0) * 2. * yF / dy,
               dWyp = (-1.0) * -(0.5 + yF) / dy;
  double dWzm = 0.0, dWz0 = 0.0, dWzp = 0.0;
  if (nz >= 3) {
    dWzm = (-1.0) * (0.5 - zF) / dz;
    dWz0 = (-1.0) * 2. * zF / dz;
    dWzp = (-1.0) * -(0.5 + zF) / dz;
  } else if (nz == 2) {
    dWzm = (-1.0) * 1.0 / dz;
    dWz0 = 0.0;
    dWzp = (-1.0) * -1.0 / dz;
  }
  int iZm = iZ - 1, iZp = iZ + 1;
  if (iZm < 0) iZm = nz - 1;
  if (iZp >= nz) iZp = 0;
  if (nz >= 3) {
    gX = sheet_grad(a, iZm, iX, iY, dWxm, dWx0, dWxp, Wym, Wy0, Wyp) * Wzm +
         sheet_grad(a, iZ, iX, iY, dWxm, dWx0, dWxp, Wym, Wy0, Wyp) * Wz0 +
         sheet_grad(a, iZp, iX, iY, dWxm, dWx0, dWxp, Wym, Wy0, Wyp) * Wzp;
    gY = sheet_grad(a, iZm, iX, iY, Wxm, Wx0, Wxp, dWym, dWy0, dWyp) * Wzm +
         sheet_grad(a, iZ, iX, iY, Wxm, Wx0, Wxp, dWym, dWy0, dWyp) * Wz0 +
         sheet_grad(a, iZp, iX, iY, Wxm, Wx0, Wxp, dWym, dWy0, dWyp) * Wzp;
    gZ = sheet_grad(a, iZm, iX, iY, Wxm, Wx0, Wxp, Wym, Wy0, Wyp) * dWzm +
         sheet_grad(a, iZ, iX, iY, Wxm, Wx0, Wxp, Wym, Wy0, Wyp) * dWz0 +
         sheet_grad(a, iZp, iX, iY, Wxm, Wx0, Wxp, Wym, Wy0, Wyp) * dWzp;
  } else if (nz == 2) {
    gX = sheet_grad(a, 0, iX, iY, dWxm, dWx0, dWxp, Wym, Wy0, Wyp) * Wzm +
         sheet_grad(a, 1, iX, iY, dWxm, dWx0, dWxp, Wym, Wy0, Wyp) * Wzp;
    gY = sheet_grad(a, 0, iX, iY, Wxm, Wx0, Wxp, dWym, dWy0, dWyp) * Wzm +
         sheet_grad(a, 1, iX, iY, Wxm, Wx0, Wxp, dWym, dWy0, dWyp) * Wzp;
    gZ = sheet_grad(a, 0, iX, iY, Wxm, Wx0, Wxp, Wym, Wy0, Wyp) * dWzm +
         sheet_grad(a, 1, iX, iY, Wxm, Wx0, Wxp, Wym, Wy0, Wyp) * dWzp;
  } else {
    gX = sheet_grad(a, 0, iX, iY, dWxm, dWx0, dWxp, Wym, Wy0, Wyp);
    gY = sheet_grad(a, 0, iX, iY, Wxm, Wx0, Wxp, dWym, dWy0, dWyp);
    gZ = 0.0;
  }
}

// Grid3D::calcGradient (Grid3D.cc:438-532) at (x, y, z) on the whole grid a.
__device__ __forceinline__ void grid3d_gradient(const double* __restrict__ a, int nx, int ny,
                                                int nz, double xMin, double dx, double yMin,
                                                double dy, double zMin, double dz, double x,
                                                double y, double z, double& gX, double& gY,
                                                double& gZ, bool wrap = false) {
  int iX, iY, iZ;
  double xF, yF, zF;
  node_ind_frac(x, xMin, dx, nx, iX, xF);
  node_ind_frac(y, yMin, dy, ny, iY, yF);
  grid3d_z_ind_frac(z, zMin, dz, nz, iZ, zF, wrap);
  const Grid3View v{a, (int64_t)ny * nz, nz, 0, 0, 0};
  grid3d_gradient_at(v, nz, iX, iY, iZ, xF, yF, zF, dx, dy, dz, gX, gY, gZ);
}

// SpaceChargeCalc2p5Drb kick loop body (SpaceChargeCalc2p5Drb.cc:146-167), and, with
// long_kick 0, the SpaceChargeCalc2p5D one (SpaceChargeCalc2p5D.cc:105-119), in three
// parts: the calculator's values for the step (read once per workgroup by the push
// kernels), the particle's loads (3 x 3 potential values, line density and its derivative
// at z), and the arithmetic.  A push issues the next particle's coordinate loads between
// the second and the third, so that the wait for one particle's loads is not a wait for the
// next one's.
// op: p[0]=SC2DState*, p[1]=phi, p[2]=z line density, p[3]=its derivative,
//     p[4]=2*log(pipe_radius), p[5]=-L*r0*m/gamma^2; i0=nx, i1=ny, i2=nz.
struct SC2DKickParams {
  double xMin, dx, yMin, dy, zMin, zMax, dz, factor, x_c, y_c, a2, lsf_in;
  int32_t skip, long_kick, bnd_type;
  gptr<const double> bnd;  // the boundary's values, read per particle only when it has one
};
__device__ __forceinline__ SC2DKickParams sc2d_kick_params(const SC2DState* state) {
  const gptr<const SC2DState> st = to_global(state);
  SC2DKickParams k;
  k.xMin = st->xMin;
  k.dx = st->dx;
  k.yMin = st->yMin;
  k.dy = st->dy;
  k.zMin = st->zMin;
  k.zMax = st->zMax;
  k.dz = st->dz;
  k.factor = st->factor;
  k.x_c = st->x_c;
  k.y_c = st->y_c;
  k.a2 = st->a2;
  k.lsf_in = st->lsf_in;
  k.skip = st->skip;
  k.long_kick = st->long_kick;
  k.bnd_type = st->bnd_type;
  k.bnd = st->bnd;
  return k;
}

struct SC2DKickLoads {
  Grad2DLoads g;
  double lam0, lamp, ez0, ezp, w0, wp;
  bool on;   // kicked: not skipped, inside the boundary (SpaceChargeCalc2p5D.cc:109)
  bool zin;  // z within [zMin, zMax] (Grid1D::getValue, Grid1D.cc:191-210)
};

// phi: row-major [nx][ny] potential (global memory or an LDS copy); zg, zd: line density
// and its derivative (nz values).
template <typename P, typename Z>
__device__ __forceinline__ void sc2p5drb_kick_loads(const SC2DKickParams& k, P phi, int nx,
                                                    int ny, Z zg, Z zd, int nz, double x,
                                                    double y, double z, SC2DKickLoads& L) {
  L.on = !k.skip && (!k.bnd_type || boundary_inside(k.bnd_type, k.bnd, x, y));
  L.zin = !(z < k.zMin || z > k.zMax);
  if (!L.on) return;
  grid2d_gradient_loads(phi, nx, ny, k.xMin, k.dx, k.yMin, k.dy, x, y, L.g);
  if (!L.zin) return;
  int i0, ip;
  grid1d_ind_wz(z, k.zMin, k.dz, nz, i0, ip, L.w0, L.wp);
  L.lam0 = zg[i0];
  L.lamp = zg[ip];
  if (k.long_kick) {
    L.ez0 = zd[i0];
    L.ezp = zd[ip];
  }
}

__device__ __forceinline__ double sc2d_line_value(const SC2DKickLoads& L, double v0, double vp,
                                                  int nz) {
  if (!L.zin) return 0.;
  if (nz > 1) return L.w0 * v0 + L.wp * vp;
  return v0;
}

__device__ __forceinline__ void sc2p5drb_kick_apply(const SC2DKickParams& k,
                                                    const SC2DKickLoads& L, int nz,
                                                    double pipe_log2, double long_coef,
                                                    double x, double& xp, double y,
                                                    double& yp, double& dE) {
  if (!L.on) return;
  double ex, ey;
  grid2d_gradient_sums(L.g, k.dx, k.dy, ex, ey);
  const double Lfactor = -sc2d_line_value(L, L.lam0, L.lamp, nz) * k.factor;
  xp += ex * Lfactor;
  yp += ey * Lfactor;
  if (!k.long_kick) return;  // SpaceChargeCalc2p5D.cc:105-119: transverse only
  const double xc = k.x_c, yc = k.y_c;
  const double r2 = (x - xc) * (x - xc) + (y - yc) * (y - yc);
  const double ez = sc2d_line_value(L, L.ez0, L.ezp, nz);
  double coeff;
  if (r2 <= k.a2)
    coeff = k.lsf_in - r2 / k.a2;
  else
    coeff = pipe_log2 - log(r2);
  dE += ez * long_coef * coeff;
}

// The whole kick of one particle (the op-program interpreter's form).
__device__ __forceinline__ void sc2p5drb_kick(const Op& op, double x, double& xp, double y,
                                              double& yp, double z, double& dE) {
  const SC2DKickParams k = sc2d_kick_params(op_ptr<const SC2DState*>(op.p[0]));
  SC2DKickLoads L;
  sc2p5drb_kick_loads(k, to_global(op_ptr<const double*>(op.p[1])), op.i0, op.i1,
                      to_global(op_ptr<const double*>(op.p[2])),
                      to_global(op_ptr<const double*>(op.p[3])), op.i2, x, y, z, L);
  sc2p5drb_kick_apply(k, L, op.i2, op.p[4], op.p[5], x, xp, y, yp, dE);
}

// SpaceChargeCalc3D kick loop body (SpaceChargeCalc3D.cc:132-147).
// op: p[0]=phi (z-fastest), p[1..6]=xMin,dx,yMin,dy,zMin,dz of the rest-frame phi grid,
//     p[7]=z_center, p[8]=gamma, p[9]=trans_factor, p[10]=long_factor, p[11]=spacing of
//     external bunches in the rest frame (0: none); i0=nx, i1=ny, i2=nz | kSC3DWrapBit when
//     the potential grid wraps periodically in z (Grid3D longWrapping).
constexpr int kSC3DWrapBit = 1 << 30;
constexpr int kSC3DNzMask = kSC3DWrapBit - 1;
__device__ __forceinline__ void sc3d_kick(const Op& op, double x, double& xp, double y,
                                          double& yp, double z, double& dE) {
  const double zc = op.p[7], gamma = op.p[8], lambda_cm = op.p[11];
  double zr = (z - zc) * gamma + zc;
  if (lambda_cm != 0.0) zr = remainder(zr, lambda_cm);  // external bunches (:138-140)
  double ex, ey, ez;
  const int nz = op.i2 & kSC3DNzMask;
  const bool wrap = (op.i2 & kSC3DWrapBit) != 0;
  grid3d_gradient(op_ptr<const double*>(op.p[0]), op.i0, op.i1, nz, op.p[1], op.p[2], op.p[3],
                  op.p[4], op.p[5], op.p[6], x, y, zr, ex, ey, ez, wrap);
  xp += -ex * op.p[9];
  yp += -ey * op.p[9];
  dE += -ez * op.p[10];
}

// Exclusive block scan of two counters (blockDim.x a multiple of 64, at most 1024).
__device__ __forceinline__ void block_scan2(uint32_t& a, uint32_t& b, uint32_t& ta,
                                           uint32_t& tb) {
  __shared__ uint32_t wa[16], wb[16];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  uint32_t ia = a, ib = b;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t xa = __shfl_up(ia, o, 64), xb = __shfl_up(ib, o, 64);
    if (lane >= o) {
      ia += xa;
      ib += xb;
    }
  }
  if (lane == 63) {
    wa[w] = ia;
    wb[w] = ib;
  }
  __syncthreads();
  uint32_t pa = 0, pb = 0;
  ta = 0;
  tb = 0;
  for (int k = 0; k < nw; k++) {
    if (k < w) {
      pa += wa[k];
      pb += wb[k];
    }
    ta += wa[k];
    tb += wb[k];
  }
  a = pa + ia - a;
  b = pb + ib - b;
}

}  // namespace ob
