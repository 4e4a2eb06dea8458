// ob_push.hip -- the fused TEAPOT map kernel.
//
// One launch applies a program of queued maps (drift, quad1, quad2, quad/multipole
// fringes, multipole kicks, wraps, rotations, plain kicks) and the deferred
// space-charge kicks to every particle, holding the six FP64 coordinates in
// registers: each particle is read once and written once per launch (96 B) however
// many maps the program holds.  When the program feeds a space-charge node it also
// emits per-block min/max of x, y, z for the next grid (BunchExtremaCalculator.cc:
// 37-83), so the extrema pass costs no extra HBM read.
//
// Arithmetic follows the reference routine of each map operation for operation
// (cited per case), compiled with -ffp-contract=off so nothing is fused: results
// are bit-identical to src/teapot/teapotbase.cc built with g++ for x86-64.
#include "ob_internal.hpp"
#include "ob_sc_device.hpp"

namespace ob {

namespace {

struct cplx {
  double re, im;
};
__device__ __forceinline__ cplx cmul(cplx a, cplx b) {  // std::complex<double> product
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ cplx cadd(cplx a, cplx b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cplx csub(cplx a, cplx b) { return {a.re - b.re, a.im - b.im}; }

// teapotbase.cc:395-490 / 508-603.  p[0]=dp_p_coeff, p[1..2]=kk (+/-kterm), i0=pole.
__device__ __forceinline__ void multpfringe(const Op& op, double& x, double& xp, double& y,
                                            double& yp, double& z, double dE) {
  const int lm1 = op.i0, l = op.i0 + 1, lp2 = op.i0 + 3;
  const cplx rootm1 = {0.0, 1.0};
  const cplx cxl = {(double)l, 0.0}, cxlp2 = {(double)lp2, 0.0};
  const cplx kk = {op.p[1], op.p[2]};
  const double px = xp, py = yp;
  const double dp_p = dE * op.p[0];
  const double KNL = 1.0 / (1.0 + dp_p);
  const cplx zz = {x, y};
  cplx zlm1 = {1.0, 0.0};
  for (int k = 0; k < lm1; k++) zlm1 = cmul(zlm1, zz);
  const cplx zl = cmul(zlm1, zz);
  const cplx fxterm = {l * x, -lp2 * y};
  const cplx dxfxterm = cmul(cxl, cadd(fxterm, zz));
  const cplx dyfxterm = cmul(rootm1, csub(cmul(cxl, fxterm), cmul(cxlp2, zz)));
  const cplx fyterm = {l * y, lp2 * x};
  const cplx dxfyterm = cadd(cmul(cxl, fyterm), cmul(cmul(rootm1, cxlp2), zz));
  const cplx dyfyterm = cmul(cxl, cadd(cmul(rootm1, fyterm), zz));
  const cplx fxcx = cmul(cmul(kk, zl), fxterm);
  const cplx dxfxcx = cmul(cmul(kk, zlm1), dxfxterm);
  const cplx dyfxcx = cmul(cmul(kk, zlm1), dyfxterm);
  const cplx fycx = cmul(cmul(kk, zl), fyterm);
  const cplx dxfycx = cmul(cmul(kk, zlm1), dxfyterm);
  const cplx dyfycx = cmul(cmul(kk, zlm1), dyfyterm);
  x -= fxcx.re * KNL;
  y -= fycx.re * KNL;
  const double M11 = 1.0 - dxfxcx.re * KNL;
  const double M12 = -dxfycx.re * KNL;
  const double M21 = -dyfxcx.re * KNL;
  const double M22 = 1.0 - dyfycx.re * KNL;
  const double detM = M11 * M22 - M12 * M21;
  const double pxnew = (M22 * px - M12 * py) / detM;
  const double pynew = (-M21 * px + M11 * py) / detM;
  xp = pxnew;
  yp = pynew;
  z -= (pxnew * fxcx.re + pynew * fycx.re) * KNL * KNL;
}

// teapotbase.cc:784-835 (IN) / 854-905 (OUT).  p[0]=dp_p_coeff, p[1]=kq.
template <bool OUT>
__device__ __forceinline__ void quadfringe(const Op& op, double& x, double& xp, double& y,
                                           double& yp, double& z, double dE) {
  const double kq = op.p[1];
  const double dp_p = dE * op.p[0];
  const double KNL = 1.0 / (1.0 + dp_p);
  const double x0 = x, xp0 = xp, y0 = y, yp0 = yp;
  const double t = (kq * KNL / 4.) * (x0 * x0 - y0 * y0);
  const double detM = 1.0 - t * t;
  const double dx = (kq * KNL / 12.) * x0 * (x0 * x0 + 3. * y0 * y0);
  const double dxp = (kq * KNL / 4.) * (xp0 * (x0 * x0 + y0 * y0) - 2. * yp0 * x0 * y0);
  const double dy = (kq * KNL / 12.) * y0 * (y0 * y0 + 3. * x0 * x0);
  const double dyp = (kq * KNL / 4.) * (-yp0 * (x0 * x0 + y0 * y0) + 2. * xp0 * x0 * y0);
  const double dz = (kq * KNL * KNL / 12.) *
                    (xp0 * x0 * (x0 * x0 + 3. * y0 * y0) - yp0 * y0 * (y0 * y0 + 3. * x0 * x0));
  if (!OUT) {
    x += dx;
    xp -= dxp;
    xp /= detM;
    y -= dy;
    yp -= dyp;
    yp /= detM;
    z += dz;
  } else {
    x -= dx;
    xp += dxp;
    xp /= detM;
    y += dy;
    yp += dyp;
    yp /= detM;
    z -= dz;
  }
}

// One particle's drift (drifti, teapotbase.cc:120-141).
__device__ __forceinline__ void drift_one(double length, double gamma2i, double dp_p_coeff,
                                          double& x, double xp, double& y, double yp,
                                          double& z, double dE) {
  const double dp_p = dE * dp_p_coeff;
  const double KNL = 1.0 / (1.0 + dp_p);
  x += KNL * length * xp;
  y += KNL * length * yp;
  double phifac = (xp * xp + yp * yp + dp_p * dp_p * gamma2i) / 2.0;
  phifac = (phifac * KNL - dp_p * gamma2i) * KNL;
  z -= length * phifac;
}

// wedgebend / wedgebendCF (teapotbase.cc:1037-1080, 1447-1509).  p[0]=ct, p[1]=rho,
// p[2]=dp_p_coeff, p[3]=gamma2i, p[4+l]=kl of multipole l; i0 = inout | n << 1 |
// skew bits << 8, i1 = nst, i2 = poles, 4 bits each.  The multipole kick of each step is
// multpi (:280-315) with kl * sm, divided by pole! per particle as there.
__device__ __forceinline__ void wedgebend(const Op& op, double& x, double& xp, double& y,
                                          double& yp, double& z, double dE) {
  const double ct = op.p[0], rho = op.p[1], dpc = op.p[2], gamma2i = op.p[3];
  const int inout = op.i0 & 1, nm = (op.i0 >> 1) & 15, skews = (op.i0 >> 8) & 255;
  const int nst = op.i1;
  double s;
  if (inout == 0) {
    s = -x / ct;
  } else {
    const double dp_p = dE * dpc;
    const double tn = xp / (1.0 + dp_p);
    s = -x / (ct + tn);
  }
  const double sm = s / nst, sm2 = sm / 2.0;
  for (int j = 0; j < nst; j++) {
    drift_one(j == 0 ? sm2 : sm, gamma2i, dpc, x, xp, y, yp, z, dE);
    xp -= sm / rho;
    for (int l = 0; l < nm; l++) {
      const int pole = (op.i2 >> (4 * l)) & 15;
      double fact = 1.0;  // teapotbase.cc:45-57
      for (int k = 1; k <= pole; k++) fact = k * fact;
      const double kl1 = (op.p[4 + l] * sm) / fact;
      const cplx zz = {x, y};
      cplx zn = {1.0, 0.0};
      for (int k = 0; k < pole; k++) zn = cmul(zn, zz);
      if ((skews >> l) & 1) {
        xp += kl1 * zn.im;
        yp += kl1 * zn.re;
      } else {
        xp -= kl1 * zn.re;
        yp += kl1 * zn.im;
      }
    }
  }
  drift_one(sm2, gamma2i, dpc, x, xp, y, yp, z, dE);
}

// soln (teapotbase.cc:1367-1425).  p: Bc, length, gamma2i, dp_p_coeff.
__device__ __forceinline__ void soln(const Op& op, double& x, double& xp, double& y,
                                     double& yp, double& z, double dE) {
  const double Bc = op.p[0], length = op.p[1], gamma2i = op.p[2];
  const double dp_p = dE * op.p[3];
  const double KNL = 1.0 / (1.0 + dp_p);
  const double cu = y / 2. - xp / Bc;
  const double cpu = x * Bc / 2. + yp;
  const double u_init = y / 2. + xp / Bc;
  const double pu_init = -x * Bc / 2. + yp;
  const double phase = KNL * Bc * length;
  const double cs = cos(phase), sn = sin(phase);
  const double u = u_init * cs + pu_init * sn / Bc;
  const double pu = -u_init * Bc * sn + pu_init * cs;
  x = (-pu + cpu) / Bc;
  xp = 0.5 * (u - cu) * Bc;
  y = u + cu;
  yp = 0.5 * (pu + cpu);
  double phifac = (pu_init * pu_init + Bc * Bc * u_init * u_init + dp_p * dp_p * gamma2i) / 2.0;
  phifac = (phifac * KNL - dp_p * gamma2i) * KNL;
  z -= length * phifac;
}

// Kernel level an op kind needs: 0 = the lite kernels, 1 = with the multipole fringes
// (a ring's sextupoles), 2 = with every register-hungry op (3D kick, wedge bend, solenoid,
// ring RF).  Each level holds the ones below it; a lower level runs with fewer VGPRs and
// so more waves per SIMD.
__host__ __device__ constexpr int op_level(int kind) {
  return (kind == OP_MULTPFRINGE_IN || kind == OP_MULTPFRINGE_OUT)
             ? 1
             : ((kind == OP_SC3D_KICK || kind == OP_WEDGEBEND || kind == OP_SOLN ||
                 kind == OP_RINGRF)
                    ? 2
                    : 0);
}

template <int LEVEL>
__device__ __forceinline__ void apply_op(const Op& op, double& x, double& xp, double& y,
                                         double& yp, double& z, double& dE) {
  if (op_level(op.kind) > LEVEL) return;  // never reached: the host picks the level
  switch (op.kind) {
    case OP_DRIFT: {  // teapotbase.cc:158-183. p: length, gamma2i, dp_p_coeff
      const double length = op.p[0], gamma2i = op.p[1];
      const double dp_p = dE * op.p[2];
      const double KNL = 1.0 / (1.0 + dp_p);
      x += KNL * length * xp;
      y += KNL * length * yp;
      double phifac = (xp * xp + yp * yp + dp_p * dp_p * gamma2i) / 2.0;
      phifac = (phifac * KNL - dp_p * gamma2i) * KNL;
      z -= length * phifac;
      break;
    }
    case OP_QUAD1: {  // teapotbase.cc:686-699. p: m11 m12 m21 m22 m33 m34 m43 m44, gamma2i,
                      // length, dp_p_coeff
      const double dp_p = dE * op.p[10];
      const double x0 = x, xp0 = xp, y0 = y, yp0 = yp;
      x = x0 * op.p[0] + xp0 * op.p[1];
      xp = x0 * op.p[2] + xp0 * op.p[3];
      y = y0 * op.p[4] + yp0 * op.p[5];
      yp = y0 * op.p[6] + yp0 * op.p[7];
      z += dp_p * op.p[8] * op.p[9];
      break;
    }
    case OP_QUAD2: {  // teapotbase.cc:719-745. p: length, gamma2i, dp_p_coeff
      const double length = op.p[0], gamma2i = op.p[1];
      const double dp_p = dE * op.p[2];
      const double KNL = 1.0 / (1.0 + dp_p);
      x -= KNL * length * dp_p * xp;
      y -= KNL * length * dp_p * yp;
      double phifac = (xp * xp + yp * yp + dp_p * dp_p * gamma2i) / 2.0;
      phifac = (phifac * KNL + dp_p * dp_p * gamma2i) * KNL;
      z -= length * phifac;
      break;
    }
    case OP_QUADFRINGE_IN:
      quadfringe<false>(op, x, xp, y, yp, z, dE);
      break;
    case OP_QUADFRINGE_OUT:
      quadfringe<true>(op, x, xp, y, yp, z, dE);
      break;
    case OP_MULTP: {  // teapotbase.cc:336-372. p[0]=kl/pole!, i0=pole, i1=skew
      const cplx zz = {x, y};
      cplx zn = {1.0, 0.0};
      for (int k = 0; k < op.i0; k++) zn = cmul(zn, zz);
      const double kl1 = op.p[0];
      if (op.i1) {
        xp += kl1 * zn.im;
        yp += kl1 * zn.re;
      } else {
        xp -= kl1 * zn.re;
        yp += kl1 * zn.im;
      }
      break;
    }
    case OP_MULTPFRINGE_IN:
    case OP_MULTPFRINGE_OUT:
      if constexpr (LEVEL >= 1) multpfringe(op, x, xp, y, yp, z, dE);
      break;
    case OP_WRAPBUNCH: {  // teapotbase.cc:202-211. p[0]=length
      const double length = op.p[0];
      if (z < -length / 2.0) z += length;
      if (z > length / 2.0) z -= length;
      break;
    }
    case OP_ROTATEXY: {  // teapotbase.cc:80-101. p: cos, sin
      const double cs = op.p[0], sn = op.p[1];
      const double xt = x, pxt = xp, yt = y, pyt = yp;
      x = cs * xt - sn * yt;
      xp = cs * pxt - sn * pyt;
      y = sn * xt + cs * yt;
      yp = sn * pxt + cs * pyt;
      break;
    }
    case OP_KICK:  // teapotbase.cc:232-257. p: kx, ky, kE
      if (op.p[0] != 0.) xp += op.p[0];
      if (op.p[1] != 0.) yp += op.p[1];
      if (op.p[2] != 0.) dE += op.p[2];
      break;
    case OP_SC2P5DRB_KICK:
      sc2p5drb_kick(op, x, xp, y, yp, z, dE);
      break;
    case OP_SC3D_KICK:
      if constexpr (LEVEL >= 2) sc3d_kick(op, x, xp, y, yp, z, dE);
      break;
    case OP_BEND1: {  // teapotbase.cc:1099-1145. p: m11 m12 m16 m21 m22 m26 m51 m52 m56,
                      // length, dp_p_coeff
      const double dp_p = dE * op.p[10];
      const double x0 = x, xp0 = xp;
      x = x0 * op.p[0] + xp0 * op.p[1] + dp_p * op.p[2];
      xp = x0 * op.p[3] + xp0 * op.p[4] + dp_p * op.p[5];
      y += op.p[9] * yp;
      z += x0 * op.p[6] + xp0 * op.p[7] + dp_p * op.p[8];
      break;
    }
    case OP_BEND3: {  // teapotbase.cc:1208-1233. p: th, gamma2i, dp_p_coeff
      const double th = op.p[0], gamma2i = op.p[1];
      const double dp_p = dE * op.p[2];
      const double KNL = 1.0 / (1.0 + dp_p);
      double phifac = (yp * yp + dp_p * dp_p * gamma2i) / 2.0;
      xp -= phifac * KNL * th;
      y += KNL * yp * x * th;
      phifac = (phifac * KNL - dp_p * gamma2i) * KNL;
      z -= th * phifac * x;
      break;
    }
    case OP_BEND4: {  // teapotbase.cc:1251-1274. p: th, dp_p_coeff
      const double th = op.p[0];
      const double dp_p = dE * op.p[1];
      const double KNL = 1.0 / (1.0 + dp_p);
      const double xfac = 1.0 + KNL * xp * th / 2.0;
      const double phifac = KNL * KNL * xp * xp / 2.0;
      x *= xfac * xfac;
      xp /= xfac;
      z -= th * phifac * x;
      break;
    }
    case OP_BENDFRINGE_IN:    // teapotbase.cc:1292-1311. p: rho, dp_p_coeff
    case OP_BENDFRINGE_OUT: {  // teapotbase.cc:1329-1348
      const double rho = op.p[0];
      const double dp_p = dE * op.p[1];
      const double KNL = 1.0 / (1.0 + dp_p);
      if (op.kind == OP_BENDFRINGE_IN) {
        x += KNL * y * y / (2. * rho);
        yp -= KNL * xp * y / rho;
        z += KNL * KNL * xp * y * y / (2. * rho);
      } else {
        x -= KNL * y * y / (2. * rho);
        yp += KNL * xp * y / rho;
        z -= KNL * KNL * xp * y * y / (2. * rho);
      }
      break;
    }
    case OP_WEDGEROTATE: {  // teapotbase.cc:923-966. p: cos e, sin e, dp_p_coeff; i0=frinout
      const double cs = op.p[0], sn = op.p[1], dpc = op.p[2];
      double dp_p, xp_temp, p0_temp, p0;
      if (op.i0 == 0) {
        dp_p = dE * dpc;
        xp_temp = xp;
        p0_temp = 1.0 + dp_p;
        x /= cs;
        xp = xp_temp * cs + p0_temp * sn;
        p0 = -xp_temp * sn + p0_temp * cs;
        dp_p = p0 - 1.0;
        z = (-x * sn + z) * cs;
      } else {
        dp_p = dE * dpc;
        p0 = 1.0 + dp_p;
        z = x * sn + z / cs;
        x *= cs;
        xp_temp = xp * cs - p0 * sn;
        p0_temp = xp * sn + p0 * cs;
        xp = xp_temp;
        dp_p = p0_temp - 1.0;
      }
      dE = dp_p / dpc;
      break;
    }
    case OP_WEDGEDRIFT: {  // teapotbase.cc:986-1012. p: ct, dp_p_coeff, gamma2i; i0=inout
      const double ct = op.p[0], dpc = op.p[1];
      double s;
      if (op.i0 == 0) {
        const double dp_p = dE * dpc;
        const double tn = xp / (1.0 + dp_p);
        s = x / (ct - tn);
      } else {
        s = x / ct;
      }
      drift_one(s, op.p[2], dpc, x, xp, y, yp, z, dE);
      break;
    }
    case OP_WEDGEBEND:
      if constexpr (LEVEL >= 2) wedgebend(op, x, xp, y, yp, z, dE);
      break;
    case OP_SOLN:
      if constexpr (LEVEL >= 2) soln(op, x, xp, y, yp, z, dE);
      break;
    case OP_RINGRF:  // teapotbase.cc:1527-1566. p: h*2pi/L, phase_s, voltage, coeff, p_in/p_out
      if constexpr (LEVEL >= 2) {
        const double deltaV = op.p[2] * (sin(op.p[0] * z + op.p[1]));
        dE += op.p[3] * deltaV;
        xp *= op.p[4];
        yp *= op.p[4];
      }
      break;
    default:
      break;
  }
}

// The kernel level a program needs (the highest of its ops).
int program_level(const OpProgram& prog) {
  int l = 0;
  for (int k = 0; k < prog.nops; k++) l = std::max(l, op_level(prog.ops[k].kind));
  return l;
}

constexpr int kPushThreads = 256;
constexpr int kMaxPushBlocks = 4096;

__device__ __forceinline__ double wave_min(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

template <typename P>
__device__ __forceinline__ void load6(P a, P b, P c, P d, P e, P f, int64_t i, double q[6]) {
  q[0] = a[i];
  q[1] = b[i];
  q[2] = c[i];
  q[3] = d[i];
  q[4] = e[i];
  q[5] = f[i];
}

// KICK0: the program's first op is a SpaceChargeCalc2p5Drb kick (the usual case: the
// deferred kick of the previous SC node); it runs outside the op interpreter, with the
// calculator's values read once, before any store (scalar loads).  Without it the kernel
// keeps the interpreter's smaller register footprint.
template <bool EXTREMA, int LEVEL, bool KICK0>
__global__ __launch_bounds__(kPushThreads) void push_kernel(DeviceBunchView v, OpProgram prog,
                                                            double* __restrict__ partials) {
  double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX};
  double hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const gptr<double> cx = to_global(v.c[0]);
  const gptr<double> cxp = to_global(v.c[1]);
  const gptr<double> cy = to_global(v.c[2]);
  const gptr<double> cyp = to_global(v.c[3]);
  const gptr<double> cz = to_global(v.c[4]);
  const gptr<double> cdE = to_global(v.c[5]);
  const Op& k0 = prog.ops[0];
  SC2DKickParams kp;
  if constexpr (KICK0) kp = sc2d_kick_params(op_ptr<const SC2DState*>(k0.p[0]));
  // The next particle's coordinates are loaded before this one is computed (after its
  // kick's loads), so their latency overlaps the kick and map arithmetic (few waves fit:
  // the map switch needs many VGPRs).
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double q[6];
  if (i < v.n) load6(cx, cxp, cy, cyp, cz, cdE, i, q);
  for (; i < v.n; i += stride) {
    double x = q[0], xp = q[1], y = q[2], yp = q[3], z = q[4], dE = q[5];
    SC2DKickLoads L;
    if constexpr (KICK0)
      sc2p5drb_kick_loads(kp, to_global(op_ptr<const double*>(k0.p[1])), k0.i0, k0.i1,
                          to_global(op_ptr<const double*>(k0.p[2])),
                          to_global(op_ptr<const double*>(k0.p[3])), k0.i2, x, y, z, L);
    if (i + stride < v.n) load6(cx, cxp, cy, cyp, cz, cdE, i + stride, q);
    if constexpr (KICK0) sc2p5drb_kick_apply(kp, L, k0.i2, k0.p[4], k0.p[5], x, xp, y, yp, dE);
    for (int k = KICK0 ? 1 : 0; k < prog.nops; k++)
      apply_op<LEVEL>(prog.ops[k], x, xp, y, yp, z, dE);
    cx[i] = x;
    cxp[i] = xp;
    cy[i] = y;
    cyp[i] = yp;
    cz[i] = z;
    cdE[i] = dE;
    if (EXTREMA) {
      lo[0] = fmin(lo[0], x);
      hi[0] = fmax(hi[0], x);
      lo[1] = fmin(lo[1], y);
      hi[1] = fmax(hi[1], y);
      lo[2] = fmin(lo[2], z);
      hi[2] = fmax(hi[2], z);
    }
  }
  if (EXTREMA) {
    __shared__ double red[kPushThreads / 64][6];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double r[6];
    for (int k = 0; k < 3; k++) {
      r[2 * k] = wave_min(lo[k]);
      r[2 * k + 1] = wave_max(hi[k]);
    }
    if (lane == 0)
      for (int k = 0; k < 6; k++) red[w][k] = r[k];
    __syncthreads();
    if (threadIdx.x < 6) {
      const int k = threadIdx.x;
      double acc = red[0][k];
      for (int j = 1; j < kPushThreads / 64; j++)
        acc = (k & 1) ? fmax(acc, red[j][k]) : fmin(acc, red[j][k]);
      partials[(int64_t)blockIdx.x * 6 + k] = acc;
    }
  }
}

// Potentials of at most this many nodes (64 x 64) are copied into LDS by each workgroup
// of the batch push: the kick's nine gathers per particle then read LDS, where 64
// random addresses cost a few cycles, instead of the L1/L2 path, where they cost a
// cache-line lookup per lane.
constexpr int kBatchPhiLds = 4096;

// All bunches of a batch in one launch: block = one work item (a particle range of one
// bunch).  The bunch's deferred 2.5D kick comes first (its op is assembled from the
// bunch's record and the batch's common factors), then the map program; extrema are
// reduced per work item.  Dynamic LDS: kBatchPhiLds doubles when the kick is active.
template <bool EXTREMA, int LEVEL>
__device__ __forceinline__ void push_batch_body(const BatchBunch* __restrict__ bb,
                                                const BatchItem* __restrict__ items,
                                                const OpProgram& prog, const BatchKick& kick,
                                                double* __restrict__ partials) {
  const BatchItem it = items[blockIdx.x];
  const BatchBunch& B = bb[it.bunch];
  // the kick's values for this step, before any store of the kernel (scalar loads)
  SC2DKickParams kp;
  if (kick.active) kp = sc2d_kick_params(B.st);
  const int nx = B.nx, ny = B.ny, nz = B.nz;
  const gptr<const double> gphi = to_global((const double*)B.phi);
  const gptr<const double> zg = to_global((const double*)B.zg);
  const gptr<const double> zd = to_global((const double*)B.zd);
  double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX};
  double hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
  const gptr<double> cx = to_global(B.v.c[0]);
  const gptr<double> cxp = to_global(B.v.c[1]);
  const gptr<double> cy = to_global(B.v.c[2]);
  const gptr<double> cyp = to_global(B.v.c[3]);
  const gptr<double> cz = to_global(B.v.c[4]);
  const gptr<double> cdE = to_global(B.v.c[5]);
  extern __shared__ double sphi[];
  const bool lds_phi = kick.active && nx * ny <= kBatchPhiLds;
  if (lds_phi) {  // one load per thread and node slot, all in flight together
    constexpr int KP = kBatchPhiLds / kPushThreads;
    const int ng = nx * ny;
    double v[KP];
#pragma unroll
    for (int u = 0; u < KP; u++) {
      const int k = (int)threadIdx.x + u * kPushThreads;
      v[u] = gphi[k < ng ? k : 0];
    }
#pragma unroll
    for (int u = 0; u < KP; u++) {
      const int k = (int)threadIdx.x + u * kPushThreads;
      if (k < ng) sphi[k] = v[u];
    }
    __syncthreads();
  }
  int64_t i = it.begin + threadIdx.x;
  double q[6];
  if (i < it.end) load6(cx, cxp, cy, cyp, cz, cdE, i, q);
  for (; i < it.end; i += kPushThreads) {
    double x = q[0], xp = q[1], y = q[2], yp = q[3], z = q[4], dE = q[5];
    // this particle's kick loads, then the next particle's coordinates, then the arithmetic
    SC2DKickLoads L;
    if (lds_phi)
      sc2p5drb_kick_loads(kp, sphi, nx, ny, zg, zd, nz, x, y, z, L);
    else if (kick.active)
      sc2p5drb_kick_loads(kp, gphi, nx, ny, zg, zd, nz, x, y, z, L);
    if (i + kPushThreads < it.end) load6(cx, cxp, cy, cyp, cz, cdE, i + kPushThreads, q);
    if (kick.active) sc2p5drb_kick_apply(kp, L, nz, kick.p4, kick.p5, x, xp, y, yp, dE);
    for (int k = 0; k < prog.nops; k++) apply_op<LEVEL>(prog.ops[k], x, xp, y, yp, z, dE);
    cx[i] = x;
    cxp[i] = xp;
    cy[i] = y;
    cyp[i] = yp;
    cz[i] = z;
    cdE[i] = dE;
    if (EXTREMA) {
      lo[0] = fmin(lo[0], x);
      hi[0] = fmax(hi[0], x);
      lo[1] = fmin(lo[1], y);
      hi[1] = fmax(hi[1], y);
      lo[2] = fmin(lo[2], z);
      hi[2] = fmax(hi[2], z);
    }
  }
  if (EXTREMA) {
    __shared__ double red[kPushThreads / 64][6];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double r[6];
    for (int k = 0; k < 3; k++) {
      r[2 * k] = wave_min(lo[k]);
      r[2 * k + 1] = wave_max(hi[k]);
    }
    if (lane == 0)
      for (int k = 0; k < 6; k++) red[w][k] = r[k];
    __syncthreads();
    if (threadIdx.x < 6) {
      const int k = threadIdx.x;
      double acc = red[0][k];
      for (int j = 1; j < kPushThreads / 64; j++)
        acc = (k & 1) ? fmax(acc, red[j][k]) : fmin(acc, red[j][k]);
      partials[(int64_t)blockIdx.x * 6 + k] = acc;
    }
  }
}


// Levels 0 and 1 fit three waves per SIMD (168 VGPRs) without spilling, at most 12 bytes
// at level 1; the bound measured 7% faster than the compiler's own 169-register, two-wave
// allocation.  Level 2 would spill 220 bytes at that bound and keeps two waves.
template <bool EXTREMA, int LEVEL>
__global__ __launch_bounds__(kPushThreads) __attribute__((amdgpu_waves_per_eu(3))) void
push_batch_kernel(const BatchBunch* __restrict__ bb, const BatchItem* __restrict__ items,
                  OpProgram prog, BatchKick kick, double* __restrict__ partials) {
  push_batch_body<EXTREMA, LEVEL>(bb, items, prog, kick, partials);
}
template <bool EXTREMA>
__global__ __launch_bounds__(kPushThreads) void push_batch_kernel_l2(
    const BatchBunch* __restrict__ bb, const BatchItem* __restrict__ items, OpProgram prog,
    BatchKick kick, double* __restrict__ partials) {
  push_batch_body<EXTREMA, 2>(bb, items, prog, kick, partials);
}

}  // namespace

void launch_push_batch(const BatchBunch* bb, const BatchItem* items, int nitems,
                       const OpProgram& prog, const BatchKick& kick, double* partials,
                       int threads, hipStream_t s) {
  OB_REQUIRE(threads == kPushThreads, "batch push: unsupported block size");
  if (nitems == 0) return;
  const int level = program_level(prog);
  const size_t lds = kick.active ? sizeof(double) * kBatchPhiLds : 0;
  auto go = [&](auto kernel, double* part) {
    hipLaunchKernelGGL(kernel, dim3(nitems), dim3(kPushThreads), lds, s, bb, items, prog, kick,
                       part);
  };
  if (prog.want_extrema) {
    OB_REQUIRE(partials != nullptr, "batch push: extrema requested without a buffer");
    ProfScope ps("push_batch_extrema");
    if (level == 2)
      go(push_batch_kernel_l2<true>, partials);
    else if (level == 1)
      go(push_batch_kernel<true, 1>, partials);
    else
      go(push_batch_kernel<true, 0>, partials);
  } else {
    if (prog.nops == 0 && !kick.active) return;
    ProfScope ps("push_batch");
    if (level == 2)
      go(push_batch_kernel_l2<false>, nullptr);
    else if (level == 1)
      go(push_batch_kernel<false, 1>, nullptr);
    else
      go(push_batch_kernel<false, 0>, nullptr);
  }
  OB_HIP(hipGetLastError());
}

int push_grid_blocks(int64_t n) {
  int64_t b = (n + kPushThreads - 1) / kPushThreads;
  if (b < 1) b = 1;
  if (b > kMaxPushBlocks) b = kMaxPushBlocks;
  return (int)b;
}

int launch_push(const DeviceBunchView& v, const OpProgram& prog, double* ext_partials,
                hipStream_t s) {
  const int blocks = push_grid_blocks(v.n);
  const int level = program_level(prog);
  // levels 0 and 1 take a leading 2.5D kick out of the interpreter
  const bool kick0 = level < 2 && prog.nops > 0 && prog.ops[0].kind == OP_SC2P5DRB_KICK;
  auto go = [&](auto kernel, double* part) {
    hipLaunchKernelGGL(kernel, dim3(blocks), dim3(kPushThreads), 0, s, v, prog, part);
  };
  if (prog.want_extrema) {
    OB_REQUIRE(ext_partials != nullptr, "push: extrema requested without a buffer");
    ProfScope ps("push_extrema");
    if (level == 2)
      go(push_kernel<true, 2, false>, ext_partials);
    else if (level == 1)
      kick0 ? go(push_kernel<true, 1, true>, ext_partials)
            : go(push_kernel<true, 1, false>, ext_partials);
    else
      kick0 ? go(push_kernel<true, 0, true>, ext_partials)
            : go(push_kernel<true, 0, false>, ext_partials);
  } else {
    if (prog.nops == 0 || v.n == 0) return 0;
    ProfScope ps("push");
    if (level == 2)
      go(push_kernel<false, 2, false>, nullptr);
    else if (level == 1)
      kick0 ? go(push_kernel<false, 1, true>, nullptr) : go(push_kernel<false, 1, false>, nullptr);
    else
      kick0 ? go(push_kernel<false, 0, true>, nullptr) : go(push_kernel<false, 0, false>, nullptr);
  }
  OB_HIP(hipGetLastError());
  return blocks;
}

}  // namespace ob
