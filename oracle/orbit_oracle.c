/*
 * orbit_oracle.c -- CPU restatement of the py-orbit TEAPOT + space-charge path.
 *
 * TEST INFRASTRUCTURE ONLY (correctness oracle and timed CPU baseline).  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it; the
 * product path never does.
 *
 * Each routine restates the cited reference routine: same formulas, same
 * evaluation order, same quirks.  Compiled with -O3 -ffp-contract=off so that
 * no multiply-add is fused (the reference's x86-64 g++ build emits none).
 *
 * Pinning (what checks this file independently of itself):
 *   Grid1D/2D/3D binning, index/fraction, interpolation and gradients: bit-exact
 *     against the reference's own Grid*.cc compiled into oracle/_ref (tests/
 *     test_oracle_ref.py and the fixtures in tests/golden).
 *   FFT: numpy.fft (pocketfft) and brute-force DFT.
 *   Poisson 2D/3D: brute-force direct convolution with the free-space Green
 *     function on small grids (exact for Hockney zero padding).
 *   calculateLongDerivative: exactness on quadratic densities and numpy.polyfit.
 *   SC kicks: analytic fields of a uniform round beam (2.5D) and a uniform
 *     sphere (3D); RB longitudinal kick sign/scale against the analytic g-factor.
 *   TEAPOT maps: textbook thick-lens matrix (quad1), exact Re/Im z^n/n! (multp),
 *     drift composition, symplecticity of every map's Jacobian.
 *   Extrema and emittance: numpy.
 * The reference's TEAPOT maps and Poisson solvers cannot be compiled here
 * (Bunch needs Python-2 C-API code; PoissonSolverFFT*.cc need fftw3.h), so no
 * function of this file is compared with them directly: see the tests above.
 */
#include "orbit_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int g_threads = 1;
void orc_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
int orc_get_threads(void) { return g_threads; }

/* ------------------------------------------------------------------------ */
/* SyncPart: wrap_syncpart.cc:70-92 (kinEnergy setter), SyncPart.cc:133-138
 * (setMomentum), SyncPart.cc:236-317 (updateKinematics, energyToMomentum). */
void orc_sync_set_kin_energy(orc_sync* s, double ek) {
  double m = s->mass;
  double p = sqrt(ek * (ek + 2.0 * m)); /* energyToMomentum */
  double pz = p;
  double p2 = 0.0 * 0.0 + 0.0 * 0.0 + pz * pz;
  double p_abs = sqrt(p2);
  double m2 = m * m;
  double w2 = m2 + p2;
  double w = sqrt(w2);
  s->T = p2 / (sqrt(m2 + p2) + m);
  s->p = p_abs;
  s->beta = p_abs / w;
  s->gamma = w / m;
}

/* ------------------------------------------------------------------------ */
/* teapot_base maps: src/teapot/teapotbase.cc */

static double g_factorial[50];
static void init_factorial(void) { /* teapotbase.cc:45-57 */
  if (g_factorial[0] == 1.0) return;
  g_factorial[0] = 1.0;
  for (int i = 1; i < 50; i++) g_factorial[i] = i * g_factorial[i - 1];
}

/* teapotbase.cc:80-101 */
void orc_rotatexy(orc_bunch* b, double anglexy) {
  double cs = cos(anglexy), sn = sin(anglexy);
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1)
  for (int64_t i = 0; i < b->n; i++) {
    double xt = b->x[i], pxt = b->xp[i], yt = b->y[i], pyt = b->yp[i];
    b->x[i] = cs * xt - sn * yt;
    b->xp[i] = cs * pxt - sn * pyt;
    b->y[i] = sn * xt + cs * yt;
    b->yp[i] = sn * pxt + cs * pyt;
  }
}

/* teapotbase.cc:158-183 */
void orc_drift(orc_bunch* b, double length) {
  orc_sync* s = &b->s;
  double v = 2.99792458e+8 * s->beta;
  s->time = s->time + length / v;
  double gamma2i = 1.0 / (s->gamma * s->gamma);
  double dp_p_coeff = 1.0 / (s->p * s->beta);
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1)
  for (int64_t i = 0; i < b->n; i++) {
    double dp_p = b->dE[i] * dp_p_coeff;
    double KNL = 1.0 / (1.0 + dp_p);
    b->x[i] += KNL * length * b->xp[i];
    b->y[i] += KNL * length * b->yp[i];
    double phifac = (b->xp[i] * b->xp[i] + b->yp[i] * b->yp[i] + dp_p * dp_p * gamma2i) / 2.0;
    phifac = (phifac * KNL - dp_p * gamma2i) * KNL;
    b->z[i] -= length * phifac;
  }
}

/* teapotbase.cc:202-211 */
void orc_wrapbunch(orc_bunch* b, double length) {
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1)
  for (int64_t i = 0; i < b->n; i++) {
    if (b->z[i] < -length / 2.0) b->z[i] += length;
    if (b->z[i] > length / 2.0) b->z[i] -= length;
  }
}

/* teapotbase.cc:232-257 */
void orc_kick(orc_bunch* b, double kx, double ky, double kE) {
  if (kx != 0.) for (int64_t i = 0; i < b->n; i++) b->xp[i] += kx;
  if (ky != 0.) for (int64_t i = 0; i < b->n; i++) b->yp[i] += ky;
  if (kE != 0.) for (int64_t i = 0; i < b->n; i++) b->dE[i] += kE;
}

/* std::complex<double> product as g++ evaluates it for finite operands. */
typedef struct { double re, im; } cplx;
static inline cplx cmul(cplx a, cplx b) {
  cplx r;
  r.re = a.re * b.re - a.im * b.im;
  r.im = a.re * b.im + a.im * b.re;
  return r;
}
static inline cplx cadd(cplx a, cplx b) { cplx r = {a.re + b.re, a.im + b.im}; return r; }
static inline cplx csub(cplx a, cplx b) { cplx r = {a.re - b.re, a.im - b.im}; return r; }
static inline cplx cneg(cplx a) { cplx r = {-a.re, -a.im}; return r; }
static inline cplx cmk(double re, double im) { cplx r = {re, im}; return r; }

/* teapotbase.cc:336-372 */
void orc_multp(orc_bunch* b, int pole, double kl, int skew) {
  if (b->s.charge == 0.) return;
  init_factorial();
  double kl1 = kl / g_factorial[pole];
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1)
  for (int64_t i = 0; i < b->n; i++) {
    cplx z = cmk(b->x[i], b->y[i]);
    cplx zn = cmk(1.0, 0.0);
    for (int k = 0; k < pole; k++) zn = cmul(zn, z);
    if (skew) {
      b->xp[i] += kl1 * zn.im;
      b->yp[i] += kl1 * zn.re;
    } else {
      b->xp[i] -= kl1 * zn.re;
      b->yp[i] += kl1 * zn.im;
    }
  }
}

/* teapotbase.cc:395-490 (sign=-1) and 508-603 (sign=+1): hard-edge multipole
 * fringe.  The two reference routines differ only in the sign of kterm in the
 * f/df terms (-kterm*... IN, kterm*... OUT). */
static void multpfringe(orc_bunch* b, int pole, double kl, int skew, int out) {
  if (b->s.charge == 0.) return;
  init_factorial();
  cplx rootm1 = cmk(0.0, 1.0);
  double dp_p_coeff = 1.0 / (b->s.p * b->s.beta);
  int lm1 = pole, l = pole + 1, lp1 = pole + 2, lp2 = pole + 3;
  cplx cxl = cmk(l, 0.0), cxlp2 = cmk(lp2, 0.0);
  double klfactlp1 = kl / (4.0 * g_factorial[lp1]);
  cplx kterm = skew ? cmk(0.0, klfactlp1) : cmk(klfactlp1, 0.0);
  cplx kk = out ? kterm : cneg(kterm);
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1)
  for (int64_t i = 0; i < b->n; i++) {
    double x = b->x[i], y = b->y[i];
    cplx z = cmk(x, y);
    double px = b->xp[i], py = b->yp[i];
    double dp_p = b->dE[i] * dp_p_coeff;
    double KNL = 1.0 / (1.0 + dp_p);
    cplx zlm1 = cmk(1., 0.);
    for (int k = 0; k < lm1; k++) zlm1 = cmul(zlm1, z);
    cplx zl = cmul(zlm1, z);
    cplx fxterm = cmk(l * x, -lp2 * y);
    cplx dxfxterm = cmul(cxl, cadd(fxterm, z));
    cplx dyfxterm = cmul(rootm1, csub(cmul(cxl, fxterm), cmul(cxlp2, z)));
    cplx fyterm = cmk(l * y, lp2 * x);
    cplx dxfyterm = cadd(cmul(cxl, fyterm), cmul(cmul(rootm1, cxlp2), z));
    cplx dyfyterm = cmul(cxl, cadd(cmul(rootm1, fyterm), z));
    cplx fxcx = cmul(cmul(kk, zl), fxterm);
    cplx dxfxcx = cmul(cmul(kk, zlm1), dxfxterm);
    cplx dyfxcx = cmul(cmul(kk, zlm1), dyfxterm);
    cplx fycx = cmul(cmul(kk, zl), fyterm);
    cplx dxfycx = cmul(cmul(kk, zlm1), dxfyterm);
    cplx dyfycx = cmul(cmul(kk, zlm1), dyfyterm);
    b->x[i] -= fxcx.re * KNL;
    b->y[i] -= fycx.re * KNL;
    double M11 = 1.0 - dxfxcx.re * KNL;
    double M12 = -dxfycx.re * KNL;
    double M21 = -dyfxcx.re * KNL;
    double M22 = 1.0 - dyfycx.re * KNL;
    double detM = M11 * M22 - M12 * M21;
    double pxnew = (M22 * px - M12 * py) / detM;
    double pynew = (-M21 * px + M11 * py) / detM;
    b->xp[i] = pxnew;
    b->yp[i] = pynew;
    b->z[i] -= (pxnew * fxcx.re + pynew * fycx.re) * KNL * KNL;
  }
}
void orc_multpfringeIN(orc_bunch* b, int pole, double kl, int skew) { multpfringe(b, pole, kl, skew, 0); }
void orc_multpfringeOUT(orc_bunch* b, int pole, double kl, int skew) { multpfringe(b, pole, kl, skew, 1); }

/* teapotbase.cc:620-700 */
void orc_quad1(orc_bunch* b, double length, double kq) {
  if (kq == 0. || b->s.charge == 0.) {
    orc_drift(b, length);
    return;
  }
  orc_sync* s = &b->s;
  double v = 2.99792458e+8 * s->beta;
  if (length > 0.) s->time = s->time + length / v;
  double gamma2i = 1.0 / (s->gamma * s->gamma);
  double dp_p_coeff = 1.0 / (s->p * s->beta);
  double sqrt_kq, kqlength, cx, sx, cy, sy;
  double m11 = 0., m12 = 0., m21 = 0., m22 = 0., m33 = 0., m34 = 0., m43 = 0., m44 = 0.;
  if (kq > 0.) {
    sqrt_kq = pow(kq, 0.5);
    kqlength = sqrt_kq * length;
    cx = cos(kqlength); sx = sin(kqlength); cy = cosh(kqlength); sy = sinh(kqlength);
    m11 = cx; m12 = sx / sqrt_kq; m21 = -sx * sqrt_kq; m22 = cx;
    m33 = cy; m34 = sy / sqrt_kq; m43 = sy * sqrt_kq; m44 = cy;
  } else if (kq < 0.) {
    sqrt_kq = pow(-kq, 0.5);
    kqlength = sqrt_kq * length;
    cx = cosh(kqlength); sx = sinh(kqlength); cy = cos(kqlength); sy = sin(kqlength);
    m11 = cx; m12 = sx / sqrt_kq; m21 = sx * sqrt_kq; m22 = cx;
    m33 = cy; m34 = sy / sqrt_kq; m43 = -sy * sqrt_kq; m44 = cy;
  }
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1)
  for (int64_t i = 0; i < b->n; i++) {
    double dp_p = b->dE[i] * dp_p_coeff;
    double x0 = b->x[i], xp0 = b->xp[i], y0 = b->y[i], yp0 = b->yp[i];
    b->x[i] = x0 * m11 + xp0 * m12;
    b->xp[i] = x0 * m21 + xp0 * m22;
    b->y[i] = y0 * m33 + yp0 * m34;
    b->yp[i] = y0 * m43 + yp0 * m44;
    b->z[i] += dp_p * gamma2i * length;
  }
}

/* teapotbase.cc:719-745 */
void orc_quad2(orc_bunch* b, double length) {
  double gamma2i = 1.0 / (b->s.gamma * b->s.gamma);
  double dp_p_coeff = 1.0 / (b->s.p * b->s.beta);
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1)
  for (int64_t i = 0; i < b->n; i++) {
    double dp_p = b->dE[i] * dp_p_coeff;
    double KNL = 1.0 / (1.0 + dp_p);
    b->x[i] -= KNL * length * dp_p * b->xp[i];
    b->y[i] -= KNL * length * dp_p * b->yp[i];
    double phifac = (b->xp[i] * b->xp[i] + b->yp[i] * b->yp[i] + dp_p * dp_p * gamma2i) / 2.0;
    phifac = (phifac * KNL + dp_p * dp_p * gamma2i) * KNL;
    b->z[i] -= length * phifac;
  }
}

/* teapotbase.cc:784-835 (IN, sgn=+1) and 854-905 (OUT, sgn=-1) */
static void quadfringe(orc_bunch* b, double kq, int out) {
  if (b->s.charge == 0.) return;
  double dp_p_coeff = 1.0 / (b->s.p * b->s.beta);
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1)
  for (int64_t i = 0; i < b->n; i++) {
    double dp_p = b->dE[i] * dp_p_coeff;
    double KNL = 1.0 / (1.0 + dp_p);
    double x0 = b->x[i], xp0 = b->xp[i], y0 = b->y[i], yp0 = b->yp[i];
    double t = (kq * KNL / 4.) * (x0 * x0 - y0 * y0);
    double detM = 1.0 - t * t; /* pow(t, 2) */
    double dx = (kq * KNL / 12.) * x0 * (x0 * x0 + 3. * y0 * y0);
    double dxp = (kq * KNL / 4.) * (xp0 * (x0 * x0 + y0 * y0) - 2. * yp0 * x0 * y0);
    double dy = (kq * KNL / 12.) * y0 * (y0 * y0 + 3. * x0 * x0);
    double dyp = (kq * KNL / 4.) * (-yp0 * (x0 * x0 + y0 * y0) + 2. * xp0 * x0 * y0);
    double dz = (kq * KNL * KNL / 12.) *
                (xp0 * x0 * (x0 * x0 + 3. * y0 * y0) - yp0 * y0 * (y0 * y0 + 3. * x0 * x0));
    if (!out) {
      b->x[i] += dx;
      b->xp[i] -= dxp; b->xp[i] /= detM;
      b->y[i] -= dy;
      b->yp[i] -= dyp; b->yp[i] /= detM;
      b->z[i] += dz;
    } else {
      b->x[i] -= dx;
      b->xp[i] += dxp; b->xp[i] /= detM;
      b->y[i] += dy;
      b->yp[i] += dyp; b->yp[i] /= detM;
      b->z[i] -= dz;
    }
  }
}
void orc_quadfringeIN(orc_bunch* b, double kq) { quadfringe(b, kq, 0); }
void orc_quadfringeOUT(orc_bunch* b, double kq) { quadfringe(b, kq, 1); }

/* ---- bends, wedges, solenoid, ring RF: teapotbase.cc:923-1569 ---- */
/* One particle's drift, teapotbase.cc:120-141 (drifti). */
static void drift_one(orc_bunch* b, int64_t i, double length, double gamma2i, double dp_p_coeff) {
  double dp_p = b->dE[i] * dp_p_coeff;
  double KNL = 1.0 / (1.0 + dp_p);
  b->x[i] += KNL * length * b->xp[i];
  b->y[i] += KNL * length * b->yp[i];
  double phifac = (b->xp[i] * b->xp[i] + b->yp[i] * b->yp[i] + dp_p * dp_p * gamma2i) / 2.0;
  phifac = (phifac * KNL - dp_p * gamma2i) * KNL;
  b->z[i] -= length * phifac;
}
/* One particle's multipole kick, teapotbase.cc:280-315 (multpi). */
static void multp_one(orc_bunch* b, int64_t i, int pole, double kl, int skew) {
  if (b->s.charge == 0.) return;
  double kl1 = kl / g_factorial[pole];
  cplx z = {b->x[i], b->y[i]}, zn = {1.0, 0.0};
  for (int k = 0; k < pole; k++) zn = cmul(zn, z);
  if (skew) {
    b->xp[i] += kl1 * zn.im;
    b->yp[i] += kl1 * zn.re;
  } else {
    b->xp[i] -= kl1 * zn.re;
    b->yp[i] += kl1 * zn.im;
  }
}
/* :923-966 */
void orc_wedgerotate(orc_bunch* b, double e, int frinout) {
  double dp_p_coeff = 1.0 / (b->s.p * b->s.beta);
  double cs = cos(e), sn = sin(e);
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1)
  for (int64_t i = 0; i < b->n; i++) {
    double dp_p, xp_temp, p0_temp, p0;
    if (frinout == 0) {
      dp_p = b->dE[i] * dp_p_coeff;
      xp_temp = b->xp[i];
      p0_temp = 1.0 + dp_p;
      b->x[i] /= cs;
      b->xp[i] = xp_temp * cs + p0_temp * sn;
      p0 = -xp_temp * sn + p0_temp * cs;
      dp_p = p0 - 1.0;
      b->z[i] = (-b->x[i] * sn + b->z[i]) * cs;
    } else {
      dp_p = b->dE[i] * dp_p_coeff;
      p0 = 1.0 + dp_p;
      b->z[i] = b->x[i] * sn + b->z[i] / cs;
      b->x[i] *= cs;
      xp_temp = b->xp[i] * cs - p0 * sn;
      p0_temp = b->xp[i] * sn + p0 * cs;
      b->xp[i] = xp_temp;
      dp_p = p0_temp - 1.0;
    }
    b->dE[i] = dp_p / dp_p_coeff;
  }
}
/* :986-1012 */
void orc_wedgedrift(orc_bunch* b, double e, int inout) {
  double gamma2i = 1.0 / (b->s.gamma * b->s.gamma);
  double dp_p_coeff = 1.0 / (b->s.p * b->s.beta);
  double ct = cos(e) / sin(e);
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1)
  for (int64_t i = 0; i < b->n; i++) {
    double s;
    if (inout == 0) {
      double dp_p = b->dE[i] * dp_p_coeff;
      double tn = b->xp[i] / (1.0 + dp_p);
      s = b->x[i] / (ct - tn);
    } else {
      s = b->x[i] / ct;
    }
    drift_one(b, i, s, gamma2i, dp_p_coeff);
  }
}
/* :1447-1509 (wedgebendCF); wedgebend (:1037-1080) is the case with no multipoles. */
void orc_wedgebendCF(orc_bunch* b, double e, int inout, double rho, int vecnum, const int* pole,
                     const double* kl, const int* skew, int nsteps) {
  init_factorial();
  double gamma2i = 1.0 / (b->s.gamma * b->s.gamma);
  double dp_p_coeff = 1.0 / (b->s.p * b->s.beta);
  int nst = nsteps / 2;
  if (nst < 1) nst = 1;
  double ct = cos(e) / sin(e);
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1)
  for (int64_t i = 0; i < b->n; i++) {
    double s;
    if (inout == 0) {
      s = -b->x[i] / ct;
    } else {
      double dp_p = b->dE[i] * dp_p_coeff;
      double tn = b->xp[i] / (1.0 + dp_p);
      s = -b->x[i] / (ct + tn);
    }
    double sm = s / nst, sm2 = sm / 2.0;
    drift_one(b, i, sm2, gamma2i, dp_p_coeff);
    b->xp[i] -= sm / rho;
    for (int l = 0; l < vecnum; l++) multp_one(b, i, pole[l], kl[l] * sm, skew[l]);
    for (int j = 1; j < nst; j++) {
      drift_one(b, i, sm, gamma2i, dp_p_coeff);
      b->xp[i] -= sm / rho;
      for (int l = 0; l < vecnum; l++) multp_one(b, i, pole[l], kl[l] * sm, skew[l]);
    }
    drift_one(b, i, sm2, gamma2i, dp_p_coeff);
  }
}
void orc_wedgebend(orc_bunch* b, double e, int inout, double rho, int nsteps) {
  orc_wedgebendCF(b, e, inout, rho, 0, NULL, NULL, NULL, nsteps);
}
/* :1099-1145 */
void orc_bend1(orc_bunch* b, double length, double th) {
  orc_sync* s = &b->s;
  double v = 2.99792458e+8 * s->beta;
  if (length > 0.) s->time = s->time + length / v;
  double betasq = s->beta * s->beta;
  double dp_p_coeff = 1.0 / (s->p * s->beta);
  double rho = length / th, cx = cos(th), sx = sin(th);
  double m11 = cx, m12 = rho * sx, m16 = rho * (1.0 - cx);
  double m21 = -sx / rho, m22 = cx, m26 = sx;
  double m51 = -sx, m52 = -rho * (1.0 - cx), m56 = -betasq * length + rho * sx;
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1)
  for (int64_t i = 0; i < b->n; i++) {
    double dp_p = b->dE[i] * dp_p_coeff;
    double x0 = b->x[i], xp0 = b->xp[i];
    b->x[i] = x0 * m11 + xp0 * m12 + dp_p * m16;
    b->xp[i] = x0 * m21 + xp0 * m22 + dp_p * m26;
    b->y[i] += length * b->yp[i];
    b->z[i] += x0 * m51 + xp0 * m52 + dp_p * m56;
  }
}
/* :1164-1190 (same arithmetic as quad2) */
void orc_bend2(orc_bunch* b, double length) { orc_quad2(b, length); }
/* :1208-1233 */
void orc_bend3(orc_bunch* b, double th) {
  double gamma2i = 1.0 / (b->s.gamma * b->s.gamma);
  double dp_p_coeff = 1.0 / (b->s.p * b->s.beta);
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1)
  for (int64_t i = 0; i < b->n; i++) {
    double dp_p = b->dE[i] * dp_p_coeff;
    double KNL = 1.0 / (1.0 + dp_p);
    double phifac = (b->yp[i] * b->yp[i] + dp_p * dp_p * gamma2i) / 2.0;
    b->xp[i] -= phifac * KNL * th;
    b->y[i] += KNL * b->yp[i] * b->x[i] * th;
    phifac = (phifac * KNL - dp_p * gamma2i) * KNL;
    b->z[i] -= th * phifac * b->x[i];
  }
}
/* :1251-1274 */
void orc_bend4(orc_bunch* b, double th) {
  double dp_p_coeff = 1.0 / (b->s.p * b->s.beta);
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1)
  for (int64_t i = 0; i < b->n; i++) {
    double dp_p = b->dE[i] * dp_p_coeff;
    double KNL = 1.0 / (1.0 + dp_p);
    double xfac = 1.0 + KNL * b->xp[i] * th / 2.0;
    double phifac = KNL * KNL * b->xp[i] * b->xp[i] / 2.0;
    b->x[i] *= xfac * xfac;
    b->xp[i] /= xfac;
    b->z[i] -= th * phifac * b->x[i];
  }
}
/* :1292-1311 (IN) and :1329-1348 (OUT) */
static void bendfringe(orc_bunch* b, double rho, int out) {
  double dp_p_coeff = 1.0 / (b->s.p * b->s.beta);
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1)
  for (int64_t i = 0; i < b->n; i++) {
    double dp_p = b->dE[i] * dp_p_coeff;
    double KNL = 1.0 / (1.0 + dp_p);
    if (!out) {
      b->x[i] += KNL * b->y[i] * b->y[i] / (2. * rho);
      b->yp[i] -= KNL * b->xp[i] * b->y[i] / rho;
      b->z[i] += KNL * KNL * b->xp[i] * b->y[i] * b->y[i] / (2. * rho);
    } else {
      b->x[i] -= KNL * b->y[i] * b->y[i] / (2. * rho);
      b->yp[i] += KNL * b->xp[i] * b->y[i] / rho;
      b->z[i] -= KNL * KNL * b->xp[i] * b->y[i] * b->y[i] / (2. * rho);
    }
  }
}
void orc_bendfringeIN(orc_bunch* b, double rho) { bendfringe(b, rho, 0); }
void orc_bendfringeOUT(orc_bunch* b, double rho) { bendfringe(b, rho, 1); }
/* :1367-1425 */
void orc_soln(orc_bunch* b, double length, double B) {
  if (fabs(B) < 1.0e-100 || b->s.charge == 0.) {
    orc_drift(b, length);
    return;
  }
  orc_sync* s = &b->s;
  double Bc = B * s->charge;
  double v = 2.99792458e+8 * s->beta;
  if (length > 0.) s->time = s->time + length / v;
  double gamma2i = 1.0 / (s->gamma * s->gamma);
  double dp_p_coeff = 1.0 / (s->p * s->beta);
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1)
  for (int64_t i = 0; i < b->n; i++) {
    double dp_p = b->dE[i] * dp_p_coeff;
    double KNL = 1.0 / (1.0 + dp_p);
    double cu = b->y[i] / 2. - b->xp[i] / Bc;
    double cpu = b->x[i] * Bc / 2. + b->yp[i];
    double u_init = b->y[i] / 2. + b->xp[i] / Bc;
    double pu_init = -b->x[i] * Bc / 2. + b->yp[i];
    double phase = KNL * Bc * length;
    double cs = cos(phase), sn = sin(phase);
    double u = u_init * cs + pu_init * sn / Bc;
    double pu = -u_init * Bc * sn + pu_init * cs;
    b->x[i] = (-pu + cpu) / Bc;
    b->xp[i] = 0.5 * (u - cu) * Bc;
    b->y[i] = u + cu;
    b->yp[i] = 0.5 * (pu + cpu);
    double phifac = (pu_init * pu_init + Bc * Bc * u_init * u_init + dp_p * dp_p * gamma2i) / 2.0;
    phifac = (phifac * KNL - dp_p * gamma2i) * KNL;
    b->z[i] -= length * phifac;
  }
}
/* :1527-1566 */
void orc_ringrf(orc_bunch* b, double ring_length, int harmonic_numb, double voltage,
                double phase_s, int useCharge) {
  double charge = +1.0;
  if (useCharge == 1) charge = b->s.charge;
  double coeff = charge;
  double Factor = 0.;
  if (ring_length > 0.) Factor = 2.0 * 3.14159265358979323846 / ring_length;
  orc_sync* s = &b->s;
  double p_synch_in = s->p;
  if (phase_s != 0.) {
    double kin_e = s->T;
    kin_e += coeff * voltage * sin(phase_s);
    double m = s->mass; /* setMomentum(energyToMomentum(kin_e)), SyncPart.cc:133-138, 314-317 */
    double pz = sqrt(kin_e * (kin_e + 2.0 * m));
    double p2 = 0.0 * 0.0 + 0.0 * 0.0 + pz * pz;
    double p_abs = sqrt(p2), m2 = m * m, w2 = m2 + p2, w = sqrt(w2);
    s->T = p2 / (sqrt(m2 + p2) + m);
    s->p = p_abs;
    s->beta = p_abs / w;
    s->gamma = w / m;
  }
  double xp_yp_coeff = p_synch_in / s->p;
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1)
  for (int64_t i = 0; i < b->n; i++) {
    double deltaV = voltage * (sin(harmonic_numb * Factor * b->z[i] + phase_s));
    b->dE[i] += coeff * deltaV;
    b->xp[i] *= xp_yp_coeff;
    b->yp[i] *= xp_yp_coeff;
  }
}

/* ------------------------------------------------------------------------ */
/* BunchExtremaCalculator.cc:37-83 */
void orc_extrema_xyz(const orc_bunch* b, double o[6]) {
  double l[6] = {DBL_MAX, -DBL_MAX, DBL_MAX, -DBL_MAX, DBL_MAX, -DBL_MAX};
  for (int64_t i = 0; i < b->n; i++) {
    if (b->x[i] < l[0]) l[0] = b->x[i];
    if (b->x[i] > l[1]) l[1] = b->x[i];
    if (b->y[i] < l[2]) l[2] = b->y[i];
    if (b->y[i] > l[3]) l[3] = b->y[i];
    if (b->z[i] < l[4]) l[4] = b->z[i];
    if (b->z[i] > l[5]) l[5] = b->z[i];
  }
  memcpy(o, l, sizeof(l));
}

/* getExtremaXpYpdE (BunchExtremaCalculator.cc:85-120): the same scan over xp, yp, dE.
   o[12] = (min, max) of x, xp, y, yp, z, dE. */
void orc_extrema6(const orc_bunch* b, double o[12]) {
  double xyz[6];
  orc_extrema_xyz(b, xyz);
  double l[6] = {DBL_MAX, -DBL_MAX, DBL_MAX, -DBL_MAX, DBL_MAX, -DBL_MAX};
  for (int64_t i = 0; i < b->n; i++) {
    if (b->xp[i] < l[0]) l[0] = b->xp[i];
    if (b->xp[i] > l[1]) l[1] = b->xp[i];
    if (b->yp[i] < l[2]) l[2] = b->yp[i];
    if (b->yp[i] > l[3]) l[3] = b->yp[i];
    if (b->dE[i] < l[4]) l[4] = b->dE[i];
    if (b->dE[i] > l[5]) l[5] = b->dE[i];
  }
  o[0] = xyz[0]; o[1] = xyz[1]; o[2] = l[0]; o[3] = l[1];
  o[4] = xyz[2]; o[5] = xyz[3]; o[6] = l[2]; o[7] = l[3];
  o[8] = xyz[4]; o[9] = xyz[5]; o[10] = l[4]; o[11] = l[5];
}

/* getExtremaR (BunchExtremaCalculator.cc:167-208): unweighted mean of x and y, then the
   largest sqrt(pow(x - <x>, 2) + pow(y - <y>, 2)). */
double orc_extrema_r(const orc_bunch* b) {
  double xa = 0., ya = 0.;
  for (int64_t i = 0; i < b->n; i++) {
    xa += b->x[i];
    ya += b->y[i];
  }
  if (b->n > 0) {
    xa = xa / (double)b->n;
    ya = ya / (double)b->n;
  }
  double rl = 0.;
  for (int64_t i = 0; i < b->n; i++) {
    const double r = sqrt(pow(b->x[i] - xa, 2.0) + pow(b->y[i] - ya, 2.0));
    if (r > rl) rl = r;
  }
  return rl;
}

/* BunchTwissAnalysis::computeBunchMoments (BunchTwissAnalysis.cc:145-337), after the
   caller's analyzeBunch: xAvg = getAverage(0) (the dispersion branch at :166 is dead, its
   test reads a local that is still 0), yAvg = getAverage(2), D = getDispersion(0),
   Em = E_kin + mass, b2 = beta^2, xb and yb the normalising terms (:238-245).
   mom[(order+1)^2], row-major [i][j]. */
void orc_compute_bunch_moments(const orc_bunch* b, int order, int dispflag, double xAvg,
                               double yAvg, double D, double Em, double b2, double xb,
                               double yb, double* mom) {
  const int o1 = order + 1;
  double momX[64], momY[64];
  for (int k = 0; k < o1 * o1; k++) mom[k] = 0.;
  for (int n = 0; n < o1; n++) momX[n] = momY[n] = 0.;
  momX[0] = 1.0;
  momY[0] = 1.0;
  double total = 0.;
  double dispterm = 0.;
  int i = 0, j = 0;
  for (int64_t ip = 0; ip < b->n; ip++) {
    const double m = b->msize ? b->msize[ip] : 1.0;
    if (b->msize) total += m;
    if (dispflag > 0) dispterm = D * b->dE[ip] / Em / b2;
    for (i = 0; i < order; i++)
      momX[i + 1] = b->msize ? momX[i] * m * ((b->x[ip] - dispterm) - xAvg)
                             : momX[i] * ((b->x[ip] - dispterm) - xAvg);
    for (i = 0; i < order; i++)
      momY[i + 1] = b->msize ? momY[i] * m * (b->y[ip] - yAvg) : momY[i] * (b->y[ip] - yAvg);
    for (j = 0; j < order; j++)
      for (i = 0; i < o1 - j; i++) {
        mom[i * o1 + j] += momX[i] / pow(xb, (double)i) * momY[j] / pow(yb, (double)j);
        if (b->msize) mom[i * o1 + j] += momX[i] * momY[j];
      }
    /* without macrosizes the second statement sits after the loops (:276-280): i = 2,
       j = order; for order 1 the reference reads and writes outside its arrays */
    if (!b->msize && order >= 2) mom[2 * o1 + order] += momX[2] * momY[order];
  }
  if (!b->msize) total += (double)b->n * 1.0;
  if (fabs(total) > 0.)
    for (i = 0; i < o1; i++)
      for (j = 0; j < o1 - i; j++) mom[i * o1 + j] /= total;
  mom[0] = 1.;
  mom[1 * o1 + 0] = xAvg;
  mom[0 * o1 + 1] = yAvg;
}

/* ------------------------------------------------------------------------ */
/* Grid2D: src/spacecharge/Grid2D.cc */
void orc_grid2d_set_x(orc_grid2d* g, double a, double b) { /* :409-413 */
  g->xMin = a; g->xMax = b; g->dx = (g->xMax - g->xMin) / (g->nx - 1);
}
void orc_grid2d_set_y(orc_grid2d* g, double a, double b) { /* :416-420 */
  g->yMin = a; g->yMax = b; g->dy = (g->yMax - g->yMin) / (g->ny - 1);
}
void orc_grid2d_ind_frac_x(const orc_grid2d* g, double x, int* ind, double* frac) { /* :337-342 */
  int i = (int)((x - g->xMin) / g->dx + 0.5);
  if (i < 1) i = 1;
  if (i > (g->nx - 2)) i = g->nx - 2;
  *ind = i;
  *frac = (x - (g->xMin + i * g->dx)) / g->dx;
}
void orc_grid2d_ind_frac_y(const orc_grid2d* g, double y, int* ind, double* frac) { /* :344-349 */
  int i = (int)((y - g->yMin) / g->dy + 0.5);
  if (i < 1) i = 1;
  if (i > (g->ny - 2)) i = g->ny - 2;
  *ind = i;
  *frac = (y - (g->yMin + i * g->dy)) / g->dy;
}
/* :142-166 */
void orc_grid2d_bin_value(orc_grid2d* g, double value, double x, double y) {
  if (x < g->xMin || x > g->xMax || y < g->yMin || y > g->yMax) return;
  int iX, iY;
  double xF, yF;
  orc_grid2d_ind_frac_x(g, x, &iX, &xF);
  orc_grid2d_ind_frac_y(g, y, &iY, &yF);
  double xF2 = xF * xF, yF2 = yF * yF;
  double Wxm = 0.5 * (0.25 - xF + xF2), Wx0 = 0.75 - xF2, Wxp = 0.5 * (0.25 + xF + xF2);
  double Wym = 0.5 * (0.25 - yF + yF2), Wy0 = 0.75 - yF2, Wyp = 0.5 * (0.25 + yF + yF2);
  int ny = g->ny;
  double* a = g->arr;
  a[(iX - 1) * ny + iY - 1] += Wxm * Wym * value;
  a[(iX - 1) * ny + iY] += Wxm * Wy0 * value;
  a[(iX - 1) * ny + iY + 1] += Wxm * Wyp * value;
  a[iX * ny + iY - 1] += Wx0 * Wym * value;
  a[iX * ny + iY] += Wx0 * Wy0 * value;
  a[iX * ny + iY + 1] += Wx0 * Wyp * value;
  a[(iX + 1) * ny + iY - 1] += Wxp * Wym * value;
  a[(iX + 1) * ny + iY] += Wxp * Wy0 * value;
  a[(iX + 1) * ny + iY + 1] += Wxp * Wyp * value;
}
/* :120-139; with threads > 1 each thread bins into a private grid and the
 * grids are summed in thread order (the reference's MPI decomposition,
 * Grid2D.cc:443-477). */
void orc_grid2d_bin_bunch(orc_grid2d* g, const orc_bunch* b) {
  int nt = g_threads;
  if (nt <= 1 || b->n < 4096) {
    for (int64_t i = 0; i < b->n; i++)
      orc_grid2d_bin_value(g, b->msize ? b->msize[i] : b->s.macrosize, b->x[i], b->y[i]);
    return;
  }
  size_t ng = (size_t)g->nx * g->ny;
  double* priv = (double*)calloc(ng * nt, sizeof(double));
#pragma omp parallel num_threads(nt)
  {
#ifdef _OPENMP
    int t = omp_get_thread_num();
#else
    int t = 0;
#endif
    orc_grid2d gg = *g;
    gg.arr = priv + ng * t;
    int64_t lo = b->n * t / nt, hi = b->n * (t + 1) / nt;
    for (int64_t i = lo; i < hi; i++)
      orc_grid2d_bin_value(&gg, b->msize ? b->msize[i] : b->s.macrosize, b->x[i], b->y[i]);
  }
  for (int t = 0; t < nt; t++)
    for (size_t k = 0; k < ng; k++) g->arr[k] += priv[ng * t + k];
  free(priv);
}
/* :212-256 */
void orc_grid2d_gradient(const orc_grid2d* g, double x, double y, double* ex, double* ey) {
  int iX, iY;
  double xF, yF;
  orc_grid2d_ind_frac_x(g, x, &iX, &xF);
  orc_grid2d_ind_frac_y(g, y, &iY, &yF);
  double xF2 = xF * xF, yF2 = yF * yF;
  double Wxm = 0.5 * (0.25 - xF + xF2), Wx0 = 0.75 - xF2, Wxp = 0.5 * (0.25 + xF + xF2);
  double Wym = 0.5 * (0.25 - yF + yF2), Wy0 = 0.75 - yF2, Wyp = 0.5 * (0.25 + yF + yF2);
  double dWxm = (-1.0) * (0.5 - xF), dWx0 = (-1.0) * (+2.) * xF, dWxp = (-1.0) * (-(0.5 + xF));
  double dWym = (-1.0) * (0.5 - yF), dWy0 = (-1.0) * (+2.) * yF, dWyp = (-1.0) * (-(0.5 + yF));
  int ny = g->ny;
  const double* a = g->arr;
  const double* r0 = a + (iX - 1) * ny + iY;
  const double* r1 = a + iX * ny + iY;
  const double* r2 = a + (iX + 1) * ny + iY;
  double e = dWxm * Wym * r0[-1] + dWxm * Wy0 * r0[0] + dWxm * Wyp * r0[1] +
             dWx0 * Wym * r1[-1] + dWx0 * Wy0 * r1[0] + dWx0 * Wyp * r1[1] +
             dWxp * Wym * r2[-1] + dWxp * Wy0 * r2[0] + dWxp * Wyp * r2[1];
  *ex = e / g->dx;
  e = Wxm * dWym * r0[-1] + Wxm * dWy0 * r0[0] + Wxm * dWyp * r0[1] +
      Wx0 * dWym * r1[-1] + Wx0 * dWy0 * r1[0] + Wx0 * dWyp * r1[1] +
      Wxp * dWym * r2[-1] + Wxp * dWy0 * r2[0] + Wxp * dWyp * r2[1];
  *ey = e / g->dy;
}
/* :81-106 */
double orc_grid2d_value(const orc_grid2d* g, double x, double y) {
  int iX, iY;
  double xF, yF;
  orc_grid2d_ind_frac_x(g, x, &iX, &xF);
  orc_grid2d_ind_frac_y(g, y, &iY, &yF);
  double xF2 = xF * xF, yF2 = yF * yF;
  double Wxm = 0.5 * (0.25 - xF + xF2), Wx0 = 0.75 - xF2, Wxp = 0.5 * (0.25 + xF + xF2);
  double Wym = 0.5 * (0.25 - yF + yF2), Wy0 = 0.75 - yF2, Wyp = 0.5 * (0.25 + yF + yF2);
  int ny = g->ny;
  const double* r0 = g->arr + (iX - 1) * ny + iY;
  const double* r1 = g->arr + iX * ny + iY;
  const double* r2 = g->arr + (iX + 1) * ny + iY;
  return Wxm * Wym * r0[-1] + Wxm * Wy0 * r0[0] + Wxm * Wyp * r0[1] + Wx0 * Wym * r1[-1] +
         Wx0 * Wy0 * r1[0] + Wx0 * Wyp * r1[1] + Wxp * Wym * r2[-1] + Wxp * Wy0 * r2[0] +
         Wxp * Wyp * r2[1];
}

/* ------------------------------------------------------------------------ */
/* Grid1D: src/spacecharge/Grid1D.cc */
void orc_grid1d_set_z(orc_grid1d* g, double a, double b) { /* :88-93 */
  g->zMin = a; g->zMax = b; g->dz = (g->zMax - g->zMin) / g->nz;
}
static inline double grid1d_gridz(const orc_grid1d* g, int i) { return g->zMin + (i + 0.5) * g->dz; } /* :132-135 */
/* :568-626 (periodic at both ends, quirk Q2) */
void orc_grid1d_ind_wz(const orc_grid1d* g, double z, int* i0, int* ip, double* w0, double* wp) {
  if (g->nz > 1) {
    double ind_dbl = ((z - g->zMin) / g->dz) - 0.5;
    int iZ0 = (int)ind_dbl;
    double zFrac = (z - grid1d_gridz(g, iZ0)) / g->dz;
    if (iZ0 == 0 && zFrac < 0.) {
      *i0 = g->nz - 1; *ip = 0; *w0 = -zFrac; *wp = 1.0 + zFrac;
      return;
    }
    if (iZ0 == (g->nz - 1)) {
      if (zFrac > 0.) {
        *i0 = iZ0; *ip = 0; *w0 = 1.0 - zFrac; *wp = zFrac;
        return;
      } else {
        *ip = iZ0; *i0 = g->nz - 2; *w0 = -zFrac; *wp = 1.0 + zFrac;
        return;
      }
    }
    *i0 = iZ0; *ip = iZ0 + 1; *w0 = 1.0 - zFrac; *wp = zFrac;
  } else {
    *i0 = 0; *ip = 0; *w0 = 1.0; *wp = 0.0;
  }
}
/* :422-440 */
void orc_grid1d_bin_value(orc_grid1d* g, double value, double z) {
  if (z < g->zMin || z > g->zMax) return;
  int i0, ip;
  double w0, wp;
  orc_grid1d_ind_wz(g, z, &i0, &ip, &w0, &wp);
  if (g->nz > 1) {
    g->arr[i0] += w0 * value;
    g->arr[ip] += wp * value;
  } else if (g->nz == 1) {
    g->arr[i0] += value;
  }
}
/* :241-275 (axis 4 = z) */
void orc_grid1d_bin_bunch(orc_grid1d* g, const orc_bunch* b) {
  int nt = g_threads;
  if (nt <= 1 || b->n < 4096) {
    for (int64_t i = 0; i < b->n; i++)
      orc_grid1d_bin_value(g, b->msize ? b->msize[i] : b->s.macrosize, b->z[i]);
    return;
  }
  double* priv = (double*)calloc((size_t)g->nz * nt, sizeof(double));
#pragma omp parallel num_threads(nt)
  {
#ifdef _OPENMP
    int t = omp_get_thread_num();
#else
    int t = 0;
#endif
    orc_grid1d gg = *g;
    gg.arr = priv + (size_t)g->nz * t;
    int64_t lo = b->n * t / nt, hi = b->n * (t + 1) / nt;
    for (int64_t i = lo; i < hi; i++)
      orc_grid1d_bin_value(&gg, b->msize ? b->msize[i] : b->s.macrosize, b->z[i]);
  }
  for (int t = 0; t < nt; t++)
    for (int k = 0; k < g->nz; k++) g->arr[k] += priv[(size_t)g->nz * t + k];
  free(priv);
}
/* binBunch(bunch, axis) / binBunchByParticle(bunch, axis) (:241-275, 335-352) */
void orc_grid1d_bin_axis(orc_grid1d* g, const orc_bunch* b, int axis, int by_particle) {
  const double* c[6] = {b->x, b->xp, b->y, b->yp, b->z, b->dE};
  for (int64_t i = 0; i < b->n; i++) {
    const double v = by_particle ? 1.0 : (b->msize ? b->msize[i] : b->s.macrosize);
    orc_grid1d_bin_value(g, v, c[axis][i]);
  }
}
/* :191-210 (quirk Q7: zero outside [zMin,zMax]) */
double orc_grid1d_value(const orc_grid1d* g, double z) {
  if (z < g->zMin || z > g->zMax) return 0.;
  int i0, ip;
  double w0, wp;
  orc_grid1d_ind_wz(g, z, &i0, &ip, &w0, &wp);
  if (g->nz > 1) return w0 * g->arr[i0] + wp * g->arr[ip];
  return g->arr[0];
}
/* :476-489 */
void orc_grid1d_gradient(const orc_grid1d* g, double z, double* ez) {
  *ez = 0.;
  if (z < g->zMin || z > g->zMax) return;
  int i0, ip;
  double w0, wp;
  orc_grid1d_ind_wz(g, z, &i0, &ip, &w0, &wp);
  if (g->nz > 1) *ez = (g->arr[ip] - g->arr[i0]) / g->dz;
}

/* ------------------------------------------------------------------------ */
/* Grid3D: src/spacecharge/Grid3D.cc */
void orc_grid3d_set_x(orc_grid3d* g, double a, double b) { g->xMin = a; g->xMax = b; g->dx = (g->xMax - g->xMin) / (g->nx - 1); }
void orc_grid3d_set_y(orc_grid3d* g, double a, double b) { g->yMin = a; g->yMax = b; g->dy = (g->yMax - g->yMin) / (g->ny - 1); }
void orc_grid3d_set_z(orc_grid3d* g, double a, double b) { g->zMin = a; g->zMax = b; g->dz = (g->zMax - g->zMin) / g->nz; }
/* :212-264 */
void orc_grid3d_ind_frac(const orc_grid3d* g, double x, double y, double z, int ind[3], double frac[3]) {
  int i = (int)((x - g->xMin) / g->dx + 0.5);
  if (i < 1) i = 1;
  if (i > (g->nx - 2)) i = g->nx - 2;
  ind[0] = i;
  frac[0] = (x - (g->xMin + i * g->dx)) / g->dx;
  i = (int)((y - g->yMin) / g->dy + 0.5);
  if (i < 1) i = 1;
  if (i > (g->ny - 2)) i = g->ny - 2;
  ind[1] = i;
  frac[1] = (y - (g->yMin + i * g->dy)) / g->dy;
  if (g->nz > 1) {
    int zi = (int)((z - g->zMin) / g->dz);
    double zf = 0.0;
    if (g->nz > 2) {
      if (!g->wrap) {
        if (zi <= 0) zi = 1;
        if (zi >= g->nz - 1) zi = g->nz - 2;
      } else {
        if (zi <= 0) zi = 0;
        if (zi >= g->nz - 1) zi = g->nz - 1;
      }
      zf = (z - (g->zMin + (zi + 0.5) * g->dz)) / g->dz;
    }
    if (g->nz == 2) {
      zf = (z - (g->zMin + 0.5 * g->dz)) / g->dz;
      if (zi < 0) { zi = 0; zf = 0.0; }
      else if (zi > 1) { zi = 1; zf = 1.0; }
    }
    ind[2] = zi;
    frac[2] = zf;
  } else {
    ind[2] = 0;
    frac[2] = 0.0;
  }
}
#define A3(g, iz, ix, iy) (g)->arr[((size_t)(iz) * (g)->nx + (ix)) * (g)->ny + (iy)]
/* :307-432 */
void orc_grid3d_bin_value(orc_grid3d* g, double m, double x, double y, double z) {
  if (x < g->xMin || x > g->xMax || y < g->yMin || y > g->yMax || z < g->zMin || z > g->zMax) return;
  int ind[3];
  double fr[3];
  orc_grid3d_ind_frac(g, x, y, z, ind, fr);
  int iX = ind[0], iY = ind[1], iZ = ind[2];
  double xF = fr[0], yF = fr[1], zF = fr[2];
  double Wzm = 0.0, Wz0 = 0.0, Wzp = 0.0;
  double Wxm = 0.5 * (0.5 - xF) * (0.5 - xF), Wx0 = 0.75 - xF * xF, Wxp = 0.5 * (0.5 + xF) * (0.5 + xF);
  double Wym = 0.5 * (0.5 - yF) * (0.5 - yF), Wy0 = 0.75 - yF * yF, Wyp = 0.5 * (0.5 + yF) * (0.5 + yF);
  int iZm = iZ - 1, iZp = iZ + 1;
  if (iZm < 0) iZm = g->nz - 1;
  if (iZp >= g->nz) iZp = 0;
  if (g->nz >= 3) {
    Wzm = 0.5 * (0.5 - zF) * (0.5 - zF);
    Wz0 = 0.75 - zF * zF;
    Wzp = 0.5 * (0.5 + zF) * (0.5 + zF);
  }
  if (g->nz == 2) { Wzm = 1.0 - zF; Wz0 = 0.0; Wzp = zF; }
  double wy[3] = {Wym, Wy0, Wyp};
  double wx[3] = {Wxm, Wx0, Wxp};
  if (g->nz >= 3) {
    int zz[3] = {iZm, iZ, iZp};
    double wz[3] = {Wzm, Wz0, Wzp};
    for (int a = 0; a < 3; a++)
      for (int c = 0; c < 3; c++) {
        double tmp = wy[c] * wz[a] * m;
        for (int d = 0; d < 3; d++) A3(g, zz[a], iX - 1 + d, iY - 1 + c) += wx[d] * tmp;
      }
  }
  if (g->nz == 2) {
    double wz[2] = {Wzm, Wzp};
    for (int a = 0; a < 2; a++)
      for (int c = 0; c < 3; c++) {
        double tmp = wy[c] * wz[a] * m;
        for (int d = 0; d < 3; d++) A3(g, a, iX - 1 + d, iY - 1 + c) += wx[d] * tmp;
      }
  }
  if (g->nz == 1) {
    for (int c = 0; c < 3; c++) {
      double tmp = wy[c] * m;
      for (int d = 0; d < 3; d++) A3(g, 0, iX - 1 + d, iY - 1 + c) += wx[d] * tmp;
    }
  }
}
/* :281-305: binBunch(bunch) sets longWrapping = 0 (:307-310) */
void orc_grid3d_bin_bunch(orc_grid3d* g, const orc_bunch* b) {
  g->wrap = 0;
  int nt = g_threads;
  if (nt <= 1 || b->n < 4096) {
    for (int64_t i = 0; i < b->n; i++)
      orc_grid3d_bin_value(g, b->msize ? b->msize[i] : b->s.macrosize, b->x[i], b->y[i], b->z[i]);
    return;
  }
  size_t ng = (size_t)g->nx * g->ny * g->nz;
  /* private grids would need nt copies of a possibly large grid: bin slabs of
   * particles per thread into private grids only when that fits in 4 GiB. */
  if (ng * nt * sizeof(double) > ((size_t)4 << 30)) nt = (int)(((size_t)4 << 30) / (ng * sizeof(double)));
  if (nt < 2) {
    for (int64_t i = 0; i < b->n; i++)
      orc_grid3d_bin_value(g, b->msize ? b->msize[i] : b->s.macrosize, b->x[i], b->y[i], b->z[i]);
    return;
  }
  double* priv = (double*)calloc(ng * nt, sizeof(double));
#pragma omp parallel num_threads(nt)
  {
#ifdef _OPENMP
    int t = omp_get_thread_num();
#else
    int t = 0;
#endif
    orc_grid3d gg = *g;
    gg.arr = priv + ng * t;
    int64_t lo = b->n * t / nt, hi = b->n * (t + 1) / nt;
    for (int64_t i = lo; i < hi; i++)
      orc_grid3d_bin_value(&gg, b->msize ? b->msize[i] : b->s.macrosize, b->x[i], b->y[i], b->z[i]);
  }
#pragma omp parallel for num_threads(g_threads)
  for (int64_t k = 0; k < (int64_t)ng; k++)
    for (int t = 0; t < nt; t++) g->arr[k] += priv[ng * t + k];
  free(priv);
}
/* binBunch(bunch, lambda) with longWrapping = 1 (:281-305, serial) */
void orc_grid3d_bin_bunch_wrapped(orc_grid3d* g, const orc_bunch* b, double lambda) {
  g->wrap = 1;
  for (int64_t i = 0; i < b->n; i++) {
    double z = remainder(b->z[i], lambda);
    orc_grid3d_bin_value(g, b->msize ? b->msize[i] : b->s.macrosize, b->x[i], b->y[i], z);
  }
}
/* :718-733 */
static double sheet_grad(const orc_grid3d* g, int iZ, int iX, int iY, double xm, double x0, double xp,
                         double ym, double y0, double yp) {
  return xm * ym * A3(g, iZ, iX - 1, iY - 1) + x0 * ym * A3(g, iZ, iX, iY - 1) +
         xp * ym * A3(g, iZ, iX + 1, iY - 1) + xm * y0 * A3(g, iZ, iX - 1, iY) +
         x0 * y0 * A3(g, iZ, iX, iY) + xp * y0 * A3(g, iZ, iX + 1, iY) +
         xm * yp * A3(g, iZ, iX - 1, iY + 1) + x0 * yp * A3(g, iZ, iX, iY + 1) +
         xp * yp * A3(g, iZ, iX + 1, iY + 1);
}
/* :438-532 */
void orc_grid3d_gradient(const orc_grid3d* g, double x, double y, double z, double* gX, double* gY,
                         double* gZ) {
  int ind[3];
  double fr[3];
  orc_grid3d_ind_frac(g, x, y, z, ind, fr);
  int iX = ind[0], iY = ind[1], iZ = ind[2];
  double xF = fr[0], yF = fr[1], zF = fr[2];
  double Wxm = 0.5 * (0.5 - xF) * (0.5 - xF), Wx0 = 0.75 - xF * xF, Wxp = 0.5 * (0.5 + xF) * (0.5 + xF);
  double Wym = 0.5 * (0.5 - yF) * (0.5 - yF), Wy0 = 0.75 - yF * yF, Wyp = 0.5 * (0.5 + yF) * (0.5 + yF);
  double Wzm = 0.0, Wz0 = 0.0, Wzp = 0.0;
  if (g->nz >= 3) {
    Wzm = 0.5 * (0.5 - zF) * (0.5 - zF);
    Wz0 = 0.75 - zF * zF;
    Wzp = 0.5 * (0.5 + zF) * (0.5 + zF);
  } else if (g->nz == 2) { Wzm = 1.0 - zF; Wz0 = 0.0; Wzp = zF; }
  double dWxm = (-1.0) * (0.5 - xF) / g->dx, dWx0 = (-1.0) * 2. * xF / g->dx, dWxp = (-1.0) * -(0.5 + xF) / g->dx;
  double dWym = (-1.0) * (0.5 - yF) / g->dy, dWy0 = (-1.0) * 2. * yF / g->dy, dWyp = (-1.0) * -(0.5 + yF) / g->dy;
  double dWzm = 0.0, dWz0 = 0.0, dWzp = 0.0;
  if (g->nz >= 3) {
    dWzm = (-1.0) * (0.5 - zF) / g->dz;
    dWz0 = (-1.0) * 2. * zF / g->dz;
    dWzp = (-1.0) * -(0.5 + zF) / g->dz;
  } else if (g->nz == 2) {
    dWzm = (-1.0) * 1.0 / g->dz;
    dWz0 = 0.0;
    dWzp = (-1.0) * -1.0 / g->dz;
  }
  int iZm = iZ - 1, iZp = iZ + 1;
  if (iZm < 0) iZm = g->nz - 1;
  if (iZp >= g->nz) iZp = 0;
  if (g->nz >= 3) {
    *gX = sheet_grad(g, iZm, iX, iY, dWxm, dWx0, dWxp, Wym, Wy0, Wyp) * Wzm +
          sheet_grad(g, iZ, iX, iY, dWxm, dWx0, dWxp, Wym, Wy0, Wyp) * Wz0 +
          sheet_grad(g, iZp, iX, iY, dWxm, dWx0, dWxp, Wym, Wy0, Wyp) * Wzp;
    *gY = sheet_grad(g, iZm, iX, iY, Wxm, Wx0, Wxp, dWym, dWy0, dWyp) * Wzm +
          sheet_grad(g, iZ, iX, iY, Wxm, Wx0, Wxp, dWym, dWy0, dWyp) * Wz0 +
          sheet_grad(g, iZp, iX, iY, Wxm, Wx0, Wxp, dWym, dWy0, dWyp) * Wzp;
    *gZ = sheet_grad(g, iZm, iX, iY, Wxm, Wx0, Wxp, Wym, Wy0, Wyp) * dWzm +
          sheet_grad(g, iZ, iX, iY, Wxm, Wx0, Wxp, Wym, Wy0, Wyp) * dWz0 +
          sheet_grad(g, iZp, iX, iY, Wxm, Wx0, Wxp, Wym, Wy0, Wyp) * dWzp;
  }
  if (g->nz == 2) {
    *gX = sheet_grad(g, 0, iX, iY, dWxm, dWx0, dWxp, Wym, Wy0, Wyp) * Wzm +
          sheet_grad(g, 1, iX, iY, dWxm, dWx0, dWxp, Wym, Wy0, Wyp) * Wzp;
    *gY = sheet_grad(g, 0, iX, iY, Wxm, Wx0, Wxp, dWym, dWy0, dWyp) * Wzm +
          sheet_grad(g, 1, iX, iY, Wxm, Wx0, Wxp, dWym, dWy0, dWyp) * Wzp;
    *gZ = sheet_grad(g, 0, iX, iY, Wxm, Wx0, Wxp, Wym, Wy0, Wyp) * dWzm +
          sheet_grad(g, 1, iX, iY, Wxm, Wx0, Wxp, Wym, Wy0, Wyp) * dWzp;
  }
  if (g->nz == 1) {
    *gX = sheet_grad(g, 0, iX, iY, dWxm, dWx0, dWxp, Wym, Wy0, Wyp);
    *gY = sheet_grad(g, 0, iX, iY, Wxm, Wx0, Wxp, dWym, dWy0, dWyp);
    *gZ = 0.0;
  }
}

/* ------------------------------------------------------------------------ */
/* FFT.  FFTW's r2c/c2r conventions (row-major, last axis halved, forward
 * exponent -1, unnormalised).  Iterative radix-2; power-of-two sizes only. */
static void fft_core(double* re, double* im, int n, int inverse) {
  for (int i = 1, j = 0; i < n; i++) { /* bit reversal */
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      double t = re[i]; re[i] = re[j]; re[j] = t;
      t = im[i]; im[i] = im[j]; im[j] = t;
    }
  }
  double sgn = inverse ? 1.0 : -1.0;
  for (int len = 2; len <= n; len <<= 1) {
    int half = len >> 1;
    for (int k = 0; k < half; k++) {
      /* twiddle exp(sgn*2*pi*i*k/len) from the exact rational angle */
      double ang = 2.0 * 3.14159265358979323846 * (double)k / (double)len;
      double wr = cos(ang), wi = sgn * sin(ang);
      for (int s = 0; s < n; s += len) {
        int a = s + k, c = s + k + half;
        double tr = re[c] * wr - im[c] * wi;
        double ti = re[c] * wi + im[c] * wr;
        re[c] = re[a] - tr; im[c] = im[a] - ti;
        re[a] += tr; im[a] += ti;
      }
    }
  }
}
void orc_fft_c2c(double* re, double* im, int n, int inverse) { fft_core(re, im, n, inverse); }

/* r2c over the last axis of a [lines][n] real block -> [lines][n/2+1] complex */
static void r2c_lines(const double* in, int64_t lines, int n, double* out) {
  int h = n / 2 + 1;
#pragma omp parallel num_threads(g_threads) if (g_threads > 1)
  {
    double* re = (double*)malloc(sizeof(double) * n);
    double* im = (double*)malloc(sizeof(double) * n);
#pragma omp for
    for (int64_t l = 0; l < lines; l++) {
      for (int j = 0; j < n; j++) { re[j] = in[l * n + j]; im[j] = 0.0; }
      fft_core(re, im, n, 0);
      for (int k = 0; k < h; k++) { out[(l * h + k) * 2] = re[k]; out[(l * h + k) * 2 + 1] = im[k]; }
    }
    free(re); free(im);
  }
}
/* complex transform along an axis of length n with element stride `st`
 * (in complex units) over `cnt` lines whose starts are given by start(l). */
static void c2c_strided(double* data, int64_t outer, int64_t inner, int n, int inverse) {
  /* data is [outer][n][inner] complex; transform along the middle axis */
#pragma omp parallel num_threads(g_threads) if (g_threads > 1)
  {
    double* re = (double*)malloc(sizeof(double) * n);
    double* im = (double*)malloc(sizeof(double) * n);
#pragma omp for
    for (int64_t l = 0; l < outer * inner; l++) {
      int64_t o = l / inner, q = l % inner;
      for (int j = 0; j < n; j++) {
        size_t idx = ((size_t)(o * n + j) * inner + q) * 2;
        re[j] = data[idx]; im[j] = data[idx + 1];
      }
      fft_core(re, im, n, inverse);
      for (int j = 0; j < n; j++) {
        size_t idx = ((size_t)(o * n + j) * inner + q) * 2;
        data[idx] = re[j]; data[idx + 1] = im[j];
      }
    }
    free(re); free(im);
  }
}
/* c2r over the last axis: [lines][n/2+1] complex -> [lines][n] real (Hermitian
 * extension; DC/Nyquist imaginary parts do not reach the real output). */
static void c2r_lines(const double* in, int64_t lines, int n, double* out) {
  int h = n / 2 + 1;
#pragma omp parallel num_threads(g_threads) if (g_threads > 1)
  {
    double* re = (double*)malloc(sizeof(double) * n);
    double* im = (double*)malloc(sizeof(double) * n);
#pragma omp for
    for (int64_t l = 0; l < lines; l++) {
      for (int k = 0; k < h; k++) { re[k] = in[(l * h + k) * 2]; im[k] = in[(l * h + k) * 2 + 1]; }
      for (int k = h; k < n; k++) { re[k] = re[n - k]; im[k] = -im[n - k]; }
      fft_core(re, im, n, 1);
      for (int j = 0; j < n; j++) out[l * n + j] = re[j];
    }
    free(re); free(im);
  }
}
void orc_fft_r2c_2d(const double* in, int n0, int n1, double* out) {
  r2c_lines(in, n0, n1, out);
  c2c_strided(out, 1, n1 / 2 + 1, n0, 0);
}
void orc_fft_c2r_2d(const double* in, int n0, int n1, double* out) {
  size_t sz = (size_t)n0 * (n1 / 2 + 1) * 2;
  double* tmp = (double*)malloc(sizeof(double) * sz);
  memcpy(tmp, in, sizeof(double) * sz);
  c2c_strided(tmp, 1, n1 / 2 + 1, n0, 1);
  c2r_lines(tmp, n0, n1, out);
  free(tmp);
}
/* Optional replacement of the two 3D transforms (same layout and conventions: FFTW's
 * row-major r2c/c2r, unnormalised).  Only the CPU-baseline timing installs one, to time the
 * oracle with a multi-threaded library FFT of the FFTW class in place of fft_core; the
 * tests never do. */
static orc_fft3_fn g_r2c3_hook = NULL, g_c2r3_hook = NULL;
void orc_set_fft3_hook(orc_fft3_fn r2c, orc_fft3_fn c2r) {
  g_r2c3_hook = r2c;
  g_c2r3_hook = c2r;
}
void orc_fft_r2c_3d(const double* in, int n0, int n1, int n2, double* out) {
  if (g_r2c3_hook) {
    g_r2c3_hook(in, n0, n1, n2, out);
    return;
  }
  int h = n2 / 2 + 1;
  r2c_lines(in, (int64_t)n0 * n1, n2, out);
  c2c_strided(out, n0, h, n1, 0);     /* along axis 1 */
  c2c_strided(out, 1, (int64_t)n1 * h, n0, 0); /* along axis 0 */
}
void orc_fft_c2r_3d(const double* in, int n0, int n1, int n2, double* out) {
  if (g_c2r3_hook) {
    g_c2r3_hook(in, n0, n1, n2, out);
    return;
  }
  int h = n2 / 2 + 1;
  size_t sz = (size_t)n0 * n1 * h * 2;
  double* tmp = (double*)malloc(sizeof(double) * sz);
  memcpy(tmp, in, sizeof(double) * sz);
  c2c_strided(tmp, 1, (int64_t)n1 * h, n0, 1);
  c2c_strided(tmp, n0, h, n1, 1);
  c2r_lines(tmp, (int64_t)n0 * n1, n2, out);
  free(tmp);
}

/* Padded FFT length: the reference pads to 2n (PoissonSolverFFT2D.cc:27-28).
 * Any length >= 2n-1 gives the same linear convolution on the n-point region;
 * power-of-two n gives exactly 2n. */
static int pad_len(int n) {
  int m = 1;
  while (m < 2 * n - 1) m <<= 1;
  return m;
}

/* ------------------------------------------------------------------------ */
/* PoissonSolverFFT2D: src/spacecharge/PoissonSolverFFT2D.cc */
struct orc_poisson2d {
  int nx, ny, M0, M1;
  double xMin, xMax, yMin, yMax, dx, dy;
  double* ghat; /* [M0][M1/2+1][2] */
  double re00;
  double *in, *spec, *res;
};
static void poisson2d_define_green(orc_poisson2d* p) { /* :115-172 */
  int M0 = p->M0, M1 = p->M1;
  double* G = (double*)malloc(sizeof(double) * M0 * M1);
  for (int iY = 0; iY <= M1 / 2; iY++) {
    double rTransY = iY * p->dy;
    for (int iX = 0; iX <= M0 / 2; iX++) {
      double rTransX = iX * p->dx;
      double rTot2 = rTransX * rTransX + rTransY * rTransY;
      G[iX * M1 + iY] = (iX != 0 || iY != 0) ? -log(rTot2) / 2 : 0.0;
    }
    for (int iX = M0 / 2 + 1; iX < M0; iX++) G[iX * M1 + iY] = G[(M0 - iX) * M1 + iY];
  }
  for (int iY = M1 / 2 + 1; iY < M1; iY++)
    for (int iX = 0; iX < M0; iX++) G[iX * M1 + iY] = G[iX * M1 + (M1 - iY)];
  orc_fft_r2c_2d(G, M0, M1, p->ghat);
  p->re00 = p->ghat[0];
  free(G);
}
orc_poisson2d* orc_poisson2d_create(int nx, int ny, double xMin, double xMax, double yMin, double yMax) {
  orc_poisson2d* p = (orc_poisson2d*)calloc(1, sizeof(orc_poisson2d));
  p->nx = nx; p->ny = ny; p->M0 = pad_len(nx); p->M1 = pad_len(ny);
  p->xMin = xMin; p->xMax = xMax; p->yMin = yMin; p->yMax = yMax;
  p->dx = (xMax - xMin) / (nx - 1);
  p->dy = (yMax - yMin) / (ny - 1);
  size_t h = (size_t)p->M0 * (p->M1 / 2 + 1) * 2;
  p->ghat = (double*)malloc(sizeof(double) * h);
  p->spec = (double*)malloc(sizeof(double) * h);
  p->in = (double*)malloc(sizeof(double) * p->M0 * p->M1);
  p->res = (double*)malloc(sizeof(double) * p->M0 * p->M1);
  poisson2d_define_green(p);
  return p;
}
void orc_poisson2d_destroy(orc_poisson2d* p) {
  if (!p) return;
  free(p->ghat); free(p->spec); free(p->in); free(p->res); free(p);
}
const double* orc_poisson2d_green_hat(const orc_poisson2d* p) { return p->ghat; }
/* :174-256 */
void orc_poisson2d_find_potential(orc_poisson2d* p, const orc_grid2d* rho, orc_grid2d* phi) {
  int M0 = p->M0, M1 = p->M1, H = M1 / 2 + 1;
  double scale_coeff = log(p->dx / rho->dx);
  double g00 = p->re00 + scale_coeff * (M0 * M1);
  memset(p->in, 0, sizeof(double) * M0 * M1);
  for (int i = 0; i < p->nx; i++)
    for (int j = 0; j < p->ny; j++) p->in[j + M1 * i] = rho->arr[i * rho->ny + j];
  orc_fft_r2c_2d(p->in, M0, M1, p->spec);
  for (int i = 0; i < M0; i++)
    for (int j = 0; j < H; j++) {
      size_t idx = (size_t)(j + H * i) * 2;
      double gr_re = ((idx == 0) ? g00 : p->ghat[idx]) - scale_coeff;
      double gr_im = p->ghat[idx + 1];
      double a = p->spec[idx], b = p->spec[idx + 1];
      p->spec[idx] = a * gr_re - b * gr_im;
      p->spec[idx + 1] = a * gr_im + b * gr_re;
    }
  orc_fft_c2r_2d(p->spec, M0, M1, p->res);
  double denom = 1.0 / (M0 * M1);
  for (int i = 0; i < p->nx; i++)
    for (int j = 0; j < p->ny; j++) phi->arr[i * phi->ny + j] = denom * p->res[j + M1 * i];
}

/* ------------------------------------------------------------------------ */
/* PoissonSolverFFT3D: src/spacecharge/PoissonSolverFFT3D.cc */
struct orc_poisson3d {
  int nx, ny, nz, M0, M1, M2;
  int nbunches;   /* nBunches_ (:34) */
  double lambda;  /* spacing of the external bunches (:108-110) */
  double xMin, xMax, yMin, yMax, zMin, zMax, dx, dy, dz;
  double* ghat; /* [M0][M1][M2/2+1][2] */
  double *in, *spec, *res;
};
static void poisson3d_define_green(orc_poisson3d* p) { /* :160-247 */
  int M0 = p->M0, M1 = p->M1, M2 = p->M2;
  double* G = p->in; /* reuse the real work array */
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1)
  for (int iZ = 0; iZ <= M2 / 2; iZ++) {
    double rZ = iZ * p->dz, rZ2 = rZ * rZ;
    for (int iY = 0; iY <= M1 / 2; iY++) {
      double rY = iY * p->dy, rY2 = rY * rY;
      for (int iX = 0; iX <= M0 / 2; iX++) {
        double rX = iX * p->dx, rX2 = rX * rX;
        double rTot = sqrt(rX2 + rY2 + rZ2);
        double externalPhi = 0.;
        if (p->nbunches != 0) { /* :184-196 */
          for (int iB = -p->nbunches / 2; iB <= p->nbunches / 2; iB++) {
            if (iB == 0) continue;
            double rZt = rZ + iB * p->lambda;
            double rTotExt = sqrt(rX2 + rY2 + rZt * rZt);
            if (rTotExt > 1.0e-10) externalPhi += 1.0 / rTotExt;
          }
        }
        G[iZ + (size_t)M2 * (iY + (size_t)M1 * iX)] =
            (iX != 0 || iY != 0 || iZ != 0) ? 1. / rTot + externalPhi : externalPhi;
      }
      for (int iX = M0 / 2 + 1; iX < M0; iX++)
        G[iZ + (size_t)M2 * (iY + (size_t)M1 * iX)] = G[iZ + (size_t)M2 * (iY + (size_t)M1 * (M0 - iX))];
    }
    for (int iY = M1 / 2 + 1; iY < M1; iY++)
      for (int iX = 0; iX < M0; iX++)
        G[iZ + (size_t)M2 * (iY + (size_t)M1 * iX)] = G[iZ + (size_t)M2 * ((M1 - iY) + (size_t)M1 * iX)];
  }
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1)
  for (int iX = 0; iX < M0; iX++)
    for (int iY = 0; iY < M1; iY++)
      for (int iZ = M2 / 2 + 1; iZ < M2; iZ++)
        G[iZ + (size_t)M2 * (iY + (size_t)M1 * iX)] = G[(M2 - iZ) + (size_t)M2 * (iY + (size_t)M1 * iX)];
  orc_fft_r2c_3d(G, M0, M1, M2, p->ghat);
}
void orc_poisson3d_set_grid(orc_poisson3d* p, double xMin, double xMax, double yMin, double yMax,
                            double zMin, double zMax) { /* :136-147 */
  p->xMin = xMin; p->xMax = xMax; p->dx = (p->xMax - p->xMin) / (p->nx - 1);
  p->yMin = yMin; p->yMax = yMax; p->dy = (p->yMax - p->yMin) / (p->ny - 1);
  p->zMin = zMin; p->zMax = zMax; p->dz = (p->zMax - p->zMin) / p->nz;
  poisson3d_define_green(p);
}
orc_poisson3d* orc_poisson3d_create(int nx, int ny, int nz, double xMin, double xMax, double yMin,
                                    double yMax, double zMin, double zMax) {
  orc_poisson3d* p = (orc_poisson3d*)calloc(1, sizeof(orc_poisson3d));
  p->nx = nx; p->ny = ny; p->nz = nz;
  p->M0 = pad_len(nx); p->M1 = pad_len(ny); p->M2 = pad_len(nz);
  if (nz == 1) p->M2 = 2;
  size_t nreal = (size_t)p->M0 * p->M1 * p->M2;
  size_t ncplx = (size_t)p->M0 * p->M1 * (p->M2 / 2 + 1) * 2;
  p->in = (double*)malloc(sizeof(double) * nreal);
  p->res = (double*)malloc(sizeof(double) * nreal);
  p->ghat = (double*)malloc(sizeof(double) * ncplx);
  p->spec = (double*)malloc(sizeof(double) * ncplx);
  orc_poisson3d_set_grid(p, xMin, xMax, yMin, yMax, zMin, zMax);
  return p;
}
void orc_poisson3d_destroy(orc_poisson3d* p) {
  if (!p) return;
  free(p->in); free(p->res); free(p->ghat); free(p->spec); free(p);
}
void orc_poisson3d_set_ext_bunches(orc_poisson3d* p, int n) { /* :101-106 */
  if (n % 2 != 0) n = n + 1;
  p->nbunches = n;
}
void orc_poisson3d_set_ext_spacing(orc_poisson3d* p, double lambda) { p->lambda = lambda; }
void orc_poisson3d_limits(const orc_poisson3d* p, double o[6]) {
  o[0] = p->xMin; o[1] = p->xMax; o[2] = p->yMin; o[3] = p->yMax; o[4] = p->zMin; o[5] = p->zMax;
}
/* :249-347 */
void orc_poisson3d_find_potential(orc_poisson3d* p, const orc_grid3d* rho, orc_grid3d* phi) {
  int M0 = p->M0, M1 = p->M1, M2 = p->M2, H = M2 / 2 + 1;
  double scale_coeff = p->dx / rho->dx;
  size_t nreal = (size_t)M0 * M1 * M2;
  memset(p->in, 0, sizeof(double) * nreal);
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1)
  for (int i = 0; i < p->nx; i++)
    for (int j = 0; j < p->ny; j++)
      for (int k = 0; k < p->nz; k++)
        p->in[k + (size_t)M2 * (j + (size_t)M1 * i)] = A3(rho, k, i, j);
  orc_fft_r2c_3d(p->in, M0, M1, M2, p->spec);
  size_t nc = (size_t)M0 * M1 * H;
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1)
  for (int64_t idx = 0; idx < (int64_t)nc; idx++) {
    double gr_re = p->ghat[2 * idx], gr_im = p->ghat[2 * idx + 1];
    double a = p->spec[2 * idx], b = p->spec[2 * idx + 1];
    p->spec[2 * idx] = a * gr_re - b * gr_im;
    p->spec[2 * idx + 1] = a * gr_im + b * gr_re;
  }
  orc_fft_c2r_3d(p->spec, M0, M1, M2, p->res);
  double denom = scale_coeff / ((double)M0 * M1 * M2);
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1)
  for (int i = 0; i < p->nx; i++)
    for (int j = 0; j < p->ny; j++)
      for (int k = 0; k < p->nz; k++)
        A3(phi, k, i, j) = denom * p->res[k + (size_t)M2 * (j + (size_t)M1 * i)];
}

/* ------------------------------------------------------------------------ */
/* SpaceChargeCalc2p5Drb: src/spacecharge/SpaceChargeCalc2p5Drb.cc */
struct orc_sc2p5drb {
  double xy_ratio;
  int n_long_avg;
  orc_poisson2d* solver;
  orc_grid2d rho, phi;
  orc_grid1d zg, zd;
  double mom[4]; /* x_c, y_c, a_bunch, totalMacrosize */
};
orc_sc2p5drb* orc_sc2p5drb_create(int nx, int ny, int nz, double xy_ratio) { /* :31-46 */
  orc_sc2p5drb* c = (orc_sc2p5drb*)calloc(1, sizeof(orc_sc2p5drb));
  c->xy_ratio = xy_ratio;
  c->n_long_avg = 3;
  c->solver = orc_poisson2d_create(nx, ny, -xy_ratio, xy_ratio, -1.0, 1.0);
  orc_grid2d* gs[2] = {&c->rho, &c->phi};
  for (int k = 0; k < 2; k++) {
    gs[k]->nx = nx; gs[k]->ny = ny;
    gs[k]->arr = (double*)calloc((size_t)nx * ny, sizeof(double));
    orc_grid2d_set_x(gs[k], -1.0, 1.0);
    orc_grid2d_set_y(gs[k], -1.0, 1.0);
  }
  orc_grid1d* zs[2] = {&c->zg, &c->zd};
  for (int k = 0; k < 2; k++) {
    zs[k]->nz = nz;
    zs[k]->arr = (double*)calloc(nz, sizeof(double));
    orc_grid1d_set_z(zs[k], -0.5, 0.5);
  }
  return c;
}
void orc_sc2p5drb_destroy(orc_sc2p5drb* c) {
  if (!c) return;
  orc_poisson2d_destroy(c->solver);
  free(c->rho.arr); free(c->phi.arr); free(c->zg.arr); free(c->zd.arr); free(c);
}
void orc_sc2p5drb_set_long_avg_n(orc_sc2p5drb* c, int n) { c->n_long_avg = n; }

/* :257-310 */
static void calc_long_derivative(orc_sc2p5drb* c) {
  int nZ = c->zg.nz;
  double z_step = c->zg.dz;
  if (nZ == 1) { c->zd.arr[0] = 0.; return; }
  if (nZ == 2) {
    double zDeriv = 0.;
    orc_grid1d_gradient(&c->zg, (grid1d_gridz(&c->zg, 0) + grid1d_gridz(&c->zg, 1)) * 0.5, &zDeriv);
    c->zd.arr[0] = zDeriv / z_step;
    c->zd.arr[1] = zDeriv / z_step;
    return;
  }
  int nAvg = c->n_long_avg;
  if (nZ < nAvg) nAvg = nZ;
  double S[5][2];
  for (int iz = 0; iz < nZ; iz++) {
    double z = grid1d_gridz(&c->zg, iz);
    int iStart = iz - nAvg / 2;
    if (iStart < 0) iStart = 0;
    if ((iStart + nAvg) >= nZ) iStart = nZ - nAvg;
    int iStop = iStart + nAvg;
    for (int j = 0; j < 5; j++)
      for (int k = 0; k < 2; k++) {
        S[j][k] = 0.;
        for (int i = iStart; i < iStop; i++) {
          double x = 1.0 * (i - iStart + 1);
          double y = c->zg.arr[i];
          S[j][k] += pow(x, j) * pow(y, k);
        }
      }
    double det = (S[0][0] * S[2][0] * S[4][0] - S[1][0] * S[1][0] * S[4][0] - S[0][0] * S[3][0] * S[3][0] +
                  2 * S[1][0] * S[2][0] * S[3][0] - S[2][0] * S[2][0] * S[2][0]);
    double a = (S[0][1] * S[1][0] * S[3][0] - S[1][1] * S[0][0] * S[3][0] - S[0][1] * S[2][0] * S[2][0] +
                S[1][1] * S[1][0] * S[2][0] + S[2][1] * S[0][0] * S[2][0] - S[2][1] * S[1][0] * S[1][0]) / det;
    double bb = (S[1][1] * S[0][0] * S[4][0] - S[0][1] * S[1][0] * S[4][0] + S[0][1] * S[2][0] * S[3][0] -
                 S[2][1] * S[0][0] * S[3][0] - S[1][1] * S[2][0] * S[2][0] + S[2][1] * S[1][0] * S[2][0]) / det;
    double y = 2 * a * (z - grid1d_gridz(&c->zg, iStart) + z_step) / z_step + bb;
    y = y / z_step;
    c->zd.arr[iz] = y / z_step;
  }
}

/* :170-255 */
static int sc2p5drb_analysis(orc_sc2p5drb* c, orc_bunch* b) {
  double e[6];
  orc_extrema_xyz(b, e);
  double xMin = e[0], xMax = e[1], yMin = e[2], yMax = e[3], zMin = e[4], zMax = e[5];
  if (xMin >= xMax || yMin >= yMax || zMin >= zMax) return -1;
  double xy_ratio_beam = (xMax - xMin) / (yMax - yMin);
  double width, center;
  if (xy_ratio_beam > c->xy_ratio) {
    center = (yMax + yMin) / 2.0;
    width = ((yMax - yMin) * (xy_ratio_beam / c->xy_ratio)) / 2.0;
    yMin = center - width;
    yMax = center + width;
  } else {
    center = (xMax + xMin) / 2.0;
    width = ((xMax - xMin) / (xy_ratio_beam / c->xy_ratio)) / 2.0;
    xMin = center - width;
    xMax = center + width;
  }
  orc_grid2d_set_x(&c->rho, xMin, xMax); orc_grid2d_set_y(&c->rho, yMin, yMax);
  orc_grid2d_set_x(&c->phi, xMin, xMax); orc_grid2d_set_y(&c->phi, yMin, yMax);
  orc_grid1d_set_z(&c->zg, zMin, zMax); orc_grid1d_set_z(&c->zd, zMin, zMax);
  memset(c->rho.arr, 0, sizeof(double) * c->rho.nx * c->rho.ny);
  memset(c->zg.arr, 0, sizeof(double) * c->zg.nz);
  memset(c->zd.arr, 0, sizeof(double) * c->zd.nz);
  orc_grid2d_bin_bunch(&c->rho, b);
  orc_grid1d_bin_bunch(&c->zg, b);
  calc_long_derivative(c);
  double x_avg = 0., x2_avg = 0., y_avg = 0., y2_avg = 0., total = 0.;
  for (int ix = 0; ix < c->rho.nx; ix++)
    for (int iy = 0; iy < c->rho.ny; iy++) {
      double val = c->rho.arr[ix * c->rho.ny + iy];
      double x = c->rho.xMin + ix * c->rho.dx;
      double y = c->rho.xMin + iy * c->rho.dx; /* quirk Q1: getGridX(iy), :238 */
      total += val;
      x_avg += x * val;
      x2_avg += x * x * val;
      y_avg += y * val;
      y2_avg += y * y * val;
    }
  x_avg /= total; x2_avg /= total; y_avg /= total; y2_avg /= total;
  c->mom[0] = x_avg;
  c->mom[1] = y_avg;
  x2_avg = fabs(x2_avg - x_avg * x_avg);
  y2_avg = fabs(y2_avg - y_avg * y_avg);
  c->mom[2] = sqrt(2 * (x2_avg + y2_avg));
  c->mom[3] = total;
  return 0;
}

/* :110-168 */
int orc_sc2p5drb_track(orc_sc2p5drb* c, orc_bunch* b, double length, double pipe_radius) {
  if (b->n < 2) return 0;
  if (sc2p5drb_analysis(c, b) != 0) return -1;
  double x_center = c->mom[0], y_center = c->mom[1], a_bunch = c->mom[2], total = c->mom[3];
  double z_step = c->zg.dz;
  orc_poisson2d_find_potential(c->solver, &c->rho, &c->phi);
  double beta = b->s.beta, gamma = b->s.gamma;
  double factor = 2 * length * b->s.classical_radius / ((beta * beta) * pow(gamma, 3));
  factor = factor / (z_step * total);
  double long_sc_factor_in = 1.0 + 2 * log(pipe_radius / a_bunch);
  double long_sc_factor_out = 2 * log(pipe_radius);
  double a_bunch_2 = a_bunch * a_bunch;
  double long_sc_factor = -length * b->s.classical_radius * b->s.mass / (gamma * gamma);
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1)
  for (int64_t i = 0; i < b->n; i++) {
    double x = b->x[i], y = b->y[i], z = b->z[i];
    double r2 = (x - x_center) * (x - x_center) + (y - y_center) * (y - y_center);
    double ex, ey;
    orc_grid2d_gradient(&c->phi, x, y, &ex, &ey);
    double ez = orc_grid1d_value(&c->zd, z);
    double Lfactor = -orc_grid1d_value(&c->zg, z) * factor;
    b->xp[i] += ex * Lfactor;
    b->yp[i] += ey * Lfactor;
    double coeff;
    if (r2 <= a_bunch_2) coeff = long_sc_factor_in - r2 / a_bunch_2;
    else coeff = long_sc_factor_out - log(r2);
    b->dE[i] += ez * long_sc_factor * coeff;
  }
  return 0;
}
int orc_sc2p5drb_get_grid(const orc_sc2p5drb* c, int which, double* out, double* lims) {
  if (which == 0 || which == 1) {
    const orc_grid2d* g = which == 0 ? &c->rho : &c->phi;
    if (out) memcpy(out, g->arr, sizeof(double) * g->nx * g->ny);
    if (lims) { lims[0] = g->xMin; lims[1] = g->xMax; lims[2] = g->yMin; lims[3] = g->yMax; }
    return g->nx * g->ny;
  }
  const orc_grid1d* g = which == 2 ? &c->zg : &c->zd;
  if (out) memcpy(out, g->arr, sizeof(double) * g->nz);
  if (lims) { lims[0] = g->zMin; lims[1] = g->zMax; }
  return g->nz;
}
void orc_sc2p5drb_get_moments(const orc_sc2p5drb* c, double o[4]) { memcpy(o, c->mom, sizeof(c->mom)); }

/* ------------------------------------------------------------------------ */
/* Boundary2D: src/spacecharge/BaseBoundary2D.cc, ShapedBoundary2D.cc */
struct orc_boundary2d {
  int np, nm, shape;        /* points, free-space modes, 0 none / 1 circle / 2 ellipse / 3 rect */
  double *bx, *by;
  double xMin, xMax, yMin, yMax, rnorm;
  double p0, p1;            /* circle r | ellipse a, b | rectangle a, b */
  double* lsq;              /* [(2 nm + 1)][np] */
  double* f;                /* [(2 nm + 1)] scratch */
};
/* the mode functions 1, (r/R)^k cos k phi, (r/R)^k sin k phi at (x, y) (:316-345) */
static void bnd_functions(const orc_boundary2d* b, double x, double y, double* f) {
  int nf = 2 * b->nm + 1;
  double r2 = x * x + y * y;
  if (r2 == 0.) {
    f[0] = 1.0;
    for (int i = 1; i < nf; i++) f[i] = 0.0;
    return;
  }
  double r = sqrt(r2);
  double sn = y / r, cs = x / r, snk = sn, csk = cs;
  double q = r / b->rnorm, rk = 1.0;
  f[0] = 1.;
  for (int k = 0; k < b->nm; k++) {
    rk *= q;
    f[2 * k + 1] = rk * csk;
    f[2 * k + 2] = rk * snk;
    double s1 = sn * csk + cs * snk;
    double c1 = cs * csk - sn * snk;
    snk = s1;
    csk = c1;
  }
}
/* In-place inverse by Gauss-Jordan elimination with full pivoting, columns swapped back
   at the end (:233-313, the Numerical Recipes scheme). 0 on success. */
static int bnd_gauss_jordan_inverse(double* a, int n) {
  int* pivoted = (int*)calloc(n, sizeof(int));
  int* prow = (int*)malloc(sizeof(int) * n);
  int* pcol = (int*)malloc(sizeof(int) * n);
  int rc = 0;
  for (int step = 0; step < n && rc == 0; step++) {
    double big = 0.0;
    int irow = 0, icol = 0;
    for (int j = 0; j < n; j++) {
      if (pivoted[j] == 1) continue;
      for (int k = 0; k < n; k++) {
        if (pivoted[k] == 0) {
          if (fabs(a[j * n + k]) >= big) {
            big = fabs(a[j * n + k]);
            irow = j;
            icol = k;
          }
        } else if (pivoted[k] > 1) {
          rc = 1;
        }
      }
    }
    if (rc) break;
    pivoted[icol] += 1;
    if (irow != icol)
      for (int l = 0; l < n; l++) {
        double t = a[irow * n + l];
        a[irow * n + l] = a[icol * n + l];
        a[icol * n + l] = t;
      }
    prow[step] = irow;
    pcol[step] = icol;
    if (a[icol * n + icol] == 0.0) { rc = 2; break; }
    double inv = 1.0 / a[icol * n + icol];
    a[icol * n + icol] = 1.0;
    for (int l = 0; l < n; l++) a[icol * n + l] *= inv;
    for (int r = 0; r < n; r++) {
      if (r == icol) continue;
      double d = a[r * n + icol];
      a[r * n + icol] = 0.0;
      for (int l = 0; l < n; l++) a[r * n + l] -= a[icol * n + l] * d;
    }
  }
  if (rc == 0)
    for (int l = n - 1; l >= 0; l--)
      if (prow[l] != pcol[l])
        for (int k = 0; k < n; k++) {
          double t = a[k * n + prow[l]];
          a[k * n + prow[l]] = a[k * n + pcol[l]];
          a[k * n + pcol[l]] = t;
        }
  free(pivoted); free(prow); free(pcol);
  return rc;
}
orc_boundary2d* orc_boundary2d_create(int np, int nm) { /* :22-95 */
  if (2 * nm + 1 >= np) return NULL;
  orc_boundary2d* b = (orc_boundary2d*)calloc(1, sizeof(orc_boundary2d));
  b->np = np; b->nm = nm; b->shape = 0;
  b->bx = (double*)malloc(sizeof(double) * np);
  b->by = (double*)malloc(sizeof(double) * np);
  for (int i = 0; i < np; i++) b->bx[i] = b->by[i] = DBL_MAX;
  b->lsq = (double*)calloc((size_t)(2 * nm + 1) * np, sizeof(double));
  b->f = (double*)calloc(2 * nm + 1, sizeof(double));
  b->xMin = DBL_MAX; b->xMax = -DBL_MAX; b->yMin = DBL_MAX; b->yMax = -DBL_MAX;
  return b;
}
void orc_boundary2d_destroy(orc_boundary2d* b) {
  if (!b) return;
  free(b->bx); free(b->by); free(b->lsq); free(b->f); free(b);
}
void orc_boundary2d_set_point(orc_boundary2d* b, int i, double x, double y) {
  if (i >= 0 && i < b->np) { b->bx[i] = x; b->by[i] = y; }
}
/* initializeBPs (:188-230): limits, norm radius, the least-squares projector
   LSQ = (F^T F)^-1 F^T with F[j][k] = f_k(point j). */
int orc_boundary2d_initialize(orc_boundary2d* b) {
  int nf = 2 * b->nm + 1;
  b->xMin = DBL_MAX; b->xMax = -DBL_MAX; b->yMin = DBL_MAX; b->yMax = -DBL_MAX;
  for (int i = 0; i < b->np; i++) {
    double x = b->bx[i], y = b->by[i];
    if (b->xMin > x) b->xMin = x;
    if (b->xMax < x) b->xMax = x;
    if (b->yMin > y) b->yMin = y;
    if (b->yMax < y) b->yMax = y;
  }
  b->rnorm = sqrt((b->xMin - b->xMax) * (b->xMin - b->xMax) + (b->yMin - b->yMax) * (b->yMin - b->yMax));
  double* m = (double*)calloc((size_t)nf * nf, sizeof(double));
  for (int p = 0; p < b->np; p++) {
    bnd_functions(b, b->bx[p], b->by[p], b->f);
    for (int i = 0; i < nf; i++)
      for (int j = 0; j < nf; j++) m[i * nf + j] += b->f[i] * b->f[j];
  }
  int rc = bnd_gauss_jordan_inverse(m, nf);
  for (int i = 0; i < nf; i++)
    for (int j = 0; j < b->np; j++) {
      double acc = 0.0;
      bnd_functions(b, b->bx[j], b->by[j], b->f);
      for (int k = 0; k < nf; k++) acc += m[i * nf + k] * b->f[k];
      b->lsq[(size_t)i * b->np + j] = acc;
    }
  free(m);
  return rc;
}
/* ShapedBoundary2D(nPoints, nModes, shape, xDim, yDim) (ShapedBoundary2D.cc:9-127) */
orc_boundary2d* orc_boundary2d_create_shaped(int np_in, int nm, int shape, double xDim, double yDim) {
  const double PI = 3.14159265358979323846;
  int np = 4 * (np_in / 4);
  orc_boundary2d* b = orc_boundary2d_create(np, nm);
  if (!b) return NULL;
  b->shape = shape;
  if (shape == 1) {
    double r = xDim / 2;
    b->p0 = r;
    for (int i = 0; i < np; i++) orc_boundary2d_set_point(b, i, r * sin(2 * i * PI / np), r * cos(2 * i * PI / np));
  } else if (shape == 2) {
    /* points of equal arc length: iterate the first quadrant until it spans pi/2 */
    double a = xDim / 2, bb = yDim / 2;
    b->p0 = a; b->p1 = bb;
    double ds = 2 * PI * sqrt(a * bb) / np, resid = 1.0;
    double* th_arr = (double*)calloc(np + 1, sizeof(double));
    while (resid > 1.0e-08 || resid < -1.0e-08) {
      double denom = bb;
      double dth = ds / denom;
      double th = dth / 2.;
      th_arr[0] = 0.;
      for (int i = 0; i < np / 4; i++) {
        denom = bb * bb + (a * a - bb * bb) * sin(th) * sin(th);
        denom = sqrt(denom);
        dth = ds / denom;
        th_arr[i + 1] = th_arr[i] + dth;
        th += dth;
      }
      resid = th_arr[np / 4] - PI / 2.;
      ds *= PI / (2. * th_arr[np / 4]);
    }
    for (int i = 0; i < np / 4; i++) th_arr[np / 2 - i] = PI - th_arr[i];
    for (int i = 0; i < np / 2; i++) th_arr[np / 2 + i] = PI + th_arr[i];
    for (int i = 0; i < np; i++) orc_boundary2d_set_point(b, i, a * cos(th_arr[i]), bb * sin(th_arr[i]));
    free(th_arr);
  } else if (shape == 3) {
    double a = xDim, bb = yDim;
    b->p0 = a; b->p1 = bb;
    double dx = a / (np / 4), dy = bb / (np / 4);
    int k = 0;
    for (int i = 0; i < np / 4; i++) orc_boundary2d_set_point(b, k++, (-a / 2.0) + i * dx, bb / 2.0);
    for (int i = 0; i < np / 4; i++) orc_boundary2d_set_point(b, k++, a / 2.0, bb / 2.0 - i * dy);
    for (int i = 0; i < np / 4; i++) orc_boundary2d_set_point(b, k++, a / 2.0 - i * dx, -bb / 2.0);
    for (int i = 0; i < np / 4; i++) orc_boundary2d_set_point(b, k++, -a / 2.0, (-bb / 2.0) + i * dy);
  } else {
    orc_boundary2d_destroy(b);
    return NULL;
  }
  orc_boundary2d_initialize(b);
  return b;
}
/* isInside: BaseBoundary2D.cc:139-147 (open box) / ShapedBoundary2D.cc:284-304 */
int orc_boundary2d_is_inside(const orc_boundary2d* b, double x, double y) {
  int in = 0;
  if (b->shape == 0) in = x > b->xMin && x < b->xMax && y > b->yMin && y < b->yMax;
  else if (b->shape == 1) in = x * x + y * y < b->p0 * b->p0;
  else if (b->shape == 2) {
    double xx = x / b->p0, yy = y / b->p1;
    in = (xx * xx + yy * yy) < 1.0;
  } else in = fabs(x / b->p0) < 0.5 && fabs(y / b->p1) < 0.5;
  return in ? 1 : -1;
}
/* addBoundaryPotential (:349-458): phi on the boundary points, least-squares mode
   coefficients, and their field subtracted at every node. */
void orc_boundary2d_add_potential(orc_boundary2d* b, orc_grid2d* phi) {
  int nf = 2 * b->nm + 1;
  double* bphi = (double*)malloc(sizeof(double) * b->np);
  double* coef = (double*)malloc(sizeof(double) * nf);
  for (int j = 0; j < b->np; j++) bphi[j] = orc_grid2d_value(phi, b->bx[j], b->by[j]);
  for (int i = 0; i < nf; i++) {
    coef[i] = 0.0;
    for (int j = 0; j < b->np; j++) coef[i] += b->lsq[(size_t)i * b->np + j] * bphi[j];
  }
  for (int i = 0; i < phi->nx; i++)
    for (int j = 0; j < phi->ny; j++) {
      double v = 0.;
      bnd_functions(b, phi->xMin + i * phi->dx, phi->yMin + j * phi->dy, b->f);
      for (int k = 0; k < nf; k++) v += coef[k] * b->f[k];
      phi->arr[i * phi->ny + j] -= v;
    }
  free(bphi); free(coef);
}
void orc_boundary2d_info(const orc_boundary2d* b, int* np, double lims[4]) {
  *np = b->np;
  lims[0] = b->xMin; lims[1] = b->xMax; lims[2] = b->yMin; lims[3] = b->yMax;
}
void orc_boundary2d_points(const orc_boundary2d* b, double* x, double* y) {
  memcpy(x, b->bx, sizeof(double) * b->np);
  memcpy(y, b->by, sizeof(double) * b->np);
}

/* ------------------------------------------------------------------------ */
/* SpaceChargeCalc2p5D: src/spacecharge/SpaceChargeCalc2p5D.cc (no boundary) */
/* PoissonSolverFFT2D::setGridXY (PoissonSolverFFT2D.cc:103-111) */
void orc_poisson2d_set_grid(orc_poisson2d* p, double xMin, double xMax, double yMin, double yMax) {
  p->xMin = xMin; p->xMax = xMax; p->yMin = yMin; p->yMax = yMax;
  p->dx = (xMax - xMin) / (p->nx - 1);
  p->dy = (yMax - yMin) / (p->ny - 1);
  poisson2d_define_green(p);
}
struct orc_sc2p5d {
  double xy_ratio;
  orc_poisson2d* solver;
  orc_grid2d rho, phi;
  orc_grid1d zg;
  double total;
  int rebuilds;
};
orc_sc2p5d* orc_sc2p5d_create(int nx, int ny, int nz, double xy_ratio) { /* :25-33 */
  orc_sc2p5d* c = (orc_sc2p5d*)calloc(1, sizeof(orc_sc2p5d));
  c->xy_ratio = xy_ratio;
  c->solver = orc_poisson2d_create(nx, ny, -xy_ratio, xy_ratio, -1.0, 1.0);
  orc_grid2d* gs[2] = {&c->rho, &c->phi};
  for (int k = 0; k < 2; k++) {
    gs[k]->nx = nx; gs[k]->ny = ny;
    gs[k]->arr = (double*)calloc((size_t)nx * ny, sizeof(double));
    orc_grid2d_set_x(gs[k], -1.0, 1.0);
    orc_grid2d_set_y(gs[k], -1.0, 1.0);
  }
  c->zg.nz = nz;
  c->zg.arr = (double*)calloc(nz, sizeof(double));
  orc_grid1d_set_z(&c->zg, -0.5, 0.5);
  return c;
}
void orc_sc2p5d_destroy(orc_sc2p5d* c) {
  if (!c) return;
  orc_poisson2d_destroy(c->solver);
  free(c->rho.arr); free(c->phi.arr); free(c->zg.arr); free(c);
}
/* bunchAnalysis without a boundary (:113-225): the grid keeps the calculator's x/y ratio
   while the beam's ratio stays within 25% of it; otherwise the ratio follows the beam and
   the solver's Green function is rebuilt on the beam's box. */
static int sc2p5d_analysis(orc_sc2p5d* c, orc_bunch* b) {
  double e[6];
  orc_extrema_xyz(b, e);
  double xMin = e[0], xMax = e[1], yMin = e[2], yMax = e[3], zMin = e[4], zMax = e[5];
  if (xMin >= xMax || yMin >= yMax || zMin >= zMax) return -1;
  double xy_ratio_beam = (xMax - xMin) / (yMax - yMin);
  if (fabs(xy_ratio_beam - c->xy_ratio) < 0.25 * c->xy_ratio) {
    double width, center;
    if (xy_ratio_beam > c->xy_ratio) {
      center = (yMax + yMin) / 2.0;
      width = ((yMax - yMin) * (xy_ratio_beam / c->xy_ratio)) / 2.0;
      yMin = center - width;
      yMax = center + width;
    } else {
      center = (xMax + xMin) / 2.0;
      width = ((xMax - xMin) / (xy_ratio_beam / c->xy_ratio)) / 2.0;
      xMin = center - width;
      xMax = center + width;
    }
    c->xy_ratio = (xMax - xMin) / (yMax - yMin);
  } else {
    c->xy_ratio = xy_ratio_beam;
    orc_poisson2d_set_grid(c->solver, xMin, xMax, yMin, yMax);
    c->rebuilds++;
  }
  orc_grid2d_set_x(&c->rho, xMin, xMax); orc_grid2d_set_y(&c->rho, yMin, yMax);
  orc_grid2d_set_x(&c->phi, xMin, xMax); orc_grid2d_set_y(&c->phi, yMin, yMax);
  orc_grid1d_set_z(&c->zg, zMin, zMax);
  {
    const orc_poisson2d* p = c->solver; /* :204-212 */
    if (fabs((p->xMax - p->xMin) / (p->yMax - p->yMin) - c->xy_ratio) > 0.00000001) {
      orc_poisson2d_set_grid(c->solver, xMin, xMax, yMin, yMax);
      c->rebuilds++;
    }
  }
  memset(c->rho.arr, 0, sizeof(double) * c->rho.nx * c->rho.ny);
  memset(c->zg.arr, 0, sizeof(double) * c->zg.nz);
  orc_grid2d_bin_bunch(&c->rho, b);
  orc_grid1d_bin_bunch(&c->zg, b);
  c->total = 0.;
  for (int iz = 0; iz < c->zg.nz; iz++) c->total += c->zg.arr[iz];
  return 0;
}
int orc_sc2p5d_track(orc_sc2p5d* c, orc_bunch* b, double length);
/* bunchAnalysis with a boundary (:174-234): the grid is the boundary's box, only z comes
   from the bunch */
static int sc2p5d_analysis_boundary(orc_sc2p5d* c, orc_bunch* b, const orc_boundary2d* bd) {
  double e[6];
  orc_extrema_xyz(b, e);
  double xMin = bd->xMin, xMax = bd->xMax, yMin = bd->yMin, yMax = bd->yMax;
  double zMin = e[4], zMax = e[5];
  c->xy_ratio = (xMax - xMin) / (yMax - yMin);
  if (zMin >= zMax) return -1;
  orc_grid2d_set_x(&c->rho, xMin, xMax); orc_grid2d_set_y(&c->rho, yMin, yMax);
  orc_grid2d_set_x(&c->phi, xMin, xMax); orc_grid2d_set_y(&c->phi, yMin, yMax);
  orc_grid1d_set_z(&c->zg, zMin, zMax);
  {
    const orc_poisson2d* p = c->solver;
    if (fabs((p->xMax - p->xMin) / (p->yMax - p->yMin) - c->xy_ratio) > 0.00000001) {
      orc_poisson2d_set_grid(c->solver, xMin, xMax, yMin, yMax);
      c->rebuilds++;
    }
  }
  memset(c->rho.arr, 0, sizeof(double) * c->rho.nx * c->rho.ny);
  memset(c->zg.arr, 0, sizeof(double) * c->zg.nz);
  orc_grid2d_bin_bunch(&c->rho, b);
  orc_grid1d_bin_bunch(&c->zg, b);
  c->total = 0.;
  for (int iz = 0; iz < c->zg.nz; iz++) c->total += c->zg.arr[iz];
  return 0;
}
/* trackBunch (:77-120), with a boundary when bd is not NULL: its potential is added to
   phi and particles outside it get no kick */
int orc_sc2p5d_track_boundary(orc_sc2p5d* c, orc_bunch* b, double length, orc_boundary2d* bd) {
  if (b->n < 2) return 0;
  if (!bd) return orc_sc2p5d_track(c, b, length);
  if (sc2p5d_analysis_boundary(c, b, bd) != 0) return -1;
  double z_step = c->zg.dz;
  orc_poisson2d_find_potential(c->solver, &c->rho, &c->phi);
  orc_boundary2d_add_potential(bd, &c->phi);
  double beta = b->s.beta, gamma = b->s.gamma;
  double factor = 2 * length * b->s.classical_radius / ((beta * beta) * pow(gamma, 3));
  factor = factor / (z_step * c->total);
  for (int64_t i = 0; i < b->n; i++) {
    if (orc_boundary2d_is_inside(bd, b->x[i], b->y[i]) != 1) continue;
    double ex, ey;
    orc_grid2d_gradient(&c->phi, b->x[i], b->y[i], &ex, &ey);
    double Lfactor = -orc_grid1d_value(&c->zg, b->z[i]) * factor;
    b->xp[i] += ex * Lfactor;
    b->yp[i] += ey * Lfactor;
  }
  return 0;
}
/* trackBunch without a boundary (:76-111) */
int orc_sc2p5d_track(orc_sc2p5d* c, orc_bunch* b, double length) {
  if (b->n < 2) return 0;
  if (sc2p5d_analysis(c, b) != 0) return -1;
  double z_step = c->zg.dz;
  orc_poisson2d_find_potential(c->solver, &c->rho, &c->phi);
  double beta = b->s.beta, gamma = b->s.gamma;
  double factor = 2 * length * b->s.classical_radius / ((beta * beta) * pow(gamma, 3));
  factor = factor / (z_step * c->total);
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1)
  for (int64_t i = 0; i < b->n; i++) {
    double ex, ey;
    orc_grid2d_gradient(&c->phi, b->x[i], b->y[i], &ex, &ey);
    double Lfactor = -orc_grid1d_value(&c->zg, b->z[i]) * factor;
    b->xp[i] += ex * Lfactor;
    b->yp[i] += ey * Lfactor;
  }
  return 0;
}
int orc_sc2p5d_get_grid(const orc_sc2p5d* c, int which, double* out, double* lims) {
  if (which == 0 || which == 1) {
    const orc_grid2d* g = which == 0 ? &c->rho : &c->phi;
    if (out) memcpy(out, g->arr, sizeof(double) * g->nx * g->ny);
    if (lims) { lims[0] = g->xMin; lims[1] = g->xMax; lims[2] = g->yMin; lims[3] = g->yMax; }
    return g->nx * g->ny;
  }
  if (out) memcpy(out, c->zg.arr, sizeof(double) * c->zg.nz);
  if (lims) { lims[0] = c->zg.zMin; lims[1] = c->zg.zMax; }
  return c->zg.nz;
}
/* out: xy_ratio, solver xMin xMax yMin yMax, totalMacrosize, Green rebuilds */
void orc_sc2p5d_state(const orc_sc2p5d* c, double o[7]) {
  o[0] = c->xy_ratio;
  o[1] = c->solver->xMin; o[2] = c->solver->xMax; o[3] = c->solver->yMin; o[4] = c->solver->yMax;
  o[5] = c->total;
  o[6] = c->rebuilds;
}

/* ------------------------------------------------------------------------ */
/* SpaceChargeCalcSliceBySlice2D: src/spacecharge/SpaceChargeCalcSliceBySlice2D.cc */
struct orc_slice2d {
  double xy_ratio;
  orc_poisson2d* solver;
  orc_grid3d rho, phi;
  int rebuilds;
};
orc_slice2d* orc_slice2d_create(int nx, int ny, int nz, double xy_ratio) { /* :15-32 */
  orc_slice2d* c = (orc_slice2d*)calloc(1, sizeof(orc_slice2d));
  c->xy_ratio = xy_ratio;
  c->solver = orc_poisson2d_create(nx, ny, -xy_ratio, xy_ratio, -1.0, 1.0);
  orc_grid3d* gs[2] = {&c->rho, &c->phi};
  for (int k = 0; k < 2; k++) {
    gs[k]->nx = nx; gs[k]->ny = ny; gs[k]->nz = nz;
    gs[k]->arr = (double*)calloc((size_t)nx * ny * nz, sizeof(double));
  }
  return c;
}
void orc_slice2d_destroy(orc_slice2d* c) {
  if (!c) return;
  orc_poisson2d_destroy(c->solver);
  free(c->rho.arr); free(c->phi.arr); free(c);
}
/* bunchAnalysis (:116-218): the 2p5D ratio rule, or the boundary's box; the z grid spans
   the bunch */
static int slice2d_analysis(orc_slice2d* c, orc_bunch* b, const orc_boundary2d* bd) {
  double e[6];
  orc_extrema_xyz(b, e);
  double xMin = e[0], xMax = e[1], yMin = e[2], yMax = e[3], zMin = e[4], zMax = e[5];
  if (!bd) {
    if (xMin >= xMax || yMin >= yMax || zMin >= zMax) return -1;
    double xy_ratio_beam = (xMax - xMin) / (yMax - yMin);
    if (fabs(xy_ratio_beam - c->xy_ratio) < 0.25 * c->xy_ratio) {
      double width, center;
      if (xy_ratio_beam > c->xy_ratio) {
        center = (yMax + yMin) / 2.0;
        width = ((yMax - yMin) * (xy_ratio_beam / c->xy_ratio)) / 2.0;
        yMin = center - width;
        yMax = center + width;
      } else {
        center = (xMax + xMin) / 2.0;
        width = ((xMax - xMin) / (xy_ratio_beam / c->xy_ratio)) / 2.0;
        xMin = center - width;
        xMax = center + width;
      }
      c->xy_ratio = (xMax - xMin) / (yMax - yMin);
    } else {
      c->xy_ratio = xy_ratio_beam;
      orc_poisson2d_set_grid(c->solver, xMin, xMax, yMin, yMax);
      c->rebuilds++;
    }
  } else {
    xMin = bd->xMin; xMax = bd->xMax; yMin = bd->yMin; yMax = bd->yMax;
    c->xy_ratio = (xMax - xMin) / (yMax - yMin);
    if (zMin >= zMax) return -1;
  }
  orc_grid3d* gs[2] = {&c->rho, &c->phi};
  for (int k = 0; k < 2; k++) {
    orc_grid3d_set_x(gs[k], xMin, xMax);
    orc_grid3d_set_y(gs[k], yMin, yMax);
    orc_grid3d_set_z(gs[k], zMin, zMax);
  }
  const orc_poisson2d* p = c->solver; /* :203-209 */
  if (fabs((p->xMax - p->xMin) / (p->yMax - p->yMin) - c->xy_ratio) > 0.00000001) {
    orc_poisson2d_set_grid(c->solver, xMin, xMax, yMin, yMax);
    c->rebuilds++;
  }
  memset(c->rho.arr, 0, sizeof(double) * c->rho.nx * c->rho.ny * c->rho.nz);
  orc_grid3d_bin_bunch(&c->rho, b);
  return 0;
}
static orc_grid2d slice_view(orc_grid3d* g, int iz) {
  orc_grid2d s;
  s.nx = g->nx; s.ny = g->ny;
  s.xMin = g->xMin; s.xMax = g->xMax; s.yMin = g->yMin; s.yMax = g->yMax;
  s.dx = g->dx; s.dy = g->dy;
  s.arr = g->arr + (size_t)iz * g->nx * g->ny;
  return s;
}
/* trackBunch (:66-113): a 2D Poisson solve per z slice (plus the boundary's potential),
   then the transverse kick from the 3D gradient of phi */
int orc_slice2d_track(orc_slice2d* c, orc_bunch* b, double length, orc_boundary2d* bd) {
  if (b->n < 2) return 0;
  if (slice2d_analysis(c, b, bd) != 0) return -1;
  for (int iz = 0; iz < c->rho.nz; iz++) {
    orc_grid2d r = slice_view(&c->rho, iz), f = slice_view(&c->phi, iz);
    orc_poisson2d_find_potential(c->solver, &r, &f);
    if (bd) orc_boundary2d_add_potential(bd, &f);
  }
  double beta = b->s.beta, gamma = b->s.gamma;
  double factor = 2 * length / c->rho.dz * b->s.classical_radius / (pow(beta, 2) * pow(gamma, 3));
  for (int64_t i = 0; i < b->n; i++) {
    if (bd && orc_boundary2d_is_inside(bd, b->x[i], b->y[i]) != 1) continue;
    double ex, ey, ez;
    orc_grid3d_gradient(&c->phi, b->x[i], b->y[i], b->z[i], &ex, &ey, &ez);
    b->xp[i] -= ex * factor;
    b->yp[i] -= ey * factor;
  }
  return 0;
}
/* which 0 rho, 1 phi ([nz][nx][ny]); lims = x, y, z limits */
void orc_slice2d_get_grid(const orc_slice2d* c, int which, double* out, double* lims) {
  const orc_grid3d* g = which == 0 ? &c->rho : &c->phi;
  if (out) memcpy(out, g->arr, sizeof(double) * g->nx * g->ny * g->nz);
  if (lims) {
    lims[0] = g->xMin; lims[1] = g->xMax; lims[2] = g->yMin;
    lims[3] = g->yMax; lims[4] = g->zMin; lims[5] = g->zMax;
  }
}
void orc_slice2d_state(const orc_slice2d* c, double o[6]) {
  o[0] = c->xy_ratio;
  o[1] = c->solver->xMin; o[2] = c->solver->xMax; o[3] = c->solver->yMin; o[4] = c->solver->yMax;
  o[5] = c->rebuilds;
}

/* ------------------------------------------------------------------------ */
/* SpaceChargeCalc3D: src/spacecharge/SpaceChargeCalc3D.cc */
struct orc_sc3d {
  double xy_ratio, xz_ratio, ratio_limit;
  double frequency; /* frequency_ (:58) */
  orc_poisson3d* solver;
  orc_grid3d rho, phi;
  int rebuilds;
};
orc_sc3d* orc_sc3d_create(int nx, int ny, int nz) { /* :32-58 */
  orc_sc3d* c = (orc_sc3d*)calloc(1, sizeof(orc_sc3d));
  c->xy_ratio = 1.0; c->xz_ratio = 1.0; c->ratio_limit = 1.2;
  c->frequency = 402.5e+6;
  c->solver = orc_poisson3d_create(nx, ny, nz, -c->xy_ratio, c->xy_ratio, -1. / c->xy_ratio,
                                   1. / c->xy_ratio, -1. / c->xz_ratio, 1. / c->xz_ratio);
  orc_grid3d* gs[2] = {&c->rho, &c->phi};
  for (int k = 0; k < 2; k++) {
    gs[k]->nx = nx; gs[k]->ny = ny; gs[k]->nz = nz;
    gs[k]->arr = (double*)calloc((size_t)nx * ny * nz, sizeof(double));
    orc_grid3d_set_x(gs[k], c->solver->xMin, c->solver->xMax);
    orc_grid3d_set_y(gs[k], c->solver->yMin, c->solver->yMax);
    orc_grid3d_set_z(gs[k], c->solver->zMin, c->solver->zMax);
  }
  return c;
}
void orc_sc3d_destroy(orc_sc3d* c) {
  if (!c) return;
  orc_poisson3d_destroy(c->solver);
  free(c->rho.arr); free(c->phi.arr); free(c);
}
void orc_sc3d_set_ratio_limit(orc_sc3d* c, double r) { c->ratio_limit = r; }
void orc_sc3d_set_num_ext_bunches(orc_sc3d* c, int n) { orc_poisson3d_set_ext_bunches(c->solver, n); }
int orc_sc3d_get_num_ext_bunches(const orc_sc3d* c) { return c->solver->nbunches; }
void orc_sc3d_set_freq(orc_sc3d* c, double f) { c->frequency = f; }
double orc_sc3d_get_freq(const orc_sc3d* c) { return c->frequency; }
#define ORC_C_LIGHT 2.99792458e+8 /* OrbitConst.cc:27 */
int orc_sc3d_green_rebuilds(const orc_sc3d* c) { return c->rebuilds; }

/* :150-249 */
static int sc3d_analysis(orc_sc3d* c, orc_bunch* b) {
  double e[6];
  orc_extrema_xyz(b, e);
  double xMin = e[0], xMax = e[1], yMin = e[2], yMax = e[3], zMin = e[4], zMax = e[5];
  double gamma = b->s.gamma;
  double center = (zMax + zMin) / 2.0;
  double width = (zMax - zMin) * gamma / 2.0;
  zMin = center - width;
  zMax = center + width;
  if (xMin >= xMax || yMin >= yMax || zMin >= zMax) return -1;
  double xy_ratio_beam = (xMax - xMin) / (yMax - yMin);
  double xz_ratio_beam = (xMax - xMin) / (zMax - zMin);
  int change = 0;
  if (c->ratio_limit < xy_ratio_beam / c->xy_ratio || 1.0 / c->ratio_limit > xy_ratio_beam / c->xy_ratio ||
      c->ratio_limit < xz_ratio_beam / c->xz_ratio || 1.0 / c->ratio_limit > xz_ratio_beam / c->xz_ratio)
    change = 1;
  orc_poisson3d* s = c->solver;
  if (change) {
    c->xy_ratio = xy_ratio_beam;
    c->xz_ratio = xz_ratio_beam;
    orc_poisson3d_set_grid(s, -(xMax - xMin) / 2.0, (xMax - xMin) / 2.0, -(yMax - yMin) / 2.0,
                           (yMax - yMin) / 2.0, -(zMax - zMin) / 2.0, (zMax - zMin) / 2.0);
    c->rebuilds++;
  }
  double scx = (xMax - xMin) / (s->xMax - s->xMin);
  double scy = (yMax - yMin) / (s->yMax - s->yMin);
  double scz = (zMax - zMin) / (s->zMax - s->zMin);
  double sc = scx;
  if (sc < scy) sc = scy;
  if (sc < scz) sc = scz;
  center = (xMax + xMin) / 2.0;
  width = (s->xMax - s->xMin) * sc / 2.0;
  xMin = center - width; xMax = center + width;
  orc_grid3d_set_x(&c->rho, xMin, xMax); orc_grid3d_set_x(&c->phi, xMin, xMax);
  center = (yMax + yMin) / 2.0;
  width = (s->yMax - s->yMin) * sc / 2.0;
  yMin = center - width; yMax = center + width;
  orc_grid3d_set_y(&c->rho, yMin, yMax); orc_grid3d_set_y(&c->phi, yMin, yMax);
  center = (zMax + zMin) / 2.0;
  width = (s->zMax - s->zMin) * sc / (2.0 * gamma);
  zMin = center - width; zMax = center + width;
  orc_grid3d_set_z(&c->rho, zMin, zMax); orc_grid3d_set_z(&c->phi, zMin, zMax);
  memset(c->rho.arr, 0, sizeof(double) * (size_t)c->rho.nx * c->rho.ny * c->rho.nz);
  orc_grid3d_bin_bunch(&c->rho, b);
  center = (zMax + zMin) / 2.0;
  width = (zMax - zMin) * gamma / 2.0;
  zMin = center - width; zMax = center + width;
  orc_grid3d_set_z(&c->rho, zMin, zMax); orc_grid3d_set_z(&c->phi, zMin, zMax);
  return 0;
}
/* wrappedBunchAnalysis :251-328 */
static int sc3d_wrapped_analysis(orc_sc3d* c, orc_bunch* b) {
  double e[6];
  orc_extrema_xyz(b, e);
  double xMin = e[0], xMax = e[1], yMin = e[2], yMax = e[3], zMin = e[4], zMax = e[5];
  c->rho.wrap = 1;
  c->phi.wrap = 1;
  double gamma = b->s.gamma, beta = b->s.beta;
  double center = (zMax + zMin) / 2.0;
  double width = (zMax - zMin) / 2.0;
  zMin = center - width;
  zMax = center + width;
  double lambda = ORC_C_LIGHT * beta / c->frequency;
  double lambda_cm = ORC_C_LIGHT * beta * gamma / c->frequency;
  if (fabs(zMax) > lambda / 2 || fabs(zMin) > lambda / 2) {
    double extension_coeff = 1.0001;
    center = 0.;
    width = extension_coeff * lambda / 2;
    zMin = center - width;
    zMax = center + width;
  }
  if (xMin >= xMax || yMin >= yMax || zMin >= zMax) return -1;
  orc_grid3d_set_x(&c->rho, xMin, xMax); orc_grid3d_set_x(&c->phi, xMin, xMax);
  orc_grid3d_set_y(&c->rho, yMin, yMax); orc_grid3d_set_y(&c->phi, yMin, yMax);
  orc_grid3d_set_z(&c->rho, zMin, zMax); orc_grid3d_set_z(&c->phi, zMin, zMax);
  memset(c->rho.arr, 0, sizeof(double) * (size_t)c->rho.nx * c->rho.ny * c->rho.nz);
  orc_grid3d_bin_bunch_wrapped(&c->rho, b, lambda);
  center = (zMax + zMin) / 2.0;
  width = (zMax - zMin) * gamma / 2.0;
  zMin = center - width; zMax = center + width;
  orc_grid3d_set_z(&c->rho, zMin, zMax); orc_grid3d_set_z(&c->phi, zMin, zMax);
  orc_poisson3d_set_ext_spacing(c->solver, lambda_cm);
  orc_poisson3d_set_grid(c->solver, c->rho.xMin, c->rho.xMax, c->rho.yMin, c->rho.yMax,
                         c->rho.zMin, c->rho.zMax);
  c->rebuilds++;
  return 0;
}
/* :100-148 */
int orc_sc3d_track(orc_sc3d* c, orc_bunch* b, double length) {
  if (b->n < 2) return 0;
  int next = c->solver->nbunches;
  if (next > 0) {
    if (sc3d_wrapped_analysis(c, b) != 0) return -1;
  } else if (sc3d_analysis(c, b) != 0) {
    return -1;
  }
  orc_poisson3d_find_potential(c->solver, &c->rho, &c->phi);
  double gamma = b->s.gamma, beta = b->s.beta;
  double trans_factor = length * b->s.classical_radius / ((beta * beta) * (gamma * gamma));
  double long_factor = length * b->s.classical_radius * b->s.mass;
  double z_center = (c->phi.zMax + c->phi.zMin) / 2.0;
  double lambda_cm = ORC_C_LIGHT * beta * gamma / c->frequency;
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1)
  for (int64_t i = 0; i < b->n; i++) {
    double x = b->x[i], y = b->y[i];
    double z = (b->z[i] - z_center) * gamma + z_center;
    if (next > 0) z = remainder(z, lambda_cm);
    double ex, ey, ez;
    orc_grid3d_gradient(&c->phi, x, y, z, &ex, &ey, &ez);
    b->xp[i] += -ex * trans_factor;
    b->yp[i] += -ey * trans_factor;
    b->dE[i] += -ez * long_factor;
  }
  return 0;
}
int64_t orc_sc3d_get_grid(const orc_sc3d* c, int which, double* out, double* lims) {
  const orc_grid3d* g = which == 0 ? &c->rho : &c->phi;
  size_t n = (size_t)g->nx * g->ny * g->nz;
  if (out) memcpy(out, g->arr, sizeof(double) * n);
  if (lims) {
    lims[0] = g->xMin; lims[1] = g->xMax; lims[2] = g->yMin; lims[3] = g->yMax;
    lims[4] = g->zMin; lims[5] = g->zMax;
  }
  return (int64_t)n;
}

/* ------------------------------------------------------------------------ */
/* BunchTwissAnalysis: analyzeBunch (:49-142, no macrosize attribute branch)
 * + getCorrelation (:344-347) + getEmittance (:372-391). */
void orc_twiss_emittances(const orc_bunch* b, double o[3]) {
  double avg[6] = {0}, corr[36] = {0};
  double* c[6] = {b->x, b->xp, b->y, b->yp, b->z, b->dE};
  for (int64_t ip = 0; ip < b->n; ip++) {
    for (int i = 0; i < 6; i++) avg[i] += c[i][ip];
    for (int i = 0; i < 6; i++)
      for (int j = 0; j < i + 1; j++) corr[i + 6 * j] += c[i][ip] * c[j][ip];
  }
  double total = (double)b->n;
  double A[6], C[36];
  for (int i = 0; i < 6; i++) A[i] = avg[i] / total;
  for (int i = 0; i < 6; i++)
    for (int j = 0; j < i + 1; j++) {
      C[i + 6 * j] = corr[i + 6 * j] / total;
      C[j + 6 * i] = corr[i + 6 * j] / total;
    }
#define CORR(i, j) (C[(i) + 6 * (j)] - A[i] * A[j])
  for (int ic = 0; ic < 3; ic++) {
    double x2 = fabs(CORR(2 * ic, 2 * ic)), xp2 = fabs(CORR(2 * ic + 1, 2 * ic + 1));
    double xxp = CORR(2 * ic, 2 * ic + 1), xdE = CORR(2 * ic, 5), xpdE = CORR(2 * ic + 1, 5);
    double dE2 = fabs(CORR(5, 5));
    if (ic == 2 || dE2 == 0)
      o[ic] = sqrt(fabs(x2 * xp2 - xxp * xxp));
    else
      o[ic] = sqrt(fabs((x2 - xdE * xdE / dE2) * (xp2 - xpdE * xpdE / dE2) -
                        (xxp - xdE * xpdE / dE2) * (xxp - xdE * xpdE / dE2)));
  }
#undef CORR
}

/* ------------------------------------------------------------------------ */
/* Flat-array entry points with the same signatures as oracle/ref_harness.cc,
 * so the restatement and the reference build can be compared call for call. */
static orc_grid2d mk2(int nx, int ny, const double* l, double* arr) {
  orc_grid2d g;
  g.nx = nx; g.ny = ny; g.arr = arr;
  orc_grid2d_set_x(&g, l[0], l[1]);
  orc_grid2d_set_y(&g, l[2], l[3]);
  return g;
}
static orc_grid3d mk3(int nx, int ny, int nz, const double* l, double* arr) {
  orc_grid3d g;
  g.nx = nx; g.ny = ny; g.nz = nz; g.arr = arr;
  g.wrap = 0;
  orc_grid3d_set_x(&g, l[0], l[1]);
  orc_grid3d_set_y(&g, l[2], l[3]);
  orc_grid3d_set_z(&g, l[4], l[5]);
  return g;
}
void orc_flat_grid2d_bin(int nx, int ny, const double* l, long n, const double* x, const double* y,
                         const double* m, double* rho) {
  memset(rho, 0, sizeof(double) * nx * ny);
  orc_grid2d g = mk2(nx, ny, l, rho);
  for (long i = 0; i < n; i++) orc_grid2d_bin_value(&g, m[i], x[i], y[i]);
}
void orc_flat_grid2d_indfrac(int nx, int ny, const double* l, long n, const double* x,
                             const double* y, int* ix, double* fx, int* iy, double* fy) {
  orc_grid2d g = mk2(nx, ny, l, NULL);
  for (long i = 0; i < n; i++) {
    orc_grid2d_ind_frac_x(&g, x[i], ix + i, fx + i);
    orc_grid2d_ind_frac_y(&g, y[i], iy + i, fy + i);
  }
}
void orc_flat_grid2d_grad(int nx, int ny, const double* l, const double* arr, long n,
                          const double* x, const double* y, double* ex, double* ey, double* val) {
  orc_grid2d g = mk2(nx, ny, l, (double*)arr);
  for (long i = 0; i < n; i++) {
    orc_grid2d_gradient(&g, x[i], y[i], ex + i, ey + i);
    val[i] = orc_grid2d_value(&g, x[i], y[i]);
  }
}
void orc_flat_grid1d_bin(int nz, double zmin, double zmax, long n, const double* z, const double* m,
                         double* out) {
  orc_grid1d g;
  g.nz = nz; g.arr = out;
  orc_grid1d_set_z(&g, zmin, zmax);
  memset(out, 0, sizeof(double) * nz);
  for (long i = 0; i < n; i++) orc_grid1d_bin_value(&g, m[i], z[i]);
}
void orc_flat_bunch_bin_axis(const orc_bunch* b, int nz, double zmin, double zmax, int axis,
                             int by_particle, double* out) {
  orc_grid1d g;
  g.nz = nz; g.arr = out;
  orc_grid1d_set_z(&g, zmin, zmax);
  memset(out, 0, sizeof(double) * nz);
  orc_grid1d_bin_axis(&g, b, axis, by_particle);
}
void orc_flat_grid1d_value(int nz, double zmin, double zmax, const double* arr, long n,
                           const double* z, double* val, double* grad) {
  orc_grid1d g;
  g.nz = nz; g.arr = (double*)arr;
  orc_grid1d_set_z(&g, zmin, zmax);
  for (long i = 0; i < n; i++) {
    val[i] = orc_grid1d_value(&g, z[i]);
    orc_grid1d_gradient(&g, z[i], grad + i);
  }
}
void orc_flat_grid3d_bin(int nx, int ny, int nz, const double* l, long n, const double* x,
                         const double* y, const double* z, const double* m, double* rho) {
  memset(rho, 0, sizeof(double) * (size_t)nx * ny * nz);
  orc_grid3d g = mk3(nx, ny, nz, l, rho);
  for (long i = 0; i < n; i++) orc_grid3d_bin_value(&g, m[i], x[i], y[i], z[i]);
}
void orc_flat_grid3d_indfrac(int nx, int ny, int nz, const double* l, long n, const double* x,
                             const double* y, const double* z, int* ind, double* frac) {
  orc_grid3d g = mk3(nx, ny, nz, l, NULL);
  for (long i = 0; i < n; i++) orc_grid3d_ind_frac(&g, x[i], y[i], z[i], ind + 3 * i, frac + 3 * i);
}
void orc_flat_grid3d_grad(int nx, int ny, int nz, const double* l, const double* arr, long n,
                          const double* x, const double* y, const double* z, double* gx,
                          double* gy, double* gz) {
  orc_grid3d g = mk3(nx, ny, nz, l, (double*)arr);
  for (long i = 0; i < n; i++) orc_grid3d_gradient(&g, x[i], y[i], z[i], gx + i, gy + i, gz + i);
}

void orc_flat_grid3d_bin_wrapped(int nx, int ny, int nz, const double* l, long n, const double* x,
                                 const double* y, const double* z, const double* m, double lambda,
                                 double* rho) {
  memset(rho, 0, sizeof(double) * (size_t)nx * ny * nz);
  orc_grid3d g = mk3(nx, ny, nz, l, rho);
  g.wrap = 1;
  for (long i = 0; i < n; i++) orc_grid3d_bin_value(&g, m[i], x[i], y[i], remainder(z[i], lambda));
}
void orc_flat_grid3d_grad_wrapped(int nx, int ny, int nz, const double* l, const double* arr,
                                  long n, const double* x, const double* y, const double* z,
                                  double* gx, double* gy, double* gz) {
  orc_grid3d g = mk3(nx, ny, nz, l, (double*)arr);
  g.wrap = 1;
  for (long i = 0; i < n; i++) orc_grid3d_gradient(&g, x[i], y[i], z[i], gx + i, gy + i, gz + i);
}

/* calculateLongDerivative on a given line density (SpaceChargeCalc2p5Drb.cc:257-310). */
void orc_flat_long_derivative(int nz, double zmin, double zmax, int n_avg, const double* zg,
                              double* zd) {
  orc_sc2p5drb c;
  memset(&c, 0, sizeof(c));
  c.n_long_avg = n_avg;
  c.zg.nz = nz; c.zg.arr = (double*)zg;
  c.zd.nz = nz; c.zd.arr = zd;
  orc_grid1d_set_z(&c.zg, zmin, zmax);
  orc_grid1d_set_z(&c.zd, zmin, zmax);
  calc_long_derivative(&c);
}

/* Direct Poisson solves on a given rho grid (solver built on solver_lims, rho/phi on
 * rho_lims), for brute-force checks of the Hockney convolution and its rescaling. */
void orc_flat_poisson2d(int nx, int ny, const double* sl, const double* rl, const double* rho,
                        double* phi) {
  orc_poisson2d* p = orc_poisson2d_create(nx, ny, sl[0], sl[1], sl[2], sl[3]);
  orc_grid2d r = mk2(nx, ny, rl, (double*)rho), f = mk2(nx, ny, rl, phi);
  orc_poisson2d_find_potential(p, &r, &f);
  orc_poisson2d_destroy(p);
}
void orc_flat_poisson3d(int nx, int ny, int nz, const double* sl, const double* rl,
                        const double* rho, double* phi) {
  orc_poisson3d* p = orc_poisson3d_create(nx, ny, nz, sl[0], sl[1], sl[2], sl[3], sl[4], sl[5]);
  orc_grid3d r = mk3(nx, ny, nz, rl, (double*)rho), f = mk3(nx, ny, nz, rl, phi);
  orc_poisson3d_find_potential(p, &r, &f);
  orc_poisson3d_destroy(p);
}
void orc_flat_poisson3d_images(int nx, int ny, int nz, const double* sl, const double* rl,
                               int nbunches, double lambda, const double* rho, double* phi) {
  orc_poisson3d* p = orc_poisson3d_create(nx, ny, nz, sl[0], sl[1], sl[2], sl[3], sl[4], sl[5]);
  orc_poisson3d_set_ext_bunches(p, nbunches);
  orc_poisson3d_set_ext_spacing(p, lambda);
  orc_poisson3d_set_grid(p, sl[0], sl[1], sl[2], sl[3], sl[4], sl[5]);
  orc_grid3d r = mk3(nx, ny, nz, rl, (double*)rho), f = mk3(nx, ny, nz, rl, phi);
  orc_poisson3d_find_potential(p, &r, &f);
  orc_poisson3d_destroy(p);
}
