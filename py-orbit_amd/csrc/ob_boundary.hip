// ob_boundary.hip -- Boundary2D (BaseBoundary2D.cc, ShapedBoundary2D.cc) for
// SpaceChargeCalc2p5D.
//
// A conducting boundary is a set of points; the free-space potential is corrected by the
// harmonic field sum_k c_k f_k(x, y) (f_0 = 1, f_2k+1/2k+2 = (r/R)^k cos/sin k phi) whose
// coefficients are the least-squares fit of phi on the boundary points, so that phi
// minus that field is (nearly) constant on the boundary.  The projector
// LSQ = (F^T F)^-1 F^T is set up once on the host (the reference's initializeBPs, a
// (2M+1)^2 Gauss-Jordan inverse); every step then runs two kernels on the device:
//   bnd_coef   one workgroup: phi interpolated at the points (Grid2D::getValue), then the
//              2M+1 coefficients, each a dot product in point order;
//   bnd_apply  one thread per node: the harmonic field from the coefficients, subtracted.
#include <cmath>
#include <string>
#include <utility>
#include <vector>

#include "ob_internal.hpp"
#include "ob_sc_device.hpp"

struct ob_boundary2d {
  int np = 0, nm = 0, type = 1;  // type: 1 point set, 2 circle, 3 ellipse, 4 rectangle
  bool shaped = false;           // points fixed by the shape
  bool initialized = false;
  std::string shape = "no shape";
  double xDim = 0., yDim = 0., p0 = 0., p1 = 0.;
  double xMin = DBL_MAX, xMax = -DBL_MAX, yMin = DBL_MAX, yMax = -DBL_MAX, rnorm = DBL_MAX;
  std::vector<double> bx, by, lsq;  // lsq: [(2 nm + 1)][np]
  double *d_bx = nullptr, *d_by = nullptr, *d_lsq = nullptr, *d_work = nullptr;
  double* d_work3 = nullptr;  // per-slice work of a stack of grids
  int64_t work3_cap = 0;
  ob::SC2DState* d_geo = nullptr;  // grid geometry of a user grid (add_potential)
  int d_np = 0;
};

namespace ob {
namespace {

constexpr int kCoefThreads = 256;

// Grid2D::getValue (Grid2D.cc:82-108).
__device__ __forceinline__ double grid2d_value(const double* __restrict__ a, int nx, int ny,
                                               double xMin, double dx, double yMin, double dy,
                                               double x, double y) {
  int iX, iY;
  double xF, yF;
  node_ind_frac(x, xMin, dx, nx, iX, xF);
  node_ind_frac(y, yMin, dy, ny, iY, yF);
  const double xF2 = xF * xF, yF2 = yF * yF;
  const double Wxm = 0.5 * (0.25 - xF + xF2), Wx0 = 0.75 - xF2, Wxp = 0.5 * (0.25 + xF + xF2);
  const double Wym = 0.5 * (0.25 - yF + yF2), Wy0 = 0.75 - yF2, Wyp = 0.5 * (0.25 + yF + yF2);
  const double* r0 = a + (int64_t)(iX - 1) * ny + iY;
  const double* r1 = a + (int64_t)iX * ny + iY;
  const double* r2 = a + (int64_t)(iX + 1) * ny + iY;
  return Wxm * Wym * r0[-1] + Wxm * Wy0 * r0[0] + Wxm * Wyp * r0[1] + Wx0 * Wym * r1[-1] +
         Wx0 * Wy0 * r1[0] + Wx0 * Wyp * r1[1] + Wxp * Wym * r2[-1] + Wxp * Wy0 * r2[0] +
         Wxp * Wyp * r2[1];
}

// work = [np] phi at the points, then [2 nm + 1] coefficients (BaseBoundary2D.cc:430-440).
__global__ __launch_bounds__(kCoefThreads) void bnd_coef(const double* __restrict__ phi, int nx,
                                                        int ny, const SC2DState* st,
                                                        const double* __restrict__ bx,
                                                        const double* __restrict__ by, int np,
                                                        const double* __restrict__ lsq, int nf,
                                                        double* __restrict__ work) {
  if (st->skip) return;
  const double xMin = st->xMin, dx = st->dx, yMin = st->yMin, dy = st->dy;
  for (int j = threadIdx.x; j < np; j += blockDim.x)
    work[j] = grid2d_value(phi, nx, ny, xMin, dx, yMin, dy, bx[j], by[j]);
  __syncthreads();
  for (int i = threadIdx.x; i < nf; i += blockDim.x) {
    double c = 0.0;
    const double* row = lsq + (size_t)i * np;
    for (int j = 0; j < np; j++) c += row[j] * work[j];
    work[np + i] = c;
  }
}

// phi -= sum_k c_k f_k at every node (BaseBoundary2D.cc:442-457, functions :316-345).
__global__ __launch_bounds__(256) void bnd_apply(double* __restrict__ phi, int nx, int ny,
                                                 const SC2DState* st,
                                                 const double* __restrict__ coef, int nm,
                                                 double rnorm) {
  if (st->skip) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nx * ny) return;
  const int i = t / ny, j = t - i * ny;
  const double x = st->xMin + i * st->dx, y = st->yMin + j * st->dy;
  const double r2 = x * x + y * y;
  double v = 0.;
  if (r2 == 0.) {
    v += coef[0] * 1.0;
    for (int k = 1; k < 2 * nm + 1; k++) v += coef[k] * 0.0;
  } else {
    const double r = sqrt(r2);
    const double sn = y / r, cs = x / r, q = r / rnorm;
    double snk = sn, csk = cs, rk = 1.0;
    v += coef[0] * 1.;
    for (int k = 0; k < nm; k++) {
      rk *= q;
      v += coef[2 * k + 1] * (rk * csk);
      v += coef[2 * k + 2] * (rk * snk);
      const double s1 = sn * csk + cs * snk;
      const double c1 = cs * csk - sn * snk;
      snk = s1;
      csk = c1;
    }
  }
  phi[t] -= v;
}

// A stack of z-fastest slices phi[(ix*ny + iy)*nz + iz]: one workgroup per slice iz.
// work = [nz][np] phi at the points, then [nz][2 nm + 1] coefficients.
__global__ __launch_bounds__(kCoefThreads) void bnd_coef_slices(
    const double* __restrict__ phi, int nx, int ny, int nz, const SC2DState* st,
    const double* __restrict__ bx, const double* __restrict__ by, int np,
    const double* __restrict__ lsq, int nf, double* __restrict__ work) {
  if (st->skip) return;
  const int iz = blockIdx.x;
  const double xMin = st->xMin, dx = st->dx, yMin = st->yMin, dy = st->dy;
  double* bphi = work + (size_t)iz * np;
  for (int j = threadIdx.x; j < np; j += blockDim.x) {
    int iX, iY;
    double xF, yF;
    node_ind_frac(bx[j], xMin, dx, nx, iX, xF);
    node_ind_frac(by[j], yMin, dy, ny, iY, yF);
    const double xF2 = xF * xF, yF2 = yF * yF;
    const double Wxm = 0.5 * (0.25 - xF + xF2), Wx0 = 0.75 - xF2, Wxp = 0.5 * (0.25 + xF + xF2);
    const double Wym = 0.5 * (0.25 - yF + yF2), Wy0 = 0.75 - yF2, Wyp = 0.5 * (0.25 + yF + yF2);
    auto a = [&](int ix, int iy) { return phi[((int64_t)ix * ny + iy) * nz + iz]; };
    bphi[j] = Wxm * Wym * a(iX - 1, iY - 1) + Wxm * Wy0 * a(iX - 1, iY) +
              Wxm * Wyp * a(iX - 1, iY + 1) + Wx0 * Wym * a(iX, iY - 1) + Wx0 * Wy0 * a(iX, iY) +
              Wx0 * Wyp * a(iX, iY + 1) + Wxp * Wym * a(iX + 1, iY - 1) +
              Wxp * Wy0 * a(iX + 1, iY) + Wxp * Wyp * a(iX + 1, iY + 1);
  }
  __syncthreads();
  double* coef = work + (size_t)nz * np + (size_t)iz * nf;
  for (int i = threadIdx.x; i < nf; i += blockDim.x) {
    double c = 0.0;
    const double* row = lsq + (size_t)i * np;
    for (int j = 0; j < np; j++) c += row[j] * bphi[j];
    coef[i] = c;
  }
}

__global__ __launch_bounds__(256) void bnd_apply_slices(double* __restrict__ phi, int nx, int ny,
                                                        int nz, const SC2DState* st,
                                                        const double* __restrict__ coefs,
                                                        int nm, double rnorm) {
  if (st->skip) return;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nx * ny * nz) return;
  const int iz = (int)(t % nz);
  const int64_t r = t / nz;
  const int i = (int)(r / ny), j = (int)(r - (int64_t)i * ny);
  const double* coef = coefs + (size_t)iz * (2 * nm + 1);
  const double x = st->xMin + i * st->dx, y = st->yMin + j * st->dy;
  const double r2 = x * x + y * y;
  double v = 0.;
  if (r2 == 0.) {
    v += coef[0] * 1.0;
    for (int k = 1; k < 2 * nm + 1; k++) v += coef[k] * 0.0;
  } else {
    const double rr = sqrt(r2);
    const double sn = y / rr, cs = x / rr, q = rr / rnorm;
    double snk = sn, csk = cs, rk = 1.0;
    v += coef[0] * 1.;
    for (int k = 0; k < nm; k++) {
      rk *= q;
      v += coef[2 * k + 1] * (rk * csk);
      v += coef[2 * k + 2] * (rk * snk);
      const double s1 = sn * csk + cs * snk;
      const double c1 = cs * csk - sn * snk;
      snk = s1;
      csk = c1;
    }
  }
  phi[t] -= v;
}

// The mode functions at one point (BaseBoundary2D.cc:316-345), host side.
void functions(const ob_boundary2d* b, double x, double y, double* f) {
  const int nf = 2 * b->nm + 1;
  const double r2 = x * x + y * y;
  if (r2 == 0.) {
    f[0] = 1.0;
    for (int i = 1; i < nf; i++) f[i] = 0.0;
    return;
  }
  const double r = sqrt(r2);
  const double sn = y / r, cs = x / r, q = r / b->rnorm;
  double snk = sn, csk = cs, rk = 1.0;
  f[0] = 1.;
  for (int k = 0; k < b->nm; k++) {
    rk *= q;
    f[2 * k + 1] = rk * csk;
    f[2 * k + 2] = rk * snk;
    const double s1 = sn * csk + cs * snk;
    const double c1 = cs * csk - sn * snk;
    snk = s1;
    csk = c1;
  }
}

// In-place inverse of the n x n matrix a by Gauss-Jordan elimination with full pivoting
// (BaseBoundary2D.cc:233-313).
void gauss_jordan_inverse(std::vector<double>& a, int n) {
  std::vector<int> pivoted(n, 0), prow(n), pcol(n);
  for (int step = 0; step < n; step++) {
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
        } else {
          OB_REQUIRE(pivoted[k] <= 1, "Boundary2D: the least-squares matrix is singular");
        }
      }
    }
    pivoted[icol] += 1;
    if (irow != icol)
      for (int l = 0; l < n; l++) std::swap(a[irow * n + l], a[icol * n + l]);
    prow[step] = irow;
    pcol[step] = icol;
    OB_REQUIRE(a[icol * n + icol] != 0.0, "Boundary2D: the least-squares matrix is singular");
    const double inv = 1.0 / a[icol * n + icol];
    a[icol * n + icol] = 1.0;
    for (int l = 0; l < n; l++) a[icol * n + l] *= inv;
    for (int r = 0; r < n; r++) {
      if (r == icol) continue;
      const double d = a[r * n + icol];
      a[r * n + icol] = 0.0;
      for (int l = 0; l < n; l++) a[r * n + l] -= a[icol * n + l] * d;
    }
  }
  for (int l = n - 1; l >= 0; l--)
    if (prow[l] != pcol[l])
      for (int k = 0; k < n; k++) std::swap(a[k * n + prow[l]], a[k * n + pcol[l]]);
}

// initializeBPs (BaseBoundary2D.cc:188-230): limits, norm radius, LSQ = (F^T F)^-1 F^T;
// the points and the projector go to the device.
void initialize(ob_boundary2d* b) {
  const int nf = 2 * b->nm + 1;
  b->xMin = DBL_MAX; b->xMax = -DBL_MAX; b->yMin = DBL_MAX; b->yMax = -DBL_MAX;
  for (int i = 0; i < b->np; i++) {
    const double x = b->bx[i], y = b->by[i];
    if (b->xMin > x) b->xMin = x;
    if (b->xMax < x) b->xMax = x;
    if (b->yMin > y) b->yMin = y;
    if (b->yMax < y) b->yMax = y;
  }
  b->rnorm = sqrt((b->xMin - b->xMax) * (b->xMin - b->xMax) +
                  (b->yMin - b->yMax) * (b->yMin - b->yMax));
  std::vector<double> m((size_t)nf * nf, 0.0), f(nf);
  for (int p = 0; p < b->np; p++) {
    functions(b, b->bx[p], b->by[p], f.data());
    for (int i = 0; i < nf; i++)
      for (int j = 0; j < nf; j++) m[i * nf + j] += f[i] * f[j];
  }
  gauss_jordan_inverse(m, nf);
  b->lsq.assign((size_t)nf * b->np, 0.0);
  for (int i = 0; i < nf; i++)
    for (int j = 0; j < b->np; j++) {
      double acc = 0.0;
      functions(b, b->bx[j], b->by[j], f.data());
      for (int k = 0; k < nf; k++) acc += m[i * nf + k] * f[k];
      b->lsq[(size_t)i * b->np + j] = acc;
    }
  ensure_init();
  hipStream_t s = stream();
  OB_HIP(hipStreamSynchronize(s));  // earlier steps may still read the old arrays
  if (b->d_np != b->np) {
    for (void* p : {(void*)b->d_bx, (void*)b->d_by, (void*)b->d_lsq, (void*)b->d_work})
      if (p) OB_HIP(hipFree(p));
    OB_HIP(hipMalloc((void**)&b->d_bx, sizeof(double) * b->np));
    OB_HIP(hipMalloc((void**)&b->d_by, sizeof(double) * b->np));
    OB_HIP(hipMalloc((void**)&b->d_lsq, sizeof(double) * nf * b->np));
    OB_HIP(hipMalloc((void**)&b->d_work, sizeof(double) * (b->np + nf)));
    b->d_np = b->np;
  }
  OB_HIP(hipMemcpy(b->d_bx, b->bx.data(), sizeof(double) * b->np, hipMemcpyHostToDevice));
  OB_HIP(hipMemcpy(b->d_by, b->by.data(), sizeof(double) * b->np, hipMemcpyHostToDevice));
  OB_HIP(hipMemcpy(b->d_lsq, b->lsq.data(), sizeof(double) * nf * b->np,
                   hipMemcpyHostToDevice));
  b->initialized = true;
}

void set_point(ob_boundary2d* b, int i, double x, double y) {
  if (i >= 0 && i < b->np) {
    b->bx[i] = x;
    b->by[i] = y;
  }
  b->initialized = false;
}
}  // namespace

void boundary_add_potential(ob_boundary2d* b, double* phi, int nx, int ny, const SC2DState* st,
                            hipStream_t s) {
  OB_REQUIRE(b->initialized,
             "Boundary2D addBoundaryPotential: the boundary was not initialized");
  const int nf = 2 * b->nm + 1;
  {
    ProfScope ps("bnd_coef", s);
    hipLaunchKernelGGL(bnd_coef, dim3(1), dim3(kCoefThreads), 0, s, phi, nx, ny, st, b->d_bx,
                       b->d_by, b->np, b->d_lsq, nf, b->d_work);
  }
  {
    ProfScope ps("bnd_apply", s);
    hipLaunchKernelGGL(bnd_apply, dim3((nx * ny + 255) / 256), dim3(256), 0, s, phi, nx, ny, st,
                       b->d_work + b->np, b->nm, b->rnorm);
  }
  OB_HIP(hipGetLastError());
}

void boundary_add_potential_slices(ob_boundary2d* b, double* phi, int nx, int ny, int nz,
                                   const SC2DState* st, hipStream_t s) {
  OB_REQUIRE(b->initialized,
             "Boundary2D addBoundaryPotential: the boundary was not initialized");
  const int nf = 2 * b->nm + 1;
  const int64_t need = (int64_t)nz * (b->np + nf);
  if (need > b->work3_cap) {
    OB_HIP(hipStreamSynchronize(s));
    if (b->d_work3) OB_HIP(hipFree(b->d_work3));
    OB_HIP(hipMalloc((void**)&b->d_work3, sizeof(double) * need));
    b->work3_cap = need;
  }
  {
    ProfScope ps("bnd_coef_slices", s);
    hipLaunchKernelGGL(bnd_coef_slices, dim3(nz), dim3(kCoefThreads), 0, s, phi, nx, ny, nz, st,
                       b->d_bx, b->d_by, b->np, b->d_lsq, nf, b->d_work3);
  }
  {
    ProfScope ps("bnd_apply_slices", s);
    const int64_t tot = (int64_t)nx * ny * nz;
    hipLaunchKernelGGL(bnd_apply_slices, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s,
                       phi, nx, ny, nz, st, b->d_work3 + (size_t)nz * b->np, b->nm, b->rnorm);
  }
  OB_HIP(hipGetLastError());
}

int boundary_describe(ob_boundary2d* b, double* bnd6) {
  OB_REQUIRE(b->initialized, "Boundary2D: the boundary was not initialized");
  bnd6[0] = b->xMin; bnd6[1] = b->xMax; bnd6[2] = b->yMin; bnd6[3] = b->yMax;
  bnd6[4] = b->p0;
  bnd6[5] = b->p1;
  return b->type;
}
}  // namespace ob

#define OB_API extern "C" __attribute__((visibility("default")))
#define OB_TRY try {
#define OB_CATCH               \
  }                            \
  catch (const ob::Error& e) { \
    ob::set_error(e.msg);      \
    return 1;                  \
  }                            \
  return 0;

// Boundary2D(nPoints, nModes) (BaseBoundary2D.cc:22-95): points to be set one by one.
OB_API int ob_boundary2d_create(int nPoints, int nModes, ob_boundary2d** out) {
  OB_TRY OB_REQUIRE(out != nullptr, "null output pointer");
  OB_REQUIRE(nModes >= 0, "Boundary2D: the number of modes must be >= 0");
  OB_REQUIRE(2 * nModes + 1 < nPoints,
             "Boundary2D: 2*nModes+1 must be below the number of points");
  ob_boundary2d* b = new ob_boundary2d();
  b->np = nPoints;
  b->nm = nModes;
  b->bx.assign(nPoints, DBL_MAX);
  b->by.assign(nPoints, DBL_MAX);
  *out = b;
  OB_CATCH
}

// Boundary2D(nPoints, nModes, shape, xDim, yDim) (ShapedBoundary2D.cc:9-127): "Circle"
// (diameter xDim = yDim), "Ellipse" (axes xDim, yDim) or "Rectangle" (sides xDim, yDim);
// nPoints is rounded down to a multiple of 4.
OB_API int ob_boundary2d_create_shaped(int nPoints, int nModes, const char* shape, double xDim,
                                       double yDim, ob_boundary2d** out) {
  OB_TRY OB_REQUIRE(out != nullptr && shape != nullptr, "null argument");
  const std::string sh(shape);
  int type = 0;
  if (sh == "Circle") type = 2;
  if (sh == "Ellipse") type = 3;
  if (sh == "Rectangle") type = 4;
  OB_REQUIRE(type != 0, "Boundary2D: the shape should be Circle or Ellipse or Rectangle");
  OB_REQUIRE(type != 2 || xDim == yDim,
             "Boundary2D: the Circle should have the same x and y sizes");
  const int np = 4 * (nPoints / 4);
  ob_boundary2d* b = nullptr;
  const int rc = ob_boundary2d_create(np, nModes, &b);
  if (rc != 0) return rc;
  b->type = type;
  b->shape = sh;
  b->xDim = xDim;
  b->yDim = yDim;
  const double PI = 3.14159265358979323846;
  if (type == 2) {
    const double r = xDim / 2;
    b->p0 = r;
    // sin and cos of one argument through sincos, as g++ combines them in the reference
    // build (ShapedBoundary2D.cc:47-51); its last bit can differ from sin's
    for (int i = 0; i < np; i++) {
      double sn, cs;
      sincos(2 * i * PI / np, &sn, &cs);
      ob::set_point(b, i, r * sn, r * cs);
    }
  } else if (type == 3) {
    // points of equal arc length: the first quadrant's steps are rescaled until they span
    // exactly pi/2, then mirrored
    const double a = xDim / 2, bb = yDim / 2;
    b->p0 = a;
    b->p1 = bb;
    double ds = 2 * PI * sqrt(a * bb) / np, resid = 1.0;
    std::vector<double> th(np + 1, 0.0);
    while (resid > 1.0e-08 || resid < -1.0e-08) {
      double denom = bb;
      double dth = ds / denom;
      double t = dth / 2.;
      th[0] = 0.;
      for (int i = 0; i < np / 4; i++) {
        denom = bb * bb + (a * a - bb * bb) * sin(t) * sin(t);
        denom = sqrt(denom);
        dth = ds / denom;
        th[i + 1] = th[i] + dth;
        t += dth;
      }
      resid = th[np / 4] - PI / 2.;
      ds *= PI / (2. * th[np / 4]);
    }
    for (int i = 0; i < np / 4; i++) th[np / 2 - i] = PI - th[i];
    for (int i = 0; i < np / 2; i++) th[np / 2 + i] = PI + th[i];
    for (int i = 0; i < np; i++) {
      double sn, cs;
      sincos(th[i], &sn, &cs);
      ob::set_point(b, i, a * cs, bb * sn);
    }
  } else {
    const double a = xDim, bb = yDim;
    b->p0 = a;
    b->p1 = bb;
    const double dx = a / (np / 4), dy = bb / (np / 4);
    int k = 0;
    for (int i = 0; i < np / 4; i++) ob::set_point(b, k++, (-a / 2.0) + i * dx, bb / 2.0);
    for (int i = 0; i < np / 4; i++) ob::set_point(b, k++, a / 2.0, bb / 2.0 - i * dy);
    for (int i = 0; i < np / 4; i++) ob::set_point(b, k++, a / 2.0 - i * dx, -bb / 2.0);
    for (int i = 0; i < np / 4; i++) ob::set_point(b, k++, -a / 2.0, (-bb / 2.0) + i * dy);
  }
  try {
    ob::initialize(b);
  } catch (...) {
    delete b;
    throw;
  }
  b->shaped = true;
  *out = b;
  OB_CATCH
}

OB_API int ob_boundary2d_destroy(ob_boundary2d* b) {
  OB_TRY if (!b) return 0;
  (void)hipStreamSynchronize(ob::stream());  // queued steps may read its arrays
  for (void* p : {(void*)b->d_bx, (void*)b->d_by, (void*)b->d_lsq, (void*)b->d_work,
                  (void*)b->d_geo, (void*)b->d_work3})
    if (p) (void)hipFree(p);
  delete b;
  OB_CATCH
}

// setBoundaryPoint (BaseBoundary2D.cc:160-184): refused once a shape fixed the points.
OB_API int ob_boundary2d_set_point(ob_boundary2d* b, int i, double x, double y) {
  OB_TRY OB_REQUIRE(b != nullptr, "null handle");
  OB_REQUIRE(!b->shaped, "Boundary2D setBoundaryPoint: the shape has been defined already");
  ob::set_point(b, i, x, y);
  OB_CATCH
}

OB_API int ob_boundary2d_get_point(ob_boundary2d* b, int i, double* x, double* y) {
  OB_TRY OB_REQUIRE(b != nullptr && x && y, "null handle");
  *x = (i >= 0 && i < b->np) ? b->bx[i] : DBL_MAX;
  *y = (i >= 0 && i < b->np) ? b->by[i] : DBL_MAX;
  OB_CATCH
}

OB_API int ob_boundary2d_initialize(ob_boundary2d* b) {
  OB_TRY OB_REQUIRE(b != nullptr, "null handle");
  ob::initialize(b);
  OB_CATCH
}

// out[4] = xMin, xMax, yMin, yMax; info[4] = nPoints, nModes, initialized, shape type
// (0 no shape, 1 Circle, 2 Ellipse, 3 Rectangle).
OB_API int ob_boundary2d_info(ob_boundary2d* b, double* lims, int* info) {
  OB_TRY OB_REQUIRE(b != nullptr && lims && info, "null handle");
  lims[0] = b->xMin; lims[1] = b->xMax; lims[2] = b->yMin; lims[3] = b->yMax;
  info[0] = b->np;
  info[1] = b->nm;
  info[2] = b->initialized ? 1 : 0;
  info[3] = b->type - 1;
  OB_CATCH
}

// isInside(x, y): +1 inside, -1 outside (BaseBoundary2D.cc:139-147, ShapedBoundary2D.cc:284-304).
OB_API int ob_boundary2d_is_inside(ob_boundary2d* b, double x, double y, int* out) {
  OB_TRY OB_REQUIRE(b != nullptr && out != nullptr, "null handle");
  double d[6] = {b->xMin, b->xMax, b->yMin, b->yMax, b->p0, b->p1};
  *out = ob::boundary_inside(b->type, d, x, y) ? 1 : -1;
  OB_CATCH
}

// addBoundaryPotential(rhoGrid, phiGrid) (BaseBoundary2D.cc:349-458) on a host phi grid
// [nx][ny] spanning lims[4], computed on the device; the boundary must lie inside it.
OB_API int ob_boundary2d_add_potential(ob_boundary2d* b, double* phi, int nx, int ny,
                                       const double* lims) {
  OB_TRY OB_REQUIRE(b != nullptr && phi != nullptr && lims != nullptr, "null handle");
  OB_REQUIRE(b->initialized,
             "Boundary2D addBoundaryPotential: the boundary was not initialized");
  OB_REQUIRE(nx >= 3 && ny >= 3, "Boundary2D addBoundaryPotential: grid too small");
  OB_REQUIRE(!(lims[0] > b->xMin || lims[1] < b->xMax || lims[2] > b->yMin || lims[3] < b->yMax),
             "Boundary2D addBoundaryPotential: some boundary points are outside the grid");
  ob::ensure_init();
  hipStream_t s = ob::stream();
  if (!b->d_geo) OB_HIP(hipMalloc((void**)&b->d_geo, sizeof(ob::SC2DState)));
  ob::SC2DState g;
  std::memset(&g, 0, sizeof(g));
  g.xMin = lims[0]; g.xMax = lims[1]; g.yMin = lims[2]; g.yMax = lims[3];
  g.dx = (lims[1] - lims[0]) / (nx - 1);  // Grid2D.cc:409-420
  g.dy = (lims[3] - lims[2]) / (ny - 1);
  double* d_phi = nullptr;
  OB_HIP(hipMalloc((void**)&d_phi, sizeof(double) * nx * ny));
  try {
    OB_HIP(hipMemcpyAsync(b->d_geo, &g, sizeof(g), hipMemcpyHostToDevice, s));
    OB_HIP(hipMemcpyAsync(d_phi, phi, sizeof(double) * nx * ny, hipMemcpyHostToDevice, s));
    ob::boundary_add_potential(b, d_phi, nx, ny, b->d_geo, s);
    OB_HIP(hipMemcpyAsync(phi, d_phi, sizeof(double) * nx * ny, hipMemcpyDeviceToHost, s));
    OB_HIP(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipFree(d_phi);
    throw;
  }
  OB_HIP(hipFree(d_phi));
  OB_CATCH
}
