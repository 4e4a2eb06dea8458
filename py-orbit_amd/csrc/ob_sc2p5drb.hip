// ob_sc2p5drb.hip -- SpaceChargeCalc2p5Drb on the device.
//
// Per trackBunch (SpaceChargeCalc2p5Drb.cc:110-168), all on the library stream with no
// host round trip:
//   1. fused map kernel flushes the bunch's queued maps and emits per-block extrema;
//   2. setup   (1 block): reduce extrema, size the grids (bunchAnalysis :170-213), clear;
//   3. deposit          : 9-point TSC charge (Grid2D.cc:120-166) with FP64 atomics, and the
//                         periodic 2-point z line density (Grid1D.cc:241-275, 568-626)
//                         accumulated per block in LDS;
//   4. post    (1 block): dlambda/dz by 3-point LSQ (:257-310), rho moments (:229-254),
//                         kick factors;
//   5. Poisson (3 kernels, PoissonSolverFFT2D.cc:174-256): row r2c of two rows per complex
//      FFT, column FFT x Green x inverse column FFT (zero-padding pruned: only the nx
//      non-zero rows are read and only the nx needed rows written), row c2r;
//   6. the kick (:146-167) is queued on the bunch and runs fused into the next map kernel.
#include <cmath>
#include <cstring>

#include "ob_fft.hpp"
#include "ob_internal.hpp"
#include "ob_sc_device.hpp"

namespace ob {

double2* make_twiddles(int nmax) {
  std::vector<double2> h(nmax);
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int k = 0; k < nmax; k++) {
    long double ang = 2.0L * pi * (long double)k / (long double)nmax;
    h[k].x = (double)cosl(ang);
    h[k].y = (double)(-sinl(ang));
  }
  double2* d = nullptr;
  OB_HIP(hipMalloc((void**)&d, sizeof(double2) * h.size()));
  OB_HIP(hipMemcpy(d, h.data(), sizeof(double2) * h.size(), hipMemcpyHostToDevice));
  return d;
}

namespace {

constexpr int kSetupThreads = 1024;
constexpr int kDepositThreads = 256;
constexpr int kMaxZBinsLDS = 4096;
constexpr int kFFTThreads = 256;

struct SC2DParams {
  int nx, ny, nz, nblk;       // grid sizes, number of extrema partial blocks
  double xy_ratio;
  double dx_solver;           // solver grid step (fixed at construction, :33)
  double macrosize;           // bunch macroSize (used when no per-particle attribute)
  double factor0;             // 2 L r0 / (beta^2 gamma^3)            (:129)
  double pipe_radius;
  int n_long_avg;
};

// ---- 2: setup -------------------------------------------------------------
// partials[pb..pe) hold per-block extrema of this bunch.  One block of kSetupThreads.
__device__ __forceinline__ void sc2d_setup_body(const SC2DParams& P,
                                                const double* __restrict__ partials, int pb,
                                                int pe, SC2DState* st, double* rho, double* zg,
                                                double* zd) {
  __shared__ double red[kSetupThreads / 64][6];
  double r[6] = {DBL_MAX, -DBL_MAX, DBL_MAX, -DBL_MAX, DBL_MAX, -DBL_MAX};
  for (int b = pb + threadIdx.x; b < pe; b += blockDim.x)
    for (int k = 0; k < 6; k++) {
      const double v = partials[b * 6 + k];
      r[k] = (k & 1) ? fmax(r[k], v) : fmin(r[k], v);
    }
  for (int k = 0; k < 6; k++)
    for (int o = 32; o > 0; o >>= 1) {
      const double v = __shfl_xor(r[k], o, 64);
      r[k] = (k & 1) ? fmax(r[k], v) : fmin(r[k], v);
    }
  if ((threadIdx.x & 63) == 0)
    for (int k = 0; k < 6; k++) red[threadIdx.x >> 6][k] = r[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kSetupThreads / 64; w++)
      for (int k = 0; k < 6; k++) r[k] = (k & 1) ? fmax(r[k], red[w][k]) : fmin(r[k], red[w][k]);
    for (int k = 0; k < 6; k++) st->ext[k] = r[k];
    double xMin = r[0], xMax = r[1], yMin = r[2], yMax = r[3];
    const double zMin = r[4], zMax = r[5];
    st->status = 0;
    st->skip = 0;
    if (xMin >= xMax || yMin >= yMax || zMin >= zMax) {  // :176-189 (reference aborts)
      st->status = 1;
      st->skip = 1;
    } else {  // :191-203
      const double xy_ratio_beam = (xMax - xMin) / (yMax - yMin);
      double width, center;
      if (xy_ratio_beam > P.xy_ratio) {
        center = (yMax + yMin) / 2.0;
        width = ((yMax - yMin) * (xy_ratio_beam / P.xy_ratio)) / 2.0;
        yMin = center - width;
        yMax = center + width;
      } else {
        center = (xMax + xMin) / 2.0;
        width = ((xMax - xMin) / (xy_ratio_beam / P.xy_ratio)) / 2.0;
        xMin = center - width;
        xMax = center + width;
      }
      st->xMin = xMin;
      st->xMax = xMax;
      st->dx = (xMax - xMin) / (P.nx - 1);  // Grid2D.cc:409-413
      st->yMin = yMin;
      st->yMax = yMax;
      st->dy = (yMax - yMin) / (P.ny - 1);
      st->zMin = zMin;
      st->zMax = zMax;
      st->dz = (zMax - zMin) / P.nz;        // Grid1D.cc:88-93
    }
  }
  const int ng = P.nx * P.ny;
  for (int i = threadIdx.x; i < ng; i += blockDim.x) rho[i] = 0.0;
  for (int i = threadIdx.x; i < P.nz; i += blockDim.x) {
    zg[i] = 0.0;
    zd[i] = 0.0;
  }
}

__global__ __launch_bounds__(kSetupThreads) void sc2d_setup(SC2DParams P,
                                                            const double* __restrict__ partials,
                                                            SC2DState* st, double* rho,
                                                            double* zg, double* zd) {
  sc2d_setup_body(P, partials, 0, P.nblk, st, rho, zg, zd);
}

// ---- 3: deposit -----------------------------------------------------------
// Particles first, first+stride, ... < end of one bunch; the block's z line density is
// summed in LDS and added to zg at the end.
__device__ __forceinline__ void sc2d_deposit_body(const SC2DParams& P, const DeviceBunchView& v,
                                                  const SC2DState* st, double* __restrict__ rho,
                                                  double* __restrict__ zg, int64_t first,
                                                  int64_t end, int64_t stride) {
  __shared__ double zl[kMaxZBinsLDS];
  const int nz = P.nz;
  for (int i = threadIdx.x; i < nz; i += blockDim.x) zl[i] = 0.0;
  __syncthreads();
  if (!st->skip) {
    const double xMin = st->xMin, xMax = st->xMax, yMin = st->yMin, yMax = st->yMax;
    const double dx = st->dx, dy = st->dy, zMin = st->zMin, zMax = st->zMax, dz = st->dz;
    const int nx = P.nx, ny = P.ny;
    for (int64_t i = first; i < end; i += stride) {
      const double x = v.c[0][i], y = v.c[2][i], z = v.c[4][i];
      const double m = v.msize ? v.msize[i] : P.macrosize;
      // Grid2D::binValue (Grid2D.cc:142-166)
      if (!(x < xMin || x > xMax || y < yMin || y > yMax)) {
        int iX, iY;
        double xF, yF;
        node_ind_frac(x, xMin, dx, nx, iX, xF);
        node_ind_frac(y, yMin, dy, ny, iY, yF);
        const double xF2 = xF * xF, yF2 = yF * yF;
        const double wx[3] = {0.5 * (0.25 - xF + xF2), 0.75 - xF2, 0.5 * (0.25 + xF + xF2)};
        const double wy[3] = {0.5 * (0.25 - yF + yF2), 0.75 - yF2, 0.5 * (0.25 + yF + yF2)};
        double* base = rho + (int64_t)(iX - 1) * ny + (iY - 1);
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
          for (int c = 0; c < 3; c++) atomicAdd(base + a * ny + c, wx[a] * wy[c] * m);
      }
      // Grid1D::binValue (Grid1D.cc:422-440)
      if (!(z < zMin || z > zMax)) {
        int i0, ip;
        double w0, wp;
        grid1d_ind_wz(z, zMin, dz, nz, i0, ip, w0, wp);
        if (nz > 1) {
          atomicAdd(&zl[i0], w0 * m);
          atomicAdd(&zl[ip], wp * m);
        } else {
          atomicAdd(&zl[0], m);
        }
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nz; i += blockDim.x)
    if (zl[i] != 0.0) atomicAdd(&zg[i], zl[i]);
}

__global__ __launch_bounds__(kDepositThreads) void sc2d_deposit(SC2DParams P, DeviceBunchView v,
                                                                const SC2DState* st,
                                                                double* __restrict__ rho,
                                                                double* __restrict__ zg) {
  sc2d_deposit_body(P, v, st, rho, zg, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, v.n,
                    (int64_t)gridDim.x * blockDim.x);
}

// ---- 4: post (one block) ---------------------------------------------------
__device__ double block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += sh[w];
  return s;  // valid in thread 0
}

__device__ __forceinline__ double ipow(double x, int j) {  // exact for the small integers used
  double r = 1.0;
  for (int k = 0; k < j; k++) r *= x;
  return r;
}

__device__ __forceinline__ void sc2d_post_body(const SC2DParams& P, SC2DState* st,
                                               const double* __restrict__ rho,
                                               const double* __restrict__ zg,
                                               double* __restrict__ zd) {
  __shared__ double sh[kSetupThreads / 64];
  if (st->skip) return;
  const int nz = P.nz;
  const double z_step = st->dz;
  const double zMin = st->zMin;
  // calculateLongDerivative (:257-310)
  if (nz == 1) {
    if (threadIdx.x == 0) zd[0] = 0.;
  } else if (nz == 2) {
    if (threadIdx.x == 0) {
      // Grid1D::calcGradient at the mid point (Grid1D.cc:476-489)
      const double zmid = ((zMin + (0 + 0.5) * z_step) + (zMin + (1 + 0.5) * z_step)) * 0.5;
      double zDeriv = 0.;
      if (!(zmid < zMin || zmid > st->zMax)) {
        int i0, ip;
        double w0, wp;
        grid1d_ind_wz(zmid, zMin, z_step, nz, i0, ip, w0, wp);
        zDeriv = (zg[ip] - zg[i0]) / z_step;
      }
      zd[0] = zDeriv / z_step;
      zd[1] = zDeriv / z_step;
    }
  } else {
    int nAvg = P.n_long_avg;
    if (nz < nAvg) nAvg = nz;
    for (int iz = threadIdx.x; iz < nz; iz += blockDim.x) {
      const double z = zMin + (iz + 0.5) * z_step;
      int iStart = iz - nAvg / 2;
      if (iStart < 0) iStart = 0;
      if ((iStart + nAvg) >= nz) iStart = nz - nAvg;
      const int iStop = iStart + nAvg;
      double S[5][2];
      for (int j = 0; j < 5; j++)
        for (int k = 0; k < 2; k++) {
          double acc = 0.;
          for (int i = iStart; i < iStop; i++) {
            const double x = 1.0 * (i - iStart + 1);
            const double y = zg[i];
            acc += ipow(x, j) * ipow(y, k);
          }
          S[j][k] = acc;
        }
      const double det = (S[0][0] * S[2][0] * S[4][0] - S[1][0] * S[1][0] * S[4][0] -
                          S[0][0] * S[3][0] * S[3][0] + 2 * S[1][0] * S[2][0] * S[3][0] -
                          S[2][0] * S[2][0] * S[2][0]);
      const double a =
          (S[0][1] * S[1][0] * S[3][0] - S[1][1] * S[0][0] * S[3][0] - S[0][1] * S[2][0] * S[2][0] +
           S[1][1] * S[1][0] * S[2][0] + S[2][1] * S[0][0] * S[2][0] - S[2][1] * S[1][0] * S[1][0]) /
          det;
      const double b =
          (S[1][1] * S[0][0] * S[4][0] - S[0][1] * S[1][0] * S[4][0] + S[0][1] * S[2][0] * S[3][0] -
           S[2][1] * S[0][0] * S[3][0] - S[1][1] * S[2][0] * S[2][0] + S[2][1] * S[1][0] * S[2][0]) /
          det;
      double y = 2 * a * (z - (zMin + (iStart + 0.5) * z_step) + z_step) / z_step + b;
      y = y / z_step;
      zd[iz] = y / z_step;
    }
  }
  // rho moments (:229-254); quirk Q1: the y position uses getGridX(iy) (:238)
  const int nx = P.nx, ny = P.ny;
  const double gx0 = st->xMin, gdx = st->dx;
  double s0 = 0., s1 = 0., s2 = 0., s3 = 0., s4 = 0.;
  for (int k = threadIdx.x; k < nx * ny; k += blockDim.x) {
    const int ix = k / ny, iy = k - ix * ny;
    const double val = rho[k];
    const double x = gx0 + ix * gdx;
    const double y = gx0 + iy * gdx;
    s0 += val;
    s1 += x * val;
    s2 += x * x * val;
    s3 += y * val;
    s4 += y * y * val;
  }
  s0 = block_sum(s0, sh);
  s1 = block_sum(s1, sh);
  s2 = block_sum(s2, sh);
  s3 = block_sum(s3, sh);
  s4 = block_sum(s4, sh);
  if (threadIdx.x == 0) {
    const double total = s0;
    double x_avg = s1 / total, x2_avg = s2 / total, y_avg = s3 / total, y2_avg = s4 / total;
    st->x_c = x_avg;
    st->y_c = y_avg;
    x2_avg = fabs(x2_avg - x_avg * x_avg);
    y2_avg = fabs(y2_avg - y_avg * y_avg);
    const double a_bunch = sqrt(2 * (x2_avg + y2_avg));
    st->a_bunch = a_bunch;
    st->total = total;
    st->factor = P.factor0 / (z_step * total);                 // :133
    st->lsf_in = 1.0 + 2 * log(P.pipe_radius / a_bunch);       // :138
    st->a2 = a_bunch * a_bunch;                                // :140
    st->scale_coeff = log(P.dx_solver / st->dx);               // PoissonSolverFFT2D.cc:216
  }
}

__global__ __launch_bounds__(kSetupThreads) void sc2d_post(SC2DParams P, SC2DState* st,
                                                           const double* __restrict__ rho,
                                                           const double* __restrict__ zg,
                                                           double* __restrict__ zd) {
  sc2d_post_body(P, st, rho, zg, zd);
}

// ---- 5: Poisson solve --------------------------------------------------------
// (a) two real rows per complex FFT along y, r2c split.  rows >= nrows_in are zero;
// columns >= ncols_in are zero (padding).  out: spectrum [*][H] complex (H = N/2+1).
// blk: row pair (rows 2 blk, 2 blk + 1).
template <int N>
__device__ __forceinline__ void rows_r2c_body(const double* __restrict__ in, int nrows_in,
                                              int ncols_in, int ld_in,
                                              double2* __restrict__ out,
                                              const double2* __restrict__ tw, int nmax,
                                              const SC2DState* st, int blk) {
  if (st && st->skip) return;
  extern __shared__ double sm[];
  double* re = sm;
  double* im = sm + N;
  const int r0 = 2 * blk, r1 = r0 + 1;
  for (int j = threadIdx.x; j < N; j += kFFTThreads) {
    re[j] = (j < ncols_in) ? in[(int64_t)r0 * ld_in + j] : 0.0;
    im[j] = (j < ncols_in && r1 < nrows_in) ? in[(int64_t)r1 * ld_in + j] : 0.0;
  }
  __syncthreads();
  lds_fft<N, 1, kFFTThreads, -1>(re, im, 1, N, tw, nmax);
  constexpr int H = N / 2 + 1;
  for (int k = threadIdx.x; k < H; k += kFFTThreads) {
    const int nk = (N - k) & (N - 1);
    const double zr = re[k], zi = im[k], nr = re[nk], ni = im[nk];
    out[(int64_t)r0 * H + k] = make_double2((zr + nr) * 0.5, (zi - ni) * 0.5);
    if (r1 < nrows_in) out[(int64_t)r1 * H + k] = make_double2((zi + ni) * 0.5, -(zr - nr) * 0.5);
  }
}

template <int N>
__global__ __launch_bounds__(kFFTThreads) void fft_rows_r2c(const double* __restrict__ in,
                                                            int nrows_in, int ncols_in,
                                                            int ld_in, double2* __restrict__ out,
                                                            const double2* __restrict__ tw,
                                                            int nmax, const SC2DState* st) {
  rows_r2c_body<N>(in, nrows_in, ncols_in, ld_in, out, tw, nmax, st, blockIdx.x);
}

// (b) column transforms along x (N = M0).  GREEN: full forward transform of all N rows,
// written transposed as ghat[k][i].  SOLVE: rows >= nrows_in are zero; forward, multiply by
// the rescaled Green spectrum, inverse, write only rows < nrows_in back.
template <bool GREEN, int N>
__device__ __forceinline__ void cols_body(double2* __restrict__ spec, int nrows_in, int H,
                                          double2* __restrict__ ghat_t,
                                          const double2* __restrict__ tw, int nmax,
                                          const SC2DState* st, double npad, int k) {
  if (st && st->skip) return;
  extern __shared__ double sm[];
  double* re = sm;
  double* im = sm + N;
  for (int i = threadIdx.x; i < N; i += kFFTThreads) {
    const double2 v = (i < nrows_in) ? spec[(int64_t)i * H + k] : make_double2(0.0, 0.0);
    re[i] = v.x;
    im[i] = v.y;
  }
  __syncthreads();
  lds_fft<N, 1, kFFTThreads, -1>(re, im, 1, N, tw, nmax);
  if constexpr (GREEN) {
    for (int i = threadIdx.x; i < N; i += kFFTThreads)
      ghat_t[(int64_t)k * N + i] = make_double2(re[i], im[i]);
    return;
  } else {
    // PoissonSolverFFT2D.cc:215-236: Green spectrum rescaled to the rho grid step
    const double s = st->scale_coeff;
    for (int i = threadIdx.x; i < N; i += kFFTThreads) {
      const double2 g = ghat_t[(int64_t)k * N + i];
      const double gr_re = ((i == 0 && k == 0) ? (g.x + s * npad) : g.x) - s;
      const double gr_im = g.y;
      const double a = re[i], b = im[i];
      re[i] = a * gr_re - b * gr_im;
      im[i] = a * gr_im + b * gr_re;
    }
    __syncthreads();
    lds_fft<N, 1, kFFTThreads, +1>(re, im, 1, N, tw, nmax);
    for (int i = threadIdx.x; i < nrows_in; i += kFFTThreads)
      spec[(int64_t)i * H + k] = make_double2(re[i], im[i]);
  }
}

template <bool GREEN, int N>
__global__ __launch_bounds__(kFFTThreads) void fft_cols(double2* __restrict__ spec, int nrows_in,
                                                        int H, double2* __restrict__ ghat_t,
                                                        const double2* __restrict__ tw, int nmax,
                                                        const SC2DState* st, double npad) {
  cols_body<GREEN, N>(spec, nrows_in, H, ghat_t, tw, nmax, st, npad, blockIdx.x);
}

// (c) c2r of two rows per complex inverse FFT; writes the first ncols_out columns scaled.
template <int N>
__device__ __forceinline__ void rows_c2r_body(const double2* __restrict__ spec, int nrows,
                                              double* __restrict__ out, int ld_out, int ncols_out,
                                              double denom, const double2* __restrict__ tw,
                                              int nmax, const SC2DState* st, int blk) {
  if (st && st->skip) return;
  extern __shared__ double sm[];
  double* re = sm;
  double* im = sm + N;
  constexpr int H = N / 2 + 1;
  const int r0 = 2 * blk, r1 = r0 + 1;
  const bool has1 = r1 < nrows;
  for (int k = threadIdx.x; k < H; k += kFFTThreads) {
    double2 a = spec[(int64_t)r0 * H + k];
    double2 b = has1 ? spec[(int64_t)r1 * H + k] : make_double2(0.0, 0.0);
    if (k == 0 || k == N / 2) {  // c2r ignores the imaginary DC/Nyquist parts
      a.y = 0.0;
      b.y = 0.0;
    }
    // Z = X0 + i X1 with X(N-k) = conj(X(k))
    re[k] = a.x - b.y;
    im[k] = a.y + b.x;
    if (k > 0 && k < N / 2) {
      re[N - k] = a.x + b.y;
      im[N - k] = b.x - a.y;
    }
  }
  __syncthreads();
  lds_fft<N, 1, kFFTThreads, +1>(re, im, 1, N, tw, nmax);
  for (int j = threadIdx.x; j < ncols_out; j += kFFTThreads) {
    out[(int64_t)r0 * ld_out + j] = denom * re[j];
    if (has1) out[(int64_t)r1 * ld_out + j] = denom * im[j];
  }
}

template <int N>
__global__ __launch_bounds__(kFFTThreads) void fft_rows_c2r(const double2* __restrict__ spec,
                                                            int nrows, double* __restrict__ out,
                                                            int ld_out, int ncols_out,
                                                            double denom,
                                                            const double2* __restrict__ tw,
                                                            int nmax, const SC2DState* st) {
  rows_c2r_body<N>(spec, nrows, out, ld_out, ncols_out, denom, tw, nmax, st, blockIdx.x);
}

// ---- batches: every bunch of a batch in one launch per phase ----------------------
__device__ __forceinline__ SC2DParams batch_params(const BatchBunch& B, double length,
                                                   double pipe_radius) {
  SC2DParams P;
  P.nx = B.nx;
  P.ny = B.ny;
  P.nz = B.nz;
  P.nblk = B.item_end - B.item_begin;
  P.xy_ratio = B.xy_ratio;
  P.dx_solver = B.dx_solver;
  P.macrosize = B.macrosize;
  P.factor0 = 2 * length * B.r0 / B.gamma_D;  // same operations as the single-bunch host path
  P.pipe_radius = pipe_radius;
  P.n_long_avg = B.n_long_avg;
  return P;
}

// Bunch whose block range [first(b), first(b+1)) holds block blk (first ascending).
__device__ __forceinline__ int batch_find(const BatchBunch* bb, int nb, int blk, bool rows) {
  int lo = 0, hi = nb - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    const int f = rows ? bb[mid].rows_begin : bb[mid].cols_begin;
    if (f <= blk) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__global__ __launch_bounds__(kSetupThreads) void sc2d_setup_batch(const BatchBunch* bb,
                                                                  const double* partials,
                                                                  double length,
                                                                  double pipe_radius) {
  const BatchBunch& B = bb[blockIdx.x];
  const SC2DParams P = batch_params(B, length, pipe_radius);
  sc2d_setup_body(P, partials, B.item_begin, B.item_end, B.st, B.rho, B.zg, B.zd);
}

// One block per work item.  Bunches whose grid fits (deposit_lds) accumulate their item's
// rho in LDS and add it to the grid with one atomic per touched node.
__global__ __launch_bounds__(kDepositThreads) void sc2d_deposit_batch(const BatchBunch* bb,
                                                                      const BatchItem* items) {
  extern __shared__ double lrho[];
  const BatchItem it = items[blockIdx.x];
  const BatchBunch& B = bb[it.bunch];
  const SC2DParams P = batch_params(B, 0.0, 1.0);
  if (B.deposit_lds) {
    const int ng = B.nx * B.ny;
    for (int i = threadIdx.x; i < ng; i += blockDim.x) lrho[i] = 0.0;
    __syncthreads();
    sc2d_deposit_body(P, B.v, B.st, lrho, B.zg, it.begin + threadIdx.x, it.end, blockDim.x);
    __syncthreads();
    for (int i = threadIdx.x; i < ng; i += blockDim.x)
      if (lrho[i] != 0.0) atomicAdd(&B.rho[i], lrho[i]);
  } else {
    sc2d_deposit_body(P, B.v, B.st, B.rho, B.zg, it.begin + threadIdx.x, it.end, blockDim.x);
  }
}

__global__ __launch_bounds__(kSetupThreads) void sc2d_post_batch(const BatchBunch* bb,
                                                                 double length,
                                                                 double pipe_radius) {
  const BatchBunch& B = bb[blockIdx.x];
  const SC2DParams P = batch_params(B, length, pipe_radius);
  sc2d_post_body(P, B.st, B.rho, B.zg, B.zd);
}

#define OB_BATCH_FFT_SIZES(X) X(8) X(16) X(32) X(64) X(128) X(256) X(512) X(1024)

__global__ __launch_bounds__(kFFTThreads) void rows_r2c_batch(const BatchBunch* bb, int nb) {
  const int b = batch_find(bb, nb, blockIdx.x, true);
  const BatchBunch& B = bb[b];
  const int blk = blockIdx.x - B.rows_begin;
  switch (B.M1) {
#define OB_CASE(n) \
  case n: rows_r2c_body<n>(B.rho, B.nx, B.ny, B.ny, B.spec, B.tw, B.nmax, B.st, blk); break;
    OB_BATCH_FFT_SIZES(OB_CASE)
#undef OB_CASE
    default: break;
  }
}

__global__ __launch_bounds__(kFFTThreads) void cols_batch(const BatchBunch* bb, int nb) {
  const int b = batch_find(bb, nb, blockIdx.x, false);
  const BatchBunch& B = bb[b];
  const int k = blockIdx.x - B.cols_begin;
  const double npad = (double)B.M0 * (double)B.M1;
  switch (B.M0) {
#define OB_CASE(n) \
  case n: cols_body<false, n>(B.spec, B.nx, B.H, B.ghat_t, B.tw, B.nmax, B.st, npad, k); break;
    OB_BATCH_FFT_SIZES(OB_CASE)
#undef OB_CASE
    default: break;
  }
}

__global__ __launch_bounds__(kFFTThreads) void rows_c2r_batch(const BatchBunch* bb, int nb) {
  const int b = batch_find(bb, nb, blockIdx.x, true);
  const BatchBunch& B = bb[b];
  const int blk = blockIdx.x - B.rows_begin;
  const double denom = 1.0 / (B.M0 * B.M1);
  switch (B.M1) {
#define OB_CASE(n) \
  case n:                                                                                    \
    rows_c2r_body<n>(B.spec, B.nx, B.phi, B.ny, B.ny, denom, B.tw, B.nmax, B.st, blk); \
    break;
    OB_BATCH_FFT_SIZES(OB_CASE)
#undef OB_CASE
    default: break;
  }
}

// Green function of PoissonSolverFFT2D::_defineGreenF (:115-172) on the padded grid.
__global__ void green2d_fill(double* __restrict__ G, int M0, int M1, double dx, double dy) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)M0 * M1) return;
  const int i = (int)(t / M1), j = (int)(t - (int64_t)i * M1);
  const int ii = (i <= M0 / 2) ? i : M0 - i;
  const int jj = (j <= M1 / 2) ? j : M1 - j;
  const double rTransX = ii * dx, rTransY = jj * dy;
  const double rTot2 = rTransX * rTransX + rTransY * rTransY;
  G[t] = (ii != 0 || jj != 0) ? -log(rTot2) / 2 : 0.0;
}

}  // namespace
}  // namespace ob

struct ob_sc2p5drb {
  int nx, ny, nz, M0, M1, H, nmax;
  double xy_ratio, dx_solver, dy_solver;
  int n_long_avg = 3;
  double *rho = nullptr, *phi = nullptr, *zg = nullptr, *zd = nullptr;
  double2 *spec = nullptr, *ghat_t = nullptr, *tw = nullptr;
  double* partials = nullptr;
  ob::SC2DState* st = nullptr;
  bool tracked = false;
};

namespace ob {
namespace {
template <int N>
size_t fft_lds() { return sizeof(double) * 2 * N; }

// Dynamic LDS above 64 KiB has to be allowed per kernel; done once per instantiation.
template <typename K>
void allow_lds(K* kernel, size_t bytes) {
  if (bytes > 64 * 1024)
    OB_HIP(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)bytes));
}

void launch_rows_r2c(ob_sc2p5drb* c, const double* in, int nrows, int ncols, int ld,
                     const SC2DState* st) {
  dispatch_pow2(c->M1, [&](auto NC) {
    constexpr int N = decltype(NC)::value;
    if constexpr (N >= 8) {
      static bool ready = false;
      if (!ready) {
        allow_lds(fft_rows_r2c<N>, fft_lds<N>());
        ready = true;
      }
      ob::ProfScope ps_("fft_rows_r2c");
      hipLaunchKernelGGL(fft_rows_r2c<N>, dim3((nrows + 1) / 2), dim3(kFFTThreads), fft_lds<N>(),
                         stream(), in, nrows, ncols, ld, c->spec, c->tw, c->nmax, st);
    }
  });
  OB_HIP(hipGetLastError());
}

template <bool GREEN>
void launch_cols(ob_sc2p5drb* c, int nrows, const SC2DState* st, double npad) {
  dispatch_pow2(c->M0, [&](auto NC) {
    constexpr int N = decltype(NC)::value;
    if constexpr (N >= 8) {
      static bool ready = false;
      if (!ready) {
        allow_lds(fft_cols<GREEN, N>, fft_lds<N>());
        ready = true;
      }
      ob::ProfScope ps_(GREEN ? "fft_cols_green" : "fft_cols");
      hipLaunchKernelGGL((fft_cols<GREEN, N>), dim3(c->H), dim3(kFFTThreads), fft_lds<N>(),
                         stream(), c->spec, nrows, c->H, c->ghat_t, c->tw, c->nmax, st, npad);
    }
  });
  OB_HIP(hipGetLastError());
}

void launch_fft_pipeline(ob_sc2p5drb* c, const SC2DState* st) {
  launch_rows_r2c(c, c->rho, c->nx, c->ny, c->ny, st);
  launch_cols<false>(c, c->nx, st, (double)c->M0 * (double)c->M1);
  const double denom = 1.0 / (c->M0 * c->M1);
  dispatch_pow2(c->M1, [&](auto NC) {
    constexpr int N = decltype(NC)::value;
    if constexpr (N >= 8) {
      static bool ready = false;
      if (!ready) {
        allow_lds(fft_rows_c2r<N>, fft_lds<N>());
        ready = true;
      }
      ob::ProfScope ps_("fft_rows_c2r");
      hipLaunchKernelGGL(fft_rows_c2r<N>, dim3((c->nx + 1) / 2), dim3(kFFTThreads), fft_lds<N>(),
                         stream(), c->spec, c->nx, c->phi, c->ny, c->ny, denom, c->tw, c->nmax,
                         st);
    }
  });
  OB_HIP(hipGetLastError());
}

void define_green(ob_sc2p5drb* c) {
  hipStream_t s = stream();
  double* G = nullptr;
  OB_HIP(hipMalloc((void**)&G, sizeof(double) * c->M0 * c->M1));
  const int64_t tot = (int64_t)c->M0 * c->M1;
  {
    ob::ProfScope ps_("green2d_fill");
    hipLaunchKernelGGL(green2d_fill, dim3((tot + 255) / 256), dim3(256), 0, s, G, c->M0, c->M1,
                       c->dx_solver, c->dy_solver);
  }
  OB_HIP(hipGetLastError());
  // full transform: all M0 rows are non-zero
  launch_rows_r2c(c, G, c->M0, c->M1, c->M1, nullptr);
  launch_cols<true>(c, c->M0, nullptr, 0.0);
  OB_HIP(hipStreamSynchronize(s));
  OB_HIP(hipFree(G));
}
}  // namespace

void sc2d_batch_record(ob_sc2p5drb* c, BatchBunch& r) {
  r.st = c->st;
  r.rho = c->rho;
  r.phi = c->phi;
  r.zg = c->zg;
  r.zd = c->zd;
  r.spec = c->spec;
  r.ghat_t = c->ghat_t;
  r.tw = c->tw;
  r.nx = c->nx;
  r.ny = c->ny;
  r.nz = c->nz;
  r.M0 = c->M0;
  r.M1 = c->M1;
  r.H = c->H;
  r.nmax = c->nmax;
  r.n_long_avg = c->n_long_avg;
  r.xy_ratio = c->xy_ratio;
  r.dx_solver = c->dx_solver;
}

void sc2d_mark_tracked(ob_sc2p5drb* c) { c->tracked = true; }

int sc2d_read_status(ob_sc2p5drb* c) {
  int v = 0;
  OB_HIP(hipMemcpy(&v, &c->st->status, sizeof(int), hipMemcpyDeviceToHost));
  return v;
}

void launch_sc2d_batch(const BatchBunch* bb, const BatchItem* items, const BatchShape& sh,
                       const double* partials, double length, double pipe_radius,
                       hipStream_t s) {
  static bool ready = false;
  if (!ready) {
    OB_HIP(hipFuncSetAttribute((const void*)sc2d_deposit_batch,
                               hipFuncAttributeMaxDynamicSharedMemorySize, 128 * 1024));
    ready = true;
  }
  OB_REQUIRE(sh.deposit_lds_bytes <= 128 * 1024, "batch deposit: LDS block too large");
  OB_REQUIRE(sh.max_M <= 1024, "batch: grids up to 512 points per axis");
  {
    ProfScope ps("sc2d_setup_batch");
    hipLaunchKernelGGL(sc2d_setup_batch, dim3(sh.nb), dim3(kSetupThreads), 0, s, bb, partials,
                       length, pipe_radius);
  }
  {
    ProfScope ps("sc2d_deposit_batch");
    hipLaunchKernelGGL(sc2d_deposit_batch, dim3(sh.nitems), dim3(kDepositThreads),
                       sh.deposit_lds_bytes, s, bb, items);
  }
  {
    ProfScope ps("sc2d_post_batch");
    hipLaunchKernelGGL(sc2d_post_batch, dim3(sh.nb), dim3(kSetupThreads), 0, s, bb, length,
                       pipe_radius);
  }
  const size_t fl = sizeof(double) * 2 * sh.max_M;
  {
    ProfScope ps("fft_rows_r2c_batch");
    hipLaunchKernelGGL(rows_r2c_batch, dim3(sh.row_blocks), dim3(kFFTThreads), fl, s, bb, sh.nb);
  }
  {
    ProfScope ps("fft_cols_batch");
    hipLaunchKernelGGL(cols_batch, dim3(sh.col_blocks), dim3(kFFTThreads), fl, s, bb, sh.nb);
  }
  {
    ProfScope ps("fft_rows_c2r_batch");
    hipLaunchKernelGGL(rows_c2r_batch, dim3(sh.row_blocks), dim3(kFFTThreads), fl, s, bb, sh.nb);
  }
  OB_HIP(hipGetLastError());
}
}  // namespace ob

#define OB_API extern "C" __attribute__((visibility("default")))
#define OB_TRY try {
#define OB_CATCH                       \
  }                                    \
  catch (const ob::Error& e) {         \
    ob::set_error(e.msg);              \
    return 1;                          \
  }                                    \
  return 0;

OB_API int ob_sc2p5drb_create(int nx, int ny, int nz, double xy_ratio, ob_sc2p5drb** out) {
  OB_TRY OB_REQUIRE(out != nullptr, "null output pointer");
  OB_REQUIRE(nx >= 3 && ny >= 3, "SpaceChargeCalc2p5Drb: grid size too small (should be >= 3)");
  OB_REQUIRE(nz >= 1, "SpaceChargeCalc2p5Drb: number of z bins should be >= 1");
  OB_REQUIRE(nz <= ob::kMaxZBinsLDS, "SpaceChargeCalc2p5Drb: more than 4096 z bins");
  OB_REQUIRE(xy_ratio > 0.0, "SpaceChargeCalc2p5Drb: xy_ratio must be positive");
  ob::ensure_init();
  ob_sc2p5drb* c = new ob_sc2p5drb();
  c->nx = nx;
  c->ny = ny;
  c->nz = nz;
  c->M0 = ob::pad_len(nx);
  c->M1 = ob::pad_len(ny);
  if (c->M0 > 4096 || c->M1 > 4096) {
    delete c;
    throw ob::Error{"SpaceChargeCalc2p5Drb: grid larger than 2048 points per axis"};
  }
  c->H = c->M1 / 2 + 1;
  c->nmax = c->M0 > c->M1 ? c->M0 : c->M1;
  c->xy_ratio = xy_ratio;
  // solver grid [-xy_ratio, xy_ratio] x [-1, 1] (:33, PoissonSolver2D.cc:31-32)
  c->dx_solver = (xy_ratio - -xy_ratio) / (nx - 1);
  c->dy_solver = (1.0 - -1.0) / (ny - 1);
  OB_HIP(hipMalloc((void**)&c->rho, sizeof(double) * nx * ny));
  OB_HIP(hipMalloc((void**)&c->phi, sizeof(double) * nx * ny));
  OB_HIP(hipMalloc((void**)&c->zg, sizeof(double) * nz));
  OB_HIP(hipMalloc((void**)&c->zd, sizeof(double) * nz));
  OB_HIP(hipMalloc((void**)&c->spec, sizeof(double2) * (size_t)c->M0 * c->H));
  OB_HIP(hipMalloc((void**)&c->ghat_t, sizeof(double2) * (size_t)c->M0 * c->H));
  OB_HIP(hipMalloc((void**)&c->partials, sizeof(double) * 6 * 4096));
  OB_HIP(hipMalloc((void**)&c->st, sizeof(ob::SC2DState)));
  OB_HIP(hipMemset(c->rho, 0, sizeof(double) * nx * ny));
  OB_HIP(hipMemset(c->phi, 0, sizeof(double) * nx * ny));
  OB_HIP(hipMemset(c->zg, 0, sizeof(double) * nz));
  OB_HIP(hipMemset(c->zd, 0, sizeof(double) * nz));
  ob::SC2DState init;
  std::memset(&init, 0, sizeof(init));
  init.xMin = -1.0; init.xMax = 1.0; init.yMin = -1.0; init.yMax = 1.0;  // Grid2D(nx, ny)
  init.dx = 2.0 / (nx - 1); init.dy = 2.0 / (ny - 1);
  init.zMin = -0.5; init.zMax = 0.5; init.dz = 1.0 / nz;                // Grid1D(nz)
  OB_HIP(hipMemcpy(c->st, &init, sizeof(init), hipMemcpyHostToDevice));
  c->tw = ob::make_twiddles(c->nmax);
  ob::define_green(c);
  *out = c;
  OB_CATCH
}

OB_API int ob_sc2p5drb_destroy(ob_sc2p5drb* c) {
  OB_TRY if (!c) return 0;
  // Pending kicks read st/phi/zg/zd: apply them before the buffers go.
  std::string err;
  try {
    ob::batch_forget_calc(c);
    ob::flush_kicks_reading(c->st);
  } catch (const ob::Error& e) {
    err = e.msg;
  }
  (void)hipStreamSynchronize(ob::stream());
  (void)hipFree(c->rho); (void)hipFree(c->phi); (void)hipFree(c->zg); (void)hipFree(c->zd);
  (void)hipFree(c->spec); (void)hipFree(c->ghat_t); (void)hipFree(c->tw); (void)hipFree(c->partials); (void)hipFree(c->st);
  delete c;
  OB_REQUIRE(err.empty(), err);
  OB_CATCH
}

OB_API int ob_sc2p5drb_track(ob_sc2p5drb* c, ob_bunch* b, double length, double pipe_radius) {
  OB_TRY OB_REQUIRE(c != nullptr && b != nullptr, "null handle");
  ob::check_statuses(false);
  const int64_t n = ob::bunch_size(b);
  if (n < 2) return 0;  // :112-113
  hipStream_t s = ob::stream();
  const int nblk = ob::bunch_flush(b, c->partials);
  ob::DeviceBunchView v = ob::bunch_device_view(b);
  const ob::SyncInfo sy = ob::bunch_sync(b);
  ob::SC2DParams P;
  P.nx = c->nx; P.ny = c->ny; P.nz = c->nz; P.nblk = nblk;
  P.xy_ratio = c->xy_ratio;
  P.dx_solver = c->dx_solver;
  P.macrosize = sy.macrosize;
  P.factor0 = 2 * length * sy.r0 / ((sy.beta * sy.beta) * pow(sy.gamma, 3));  // :129
  P.pipe_radius = pipe_radius;
  P.n_long_avg = c->n_long_avg;
  {
    ob::ProfScope ps_("sc2d_setup");
    hipLaunchKernelGGL(ob::sc2d_setup, dim3(1), dim3(ob::kSetupThreads), 0, s, P, c->partials,
                       c->st, c->rho, c->zg, c->zd);
  }
  int dblk = (int)((n + ob::kDepositThreads - 1) / ob::kDepositThreads);
  const int cap = 8 * ob::device_cu_count();
  if (dblk > cap) dblk = cap;
  {
    ob::ProfScope ps_("sc2d_deposit");
    hipLaunchKernelGGL(ob::sc2d_deposit, dim3(dblk), dim3(ob::kDepositThreads), 0, s, P, v, c->st,
                       c->rho, c->zg);
  }
  {
    ob::ProfScope ps_("sc2d_post");
    hipLaunchKernelGGL(ob::sc2d_post, dim3(1), dim3(ob::kSetupThreads), 0, s, P, c->st, c->rho,
                       c->zg, c->zd);
  }
  ob::launch_fft_pipeline(c, c->st);
  OB_HIP(hipGetLastError());
  ob::watch_status(&c->st->status,
                   "SpaceChargeCalc2p5Drb::bunchAnalysis: the bunch min and max sizes are wrong "
                   "(zero extent); cannot calculate space charge");
  // kick (:146-167), fused into the next map kernel
  ob::Op op;
  std::memset(&op, 0, sizeof(op));
  op.kind = ob::OP_SC2P5DRB_KICK;
  op.i0 = c->nx;
  op.i1 = c->ny;
  op.i2 = c->nz;
  op.p[0] = ob::ptr_bits(c->st);
  op.p[1] = ob::ptr_bits(c->phi);
  op.p[2] = ob::ptr_bits(c->zg);
  op.p[3] = ob::ptr_bits(c->zd);
  op.p[4] = 2 * log(pipe_radius);                                             // :139
  op.p[5] = -length * sy.r0 * sy.mass / (sy.gamma * sy.gamma);                // :141
  ob::bunch_queue(b, op);
  c->tracked = true;
  OB_CATCH
}

OB_API int ob_sc2p5drb_get_grid(ob_sc2p5drb* c, int which, double* out, double* lims) {
  OB_TRY OB_REQUIRE(c != nullptr, "null handle");
  OB_REQUIRE(which >= 0 && which <= 3, "grid selector must be 0..3");
  hipStream_t s = ob::stream();
  ob::SC2DState h;
  OB_HIP(hipMemcpyAsync(&h, c->st, sizeof(h), hipMemcpyDeviceToHost, s));
  const double* src = which == 0 ? c->rho : which == 1 ? c->phi : which == 2 ? c->zg : c->zd;
  const size_t cnt = which <= 1 ? (size_t)c->nx * c->ny : (size_t)c->nz;
  if (out) OB_HIP(hipMemcpyAsync(out, src, sizeof(double) * cnt, hipMemcpyDeviceToHost, s));
  OB_HIP(hipStreamSynchronize(s));
  ob::check_statuses(true);
  if (lims) {
    if (which <= 1) {
      lims[0] = h.xMin; lims[1] = h.xMax; lims[2] = h.yMin; lims[3] = h.yMax;
    } else {
      lims[0] = h.zMin; lims[1] = h.zMax;
    }
  }
  OB_CATCH
}

OB_API int ob_sc2p5drb_get_moments(ob_sc2p5drb* c, double* out) {
  OB_TRY OB_REQUIRE(c != nullptr, "null handle");
  ob::SC2DState h;
  OB_HIP(hipMemcpyAsync(&h, c->st, sizeof(h), hipMemcpyDeviceToHost, ob::stream()));
  OB_HIP(hipStreamSynchronize(ob::stream()));
  out[0] = h.x_c; out[1] = h.y_c; out[2] = h.a_bunch; out[3] = h.total;
  OB_CATCH
}

OB_API int ob_sc2p5drb_set_long_avg_n(ob_sc2p5drb* c, int n) {
  OB_TRY OB_REQUIRE(c != nullptr, "null handle");
  OB_REQUIRE(n >= 1, "number of averaging points must be >= 1");
  c->n_long_avg = n;
  OB_CATCH
}

OB_API int ob_sc2p5drb_get_long_avg_n(ob_sc2p5drb* c, int* n) {
  OB_TRY OB_REQUIRE(c != nullptr, "null handle");
  *n = c->n_long_avg;
  OB_CATCH
}
