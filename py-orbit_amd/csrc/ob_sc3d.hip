// ob_sc3d.hip -- SpaceChargeCalc3D on the device.
//
// Per trackBunch (SpaceChargeCalc3D.cc:100-148):
//   1. fused map kernel flushes queued maps and emits per-block extrema; one block reduces
//      them and the six numbers come to the host (the Green-function rebuild decision of
//      bunchAnalysis, :186-203, launches kernels, so it is taken on the host);
//   2. host: bunchAnalysis grid sizing (:150-249), identical scalar arithmetic;
//   3. (on a shape change) Green function 1/r and its FFT, built on the device (:160-247);
//   4. 27-point TSC deposit (Grid3D.cc:281-432) with FP64 atomics;
//   5. Hockney solve (PoissonSolverFFT3D.cc:249-347) in five passes over a z-fastest
//      layout: z r2c (two real lines per complex FFT), y FFT, x FFT x Green x inverse x FFT,
//      inverse y FFT, z c2r.  Zero padding is pruned: lines that are entirely zero are
//      never read, and only the nx*ny*nz outputs that are kept are written;
//   6. the kick (:120-147) is queued on the bunch and runs fused into the next map kernel.
//
// The Green spectrum of the even function 1/r is real and even in kx and ky, so only the
// quadrant kx <= M0/2, ky <= M1/2 of its real part is stored (8x less than the complex
// spectrum the reference keeps; the discarded imaginary part is rounding noise).
#include <chrono>
#include <cmath>
#include <cstring>

#include "ob_fft.hpp"
#include "ob_internal.hpp"
#include "ob_sc_device.hpp"
#include "ob_tiled3d.hpp"

namespace ob {
namespace {

constexpr int kThreads = 256;
constexpr double kSpeedOfLight = 2.99792458e+8;  // OrbitConst.cc:27
// Complex lines (= pairs of real z-lines) per block in the z passes.
__host__ __device__ constexpr int zpass_pairs(int n) { return n >= 256 ? (n >= 2048 ? 1 : 2048 / n) : 8; }
// Consecutive kz lines per block in the y and x passes (LDS <= 131 KiB), and threads.
__host__ __device__ constexpr int strided_lines(int n) { return n >= 2048 ? 4 : 8; }
__host__ __device__ constexpr int strided_threads(int n) { return n >= 256 ? 512 : 256; }

// CONV2: the x pass of a stack of independent 2D solves (SpaceChargeCalcSliceBySlice2D):
// the Green spectrum is 2D, ghat[kx * qy + ky], the same for every line.
// GREEN: an even-symmetric transform along one axis of the real Green quadrant.  Two
// adjacent kz lines (2c, 2c+1) read as one double2 are the complex line a + i b; the
// N/2+1 stored samples are extended evenly to N points, and the first N/2+1 outputs
// A + i B (A, B real: the transforms of even real lines) go back as the same double2.
// n_in = the number of real kz lines; a partner line at or past it is padding, zero.
enum StridedMode { FWD = 0, INV = 1, CONV = 2, CONV2 = 3, GREEN = 4 };

// LDS doubles per line (re or im) of a strided block: the larger of the forward and the
// reversed plan's padded layouts.
template <int N>
__host__ __device__ constexpr int strided_line_doubles() {
  return RegPlan<N, false, 1, 1>::LDS_LINE > RegPlan<N, true, 1, 1>::LDS_LINE
             ? RegPlan<N, false, 1, 1>::LDS_LINE
             : RegPlan<N, true, 1, 1>::LDS_LINE;
}

struct GreenGeom {
  int M0, M1, M2;
  double dx, dy, dz;
  int nbunches;   // external bunches (images), even
  double lambda;  // their spacing
};

__device__ __forceinline__ double green3d(int i, int j, int k, const GreenGeom& g) {
  // PoissonSolverFFT3D::_defineGreenF (:160-247)
  const int ii = (i <= g.M0 / 2) ? i : g.M0 - i;
  const int jj = (j <= g.M1 / 2) ? j : g.M1 - j;
  const int kk = (k <= g.M2 / 2) ? k : g.M2 - k;
  const double rZ = kk * g.dz, rZ2 = rZ * rZ;
  const double rY = jj * g.dy, rY2 = rY * rY;
  const double rX = ii * g.dx, rX2 = rX * rX;
  const double rTot = sqrt(rX2 + rY2 + rZ2);
  double externalPhi = 0.;
  for (int b = -g.nbunches / 2; b <= g.nbunches / 2; b++) {  // :184-196
    if (b == 0) continue;
    const double rZt = rZ + b * g.lambda;
    const double rTotExt = sqrt(rX2 + rY2 + rZt * rZt);
    if (rTotExt > 1.0e-10) externalPhi += 1.0 / rTotExt;
  }
  return (ii != 0 || jj != 0 || kk != 0) ? 1. / rTot + externalPhi : externalPhi;
}

// ---- extrema reduction (one block) -------------------------------------------
// Writes the six extrema to out (pinned host memory) and then seq to *seq_out with a
// system-scope release, all from one thread, so a host that reads *seq_out == seq (with
// acquire) reads the extrema of this launch.
__global__ __launch_bounds__(1024) void reduce_extrema(const double* __restrict__ partials,
                                                       int nblk, double* __restrict__ out,
                                                       uint64_t* __restrict__ seq_out,
                                                       uint64_t seq) {
  __shared__ double red[16][6];
  double r[6] = {DBL_MAX, -DBL_MAX, DBL_MAX, -DBL_MAX, DBL_MAX, -DBL_MAX};
  // four rows per thread and round, their loads in flight together
  constexpr int U = 4;
  for (int b0 = threadIdx.x; b0 < nblk; b0 += U * blockDim.x) {
    double v[U][6];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int b = b0 + u * (int)blockDim.x;
#pragma unroll
      for (int k = 0; k < 6; k++) v[u][k] = b < nblk ? partials[b * 6 + k] : r[k];
    }
#pragma unroll
    for (int u = 0; u < U; u++)
#pragma unroll
      for (int k = 0; k < 6; k++) r[k] = (k & 1) ? fmax(r[k], v[u][k]) : fmin(r[k], v[u][k]);
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
    double acc[6];
    for (int k = 0; k < 6; k++) {
      acc[k] = red[0][k];
      for (int w = 1; w < (int)(blockDim.x >> 6); w++)
        acc[k] = (k & 1) ? fmax(acc[k], red[w][k]) : fmin(acc[k], red[w][k]);
    }
    for (int k = 0; k < 6; k++) out[k] = acc[k];
    __hip_atomic_store(seq_out, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// The bunch extrema on the host for the grid sizing of a step (bunchAnalysis runs on the
// host, as its Green-function rebuild decision launches kernels).  reduce_extrema writes
// them straight into a pinned coherent buffer, so there is no copy kernel, and the host
// watches the sequence number instead of waiting for the stream to drain: it spins for at
// most 20 ms, then synchronises the stream.
struct ExtremaLink {
  double* host = nullptr;  // [0..5] xMin, xMax, yMin, yMax, zMin, zMax; [6] sequence number
  double* dev = nullptr;   // device address of host
  uint64_t seq = 0;
};

void extrema_link_create(ExtremaLink& e) {
  OB_HIP(hipHostMalloc((void**)&e.host, sizeof(double) * 8, hipHostMallocCoherent));
  std::memset(e.host, 0, sizeof(double) * 8);
  OB_HIP(hipHostGetDevicePointer((void**)&e.dev, e.host, 0));
  e.seq = 0;
}

void extrema_link_destroy(ExtremaLink& e) {
  if (e.host) (void)hipHostFree(e.host);
  e.host = e.dev = nullptr;
}

const double* extrema_to_host(ExtremaLink& e, const double* partials, int nblk, hipStream_t s) {
  const uint64_t want = ++e.seq;
  uint64_t* seq_host = reinterpret_cast<uint64_t*>(e.host + 6);
  {
    ProfScope ps_("reduce_extrema");
    hipLaunchKernelGGL(reduce_extrema, dim3(1), dim3(1024), 0, s, partials, nblk, e.dev,
                       reinterpret_cast<uint64_t*>(e.dev + 6), want);
  }
  OB_HIP(hipGetLastError());
  const auto t0 = std::chrono::steady_clock::now();
  while (__atomic_load_n(seq_host, __ATOMIC_ACQUIRE) != want) {
    if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(20)) {
      OB_HIP(hipStreamSynchronize(s));
      OB_REQUIRE(__atomic_load_n(seq_host, __ATOMIC_ACQUIRE) == want,
                 "bunch extrema: the reduction did not reach the host");
      break;
    }
  }
  return e.host;
}

__global__ void fill_zero(double* __restrict__ a, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) a[i] = 0.0;
}

// ---- deposit: Grid3D::binValue (Grid3D.cc:307-432), z-fastest layout ----------
struct Grid3 {
  int nx, ny, nz;
  double xMin, xMax, yMin, yMax, zMin, zMax, dx, dy, dz;
};

// lambda != 0: binBunch(bunch, lambda) with longWrapping = 1 (Grid3D.cc:281-305).
__global__ __launch_bounds__(kThreads) void deposit3d(DeviceBunchView v, Grid3 g,
                                                      double macrosize, double lambda,
                                                      double* __restrict__ rho) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t sx = (int64_t)g.ny * g.nz, sy = g.nz;
  const bool wrap = lambda != 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < v.n; i += stride) {
    const double x = v.c[0][i], y = v.c[2][i];
    const double z = wrap ? remainder(v.c[4][i], lambda) : v.c[4][i];
    if (x < g.xMin || x > g.xMax || y < g.yMin || y > g.yMax || z < g.zMin || z > g.zMax) continue;
    const double m = v.msize ? v.msize[i] : macrosize;
    int iX, iY, iZ;
    double xF, yF, zF;
    node_ind_frac(x, g.xMin, g.dx, g.nx, iX, xF);
    node_ind_frac(y, g.yMin, g.dy, g.ny, iY, yF);
    grid3d_z_ind_frac(z, g.zMin, g.dz, g.nz, iZ, zF, wrap);
    const double wx[3] = {0.5 * (0.5 - xF) * (0.5 - xF), 0.75 - xF * xF, 0.5 * (0.5 + xF) * (0.5 + xF)};
    const double wy[3] = {0.5 * (0.5 - yF) * (0.5 - yF), 0.75 - yF * yF, 0.5 * (0.5 + yF) * (0.5 + yF)};
    int iZm = iZ - 1, iZp = iZ + 1;
    if (iZm < 0) iZm = g.nz - 1;
    if (iZp >= g.nz) iZp = 0;
    double* base = rho + (int64_t)(iX - 1) * sx + (int64_t)(iY - 1) * sy;
    if (g.nz >= 3) {
      const double wz[3] = {0.5 * (0.5 - zF) * (0.5 - zF), 0.75 - zF * zF,
                            0.5 * (0.5 + zF) * (0.5 + zF)};
      const int zz[3] = {iZm, iZ, iZp};
#pragma unroll
      for (int a = 0; a < 3; a++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
          const double tmp = wy[c] * wz[a] * m;
#pragma unroll
          for (int d = 0; d < 3; d++) atomicAdd(base + d * sx + c * sy + zz[a], wx[d] * tmp);
        }
    } else if (g.nz == 2) {
      const double wz[2] = {1.0 - zF, zF};
      for (int a = 0; a < 2; a++)
        for (int c = 0; c < 3; c++) {
          const double tmp = wy[c] * wz[a] * m;
          for (int d = 0; d < 3; d++) atomicAdd(base + d * sx + c * sy + a, wx[d] * tmp);
        }
    } else {
      for (int c = 0; c < 3; c++) {
        const double tmp = wy[c] * m;
        for (int d = 0; d < 3; d++) atomicAdd(base + d * sx + c * sy, wx[d] * tmp);
      }
    }
  }
}

// Line q = i*NY + j of a z pass (0 <= q < 2^31, checked on the host): (i, j) of q with one
// 32-bit division, and of q + 1 from it.
__device__ __forceinline__ void line_ij(int64_t q, int NY, int& i, int& j) {
  const uint32_t u = (uint32_t)q;
  i = (int)(u / (uint32_t)NY);
  j = (int)(u - (uint32_t)i * (uint32_t)NY);
}
__device__ __forceinline__ void line_next(int NY, int i, int j, int& i1, int& j1) {
  j1 = j + 1;
  i1 = i;
  if (j1 == NY) {
    j1 = 0;
    i1++;
  }
}

// ---- z pass r2c: two real z-lines per complex FFT -------------------------------
// Real line q = i*NY + j (q < nlines) holds nz_in values, zero-padded to N = M2.  The
// half spectrum kz <= N/2 of line q goes to spec[(i*M1 + j)*H2p + kz].
template <int N>
__global__ __launch_bounds__(kThreads) void fft3_z_r2c(const double* __restrict__ rho,
                                                       int64_t nlines, int NY, int nz_in, int M1,
                                                       int H2p, double2* __restrict__ spec,
                                                       const double2* __restrict__ tw, int nmax) {
  constexpr int P = zpass_pairs(N), ls = fft_line_stride(N), H2 = N / 2 + 1;
  extern __shared__ double sm[];
  double* re = sm;
  double* im = sm + P * ls;
  const int64_t pair0 = (int64_t)blockIdx.x * P;
  // All of a thread's loads are issued before any is stored (a masked lane loads rho[0]
  // and stores 0), so their latencies overlap instead of adding up.
  constexpr int IT = (P * N + kThreads - 1) / kThreads;  // <= 8
  double a0[IT], a1[IT];
  uint32_t m0 = 0, m1 = 0;
#pragma unroll
  for (int u = 0; u < IT; u++) {
    const int idx = (int)threadIdx.x + u * kThreads;
    const int lp = idx / N, k = idx - lp * N;
    const int64_t q0 = 2 * (pair0 + lp), q1 = q0 + 1;
    const bool ok = idx < P * N && k < nz_in;
    const bool ok0 = ok && q0 < nlines, ok1 = ok && q1 < nlines;
    a0[u] = rho[ok0 ? q0 * nz_in + k : 0];
    a1[u] = rho[ok1 ? q1 * nz_in + k : 0];
    m0 |= (ok0 ? 1u : 0u) << u;
    m1 |= (ok1 ? 1u : 0u) << u;
  }
#pragma unroll
  for (int u = 0; u < IT; u++) {
    const int idx = (int)threadIdx.x + u * kThreads;
    if (idx < P * N) {
      const int lp = idx / N, k = idx - lp * N;
      re[lp * ls + k] = ((m0 >> u) & 1u) ? a0[u] : 0.0;
      im[lp * ls + k] = ((m1 >> u) & 1u) ? a1[u] : 0.0;
    }
  }
  __syncthreads();
  lds_fft<N, P, kThreads, -1>(re, im, P, ls, tw, nmax);
  for (int idx = threadIdx.x; idx < P * H2; idx += kThreads) {
    const int lp = idx / H2, k = idx - lp * H2;
    const int64_t q0 = 2 * (pair0 + lp), q1 = q0 + 1;
    if (q0 >= nlines) continue;
    const int nk = (N - k) & (N - 1);
    const double zr = re[lp * ls + k], zi = im[lp * ls + k];
    const double nr = re[lp * ls + nk], ni = im[lp * ls + nk];
    int i0, j0;
    line_ij(q0, NY, i0, j0);
    spec[((int64_t)i0 * M1 + j0) * H2p + k] = make_double2((zr + nr) * 0.5, (zi - ni) * 0.5);
    if (q1 < nlines) {
      int i1, j1;
      line_next(NY, i0, j0, i1, j1);
      spec[((int64_t)i1 * M1 + j1) * H2p + k] = make_double2((zi + ni) * 0.5, -(zr - nr) * 0.5);
    }
  }
}

// ---- z pass c2r: two half spectra per complex inverse FFT ------------------------
template <int N>
__global__ __launch_bounds__(kThreads) void fft3_z_c2r(const double2* __restrict__ spec,
                                                       int64_t nlines, int NY, int M1, int H2p,
                                                       int nz_out, double denom,
                                                       double* __restrict__ phi,
                                                       const double2* __restrict__ tw, int nmax) {
  constexpr int P = zpass_pairs(N), ls = fft_line_stride(N), H2 = N / 2 + 1;
  extern __shared__ double sm[];
  double* re = sm;
  double* im = sm + P * ls;
  const int64_t pair0 = (int64_t)blockIdx.x * P;
  // loads issued together before the stores, as in fft3_z_r2c
  constexpr int IT = (P * H2 + kThreads - 1) / kThreads;
  double2 av[IT], bv[IT];
  uint32_t ma = 0, mb = 0;
#pragma unroll
  for (int u = 0; u < IT; u++) {
    const int idx = (int)threadIdx.x + u * kThreads;
    const int lp = idx / H2, k = idx - lp * H2;
    const int64_t q0 = 2 * (pair0 + lp), q1 = q0 + 1;
    const bool ok0 = idx < P * H2 && q0 < nlines, ok1 = idx < P * H2 && q1 < nlines;
    int i0, j0, i1, j1;
    line_ij(q0, NY, i0, j0);
    line_next(NY, i0, j0, i1, j1);
    av[u] = spec[ok0 ? ((int64_t)i0 * M1 + j0) * H2p + k : 0];
    bv[u] = spec[ok1 ? ((int64_t)i1 * M1 + j1) * H2p + k : 0];
    ma |= (ok0 ? 1u : 0u) << u;
    mb |= (ok1 ? 1u : 0u) << u;
  }
#pragma unroll
  for (int u = 0; u < IT; u++) {
    const int idx = (int)threadIdx.x + u * kThreads;
    if (idx >= P * H2) continue;
    const int lp = idx / H2, k = idx - lp * H2;
    double2 a = ((ma >> u) & 1u) ? av[u] : make_double2(0.0, 0.0);
    double2 b = ((mb >> u) & 1u) ? bv[u] : make_double2(0.0, 0.0);
    if (k == 0 || k == N / 2) {  // c2r ignores the imaginary DC/Nyquist parts
      a.y = 0.0;
      b.y = 0.0;
    }
    // Z = X0 + i X1 with X(N-k) = conj(X(k))
    re[lp * ls + k] = a.x - b.y;
    im[lp * ls + k] = a.y + b.x;
    if (k > 0 && k < N / 2) {
      re[lp * ls + N - k] = a.x + b.y;
      im[lp * ls + N - k] = b.x - a.y;
    }
  }
  __syncthreads();
  lds_fft<N, P, kThreads, +1>(re, im, P, ls, tw, nmax);
  for (int idx = threadIdx.x; idx < P * nz_out; idx += kThreads) {
    const int lp = idx / nz_out, k = idx - lp * nz_out;
    const int64_t q0 = 2 * (pair0 + lp), q1 = q0 + 1;
    if (q0 < nlines) phi[q0 * nz_out + k] = denom * re[lp * ls + k];
    if (q1 < nlines) phi[q1 * nz_out + k] = denom * im[lp * ls + k];
  }
}

// ---- strided passes (y and x axes) -----------------------------------------------
// Block = (outer index o, tile of up to L consecutive kz).  Element p of line (o, kz) is
// spec[o*outer_stride + p*line_stride + kz]; a tile row is L*16 contiguous bytes, aligned
// because H2p is a multiple of 8.  Inputs p >= n_in are zero; outputs p < n_out are written.
// The first FFT stage reads global memory straight into registers and the last one writes
// from registers; in the x pass (CONV) the Green multiply sits between the forward and the
// inverse transform on the registers, so only the middle stages use LDS.
//
// Zero padding is pruned at compile time: the forward y pass and the x pass read only
// p < N/2 (n_in <= N/2 because N >= 2 n_in - 1), and the x pass and the inverse y pass keep
// only p < N/2, so half of the first-stage inputs are known zeros and half of the
// last-stage outputs are never formed.  The x pass handles ky = o and its mirror M1 - o in
// one block: both rows multiply by the same Green values, which stay in registers.
template <int MODE, int N>
using StridedIn = RegPlan<N, MODE == INV, strided_lines(N), strided_threads(N)>;
template <int N>
using StridedFwd = RegPlan<N, false, strided_lines(N), strided_threads(N)>;

// Global -> registers in the first-stage layout of the row's plan.
template <int MODE, int N>
__device__ __forceinline__ void strided_load(const double2* __restrict__ base, int nl, int n_in,
                                             int pair0, int64_t line_stride,
                                             double (&vr)[StridedIn<MODE, N>::B0][StridedIn<MODE, N>::R0],
                                             double (&vi)[StridedIn<MODE, N>::B0][StridedIn<MODE, N>::R0]) {
  constexpr int L = strided_lines(N), NT = strided_threads(N);
  constexpr bool HALF_IN = (MODE != INV && MODE != GREEN);
  using P0 = StridedIn<MODE, N>;
#pragma unroll
  for (int q = 0; q < P0::B0; q++) {
    const int b = (int)threadIdx.x + q * NT;
    const int line = b % L, j = b / L;
    const bool ok = b < L * P0::NB0 && line < nl;
#pragma unroll
    for (int m = 0; m < P0::R0; m++) {
      const int p = j + m * P0::NB0;
      double2 v = make_double2(0.0, 0.0);
      if constexpr (MODE == GREEN) {  // even extension: point p holds sample min(p, N - p)
        const int pp = p <= N / 2 ? p : N - p;
        if (ok) v = base[(int64_t)pp * line_stride + line];
        // n_in = real lines: a partner line past the last one is padding, read as zero
        // (else the rounding noise left there would grow by ~N^2 with every rebuild)
        if (2 * (pair0 + line) + 1 >= n_in) v.y = 0.0;
      } else {
        if (!(HALF_IN && 2 * m >= P0::R0) && ok && p < n_in)
          v = base[(int64_t)p * line_stride + line];
      }
      vr[q][m] = v.x;
      vi[q][m] = v.y;
    }
  }
}

// Transform(s) of one row already in registers, then registers -> global.
template <int MODE, int N>
__device__ __forceinline__ void strided_compute(
    double* re, double* im, double2* __restrict__ base, int nl, int n_out, int64_t line_stride,
    const double2* __restrict__ tw, int nmax,
    const double (&g)[StridedFwd<N>::BL][StridedFwd<N>::RL],
    double (&vr)[StridedIn<MODE, N>::B0][StridedIn<MODE, N>::R0],
    double (&vi)[StridedIn<MODE, N>::B0][StridedIn<MODE, N>::R0]) {
  constexpr int L = strided_lines(N), NT = strided_threads(N);
  constexpr bool REV0 = (MODE == INV);
  constexpr int SIGN0 = (MODE == INV) ? +1 : -1;
  constexpr bool HALF_OUT = (MODE != FWD);
  using P0 = StridedIn<MODE, N>;
  double ur[P0::BL][P0::RL], ui[P0::BL][P0::RL];
  plan_regs<N, REV0, L, NT, SIGN0>(re, im, nl, tw, nmax, vr, vi, ur, ui);
  if constexpr (MODE == CONV || MODE == CONV2) {
    using PI = RegPlan<N, true, L, NT>;
#pragma unroll
    for (int q = 0; q < P0::BL; q++)
#pragma unroll
      for (int m = 0; m < P0::RL; m++) {
        ur[q][m] *= g[q][m];
        ui[q][m] *= g[q][m];
      }
    __syncthreads();  // every thread has read its last-stage inputs from LDS
    double wr[PI::BL][PI::RL], wi[PI::BL][PI::RL];
    plan_regs<N, true, L, NT, +1>(re, im, nl, tw, nmax, ur, ui, wr, wi);
#pragma unroll
    for (int q = 0; q < PI::BL; q++) {
      const int b = (int)threadIdx.x + q * NT;
      const int line = b % L, j = b / L;
      if (b < L * PI::NBL && line < nl) {
#pragma unroll
        for (int m = 0; m < PI::RL / 2; m++) {
          const int p = j + m * PI::NBL;
          if (p < n_out) base[(int64_t)p * line_stride + line] = make_double2(wr[q][m], wi[q][m]);
        }
      }
    }
  } else if constexpr (MODE == GREEN) {  // outputs p <= N/2
#pragma unroll
    for (int q = 0; q < P0::BL; q++) {
      const int b = (int)threadIdx.x + q * NT;
      const int line = b % L, j = b / L;
      if (b < L * P0::NBL && line < nl) {
        const bool pad = 2 * line + 1 >= n_out;  // the partner line stays zero
#pragma unroll
        for (int m = 0; m <= P0::RL / 2; m++) {
          const int p = j + m * P0::NBL;
          if (p <= N / 2)
            base[(int64_t)p * line_stride + line] = make_double2(ur[q][m], pad ? 0.0 : ui[q][m]);
        }
      }
    }
  } else {
#pragma unroll
    for (int q = 0; q < P0::BL; q++) {
      const int b = (int)threadIdx.x + q * NT;
      const int line = b % L, j = b / L;
      if (b < L * P0::NBL && line < nl) {
#pragma unroll
        for (int m = 0; m < (HALF_OUT ? P0::RL / 2 : P0::RL); m++) {
          const int p = j + m * P0::NBL;
          if (p < n_out) base[(int64_t)p * line_stride + line] = make_double2(ur[q][m], ui[q][m]);
        }
      }
    }
  }
}

// Each block transforms two rows of its kz tile: the x pass takes ky = o and M1 - o, the
// y passes take outer rows 2o and 2o+1.  The second row's loads are issued before the
// first row is transformed, so they overlap its FFT work.
template <int MODE, int N>
__global__ __launch_bounds__(strided_threads(N)) void fft3_strided(
    double2* __restrict__ spec, int n_in, int n_out, int64_t line_stride, int64_t outer_stride,
    int outers, int H2, const double* __restrict__ ghat, int qy, int H2p, int M1,
    const double2* __restrict__ tw, int nmax, double s_coeff, double npad) {
  constexpr int L = strided_lines(N), NT = strided_threads(N);
  using PF = StridedFwd<N>;
  using P0 = StridedIn<MODE, N>;
  extern __shared__ double sm[];
  double* re = sm;
  double* im = sm + L * strided_line_doubles<N>();
  const int tiles = (H2 + L - 1) / L;
  const int o = blockIdx.x / tiles;
  const int k0 = (blockIdx.x - o * tiles) * L;
  const int nl = min(L, H2 - k0);
  int r0, r1;
  bool two;
  if constexpr (MODE == CONV || MODE == CONV2) {
    r0 = o;
    r1 = M1 - o;
    two = (o != 0 && 2 * o != M1);
  } else {
    r0 = 2 * o;
    r1 = 2 * o + 1;
    two = r1 < outers;
  }
  double2* base0 = spec + (int64_t)r0 * outer_stride + k0;
  double2* base1 = spec + (int64_t)r1 * outer_stride + k0;
  double ar[P0::B0][P0::R0], ai[P0::B0][P0::R0], br[P0::B0][P0::R0], bi[P0::B0][P0::R0];
  strided_load<MODE, N>(base0, nl, n_in, k0, line_stride, ar, ai);
  if (two) strided_load<MODE, N>(base1, nl, n_in, k0, line_stride, br, bi);
  double g[PF::BL][PF::RL];
  if constexpr (MODE == CONV) {  // Green values of (ky = o, this kz tile), shared by both rows
#pragma unroll
    for (int q = 0; q < PF::BL; q++) {
      const int b = (int)threadIdx.x + q * NT;
      const int line = b % L, j = b / L;
      const bool ok = b < L * PF::NBL && line < nl;
#pragma unroll
      for (int m = 0; m < PF::RL; m++) {
        const int p = j + m * PF::NBL;
        const int gx = (p <= N / 2) ? p : N - p;
        g[q][m] = ok ? ghat[((int64_t)gx * qy + o) * H2p + k0 + line] : 0.0;
      }
    }
  }
  if constexpr (MODE == CONV2) {  // PoissonSolverFFT2D.cc:215-228: rescaled to the rho step
#pragma unroll
    for (int q = 0; q < PF::BL; q++) {
      const int b = (int)threadIdx.x + q * NT;
      const int j = b / L;
      const bool ok = b < L * PF::NBL;
#pragma unroll
      for (int m = 0; m < PF::RL; m++) {
        const int p = j + m * PF::NBL;
        const int gx = (p <= N / 2) ? p : N - p;
        double v = ok ? ghat[(int64_t)gx * qy + o] : 0.0;
        if (gx == 0 && o == 0) v = v + s_coeff * npad;
        g[q][m] = v - s_coeff;
      }
    }
  }
  // GREEN: n_out passes the real lines from this tile's first pair on
  const int nout = MODE == GREEN ? n_in - 2 * k0 : n_out;
  strided_compute<MODE, N>(re, im, base0, nl, nout, line_stride, tw, nmax, g, ar, ai);
  if (two) {
    __syncthreads();  // LDS reads of the first row are done
    strided_compute<MODE, N>(re, im, base1, nl, nout, line_stride, tw, nmax, g, br, bi);
  }
}

// ---- Green spectrum (PoissonSolverFFT3D.cc:160-247) ----------------------------------
// 1/r on the padded grid is real and even along every axis, so its spectrum is real and
// even: it is the product of three even-symmetric transforms over the quadrant
// i <= M0/2, j <= M1/2, k <= M2/2.  Two real even lines share one complex FFT: the
// transform of a + i b is A + i B with A, B real.  ghat[(i*qy + j)*H2p + k].
template <int N>  // N = M2
__global__ __launch_bounds__(kThreads) void green3d_z(int qx, int qy, int H2p,
                                                      double* __restrict__ ghat, GreenGeom gg,
                                                      const double2* __restrict__ tw, int nmax) {
  constexpr int P = zpass_pairs(N), ls = fft_line_stride(N), H2 = N / 2 + 1;
  extern __shared__ double sm[];
  double* re = sm;
  double* im = sm + P * ls;
  const int64_t nlines = (int64_t)qx * qy;
  const int64_t pair0 = (int64_t)blockIdx.x * P;
  for (int idx = threadIdx.x; idx < P * N; idx += kThreads) {
    const int lp = idx / N, k = idx - lp * N;
    const int64_t q0 = 2 * (pair0 + lp), q1 = q0 + 1;
    int i0, j0, i1, j1;
    line_ij(q0, qy, i0, j0);
    line_next(qy, i0, j0, i1, j1);
    re[lp * ls + k] = q0 < nlines ? green3d(i0, j0, k, gg) : 0.0;
    im[lp * ls + k] = q1 < nlines ? green3d(i1, j1, k, gg) : 0.0;
  }
  __syncthreads();
  lds_fft<N, P, kThreads, -1>(re, im, P, ls, tw, nmax);
  for (int idx = threadIdx.x; idx < P * H2; idx += kThreads) {
    const int lp = idx / H2, k = idx - lp * H2;
    const int64_t q0 = 2 * (pair0 + lp), q1 = q0 + 1;
    if (q0 < nlines) ghat[q0 * H2p + k] = re[lp * ls + k];
    if (q1 < nlines) ghat[q1 * H2p + k] = im[lp * ls + k];
  }
}

// In-place even-symmetric transform along one axis of the real quadrant array: line
// (o, kz) holds N/2+1 samples at ghat[o*outer_stride + p*line_stride + kz], extended
// evenly to N points; its first N/2+1 spectral values replace them.
template <int N>
__global__ __launch_bounds__(kThreads) void green3d_axis(double* __restrict__ ghat,
                                                         int64_t line_stride,
                                                         int64_t outer_stride, int H2,
                                                         const double2* __restrict__ tw,
                                                         int nmax) {
  constexpr int L = strided_lines(N), C = L / 2, ls = fft_line_stride(N), Q = N / 2 + 1;
  extern __shared__ double sm[];
  double* re = sm;
  double* im = sm + C * ls;
  const int tiles = (H2 + L - 1) / L;
  const int o = blockIdx.x / tiles;
  const int k0 = (blockIdx.x - o * tiles) * L;
  const int nl = min(L, H2 - k0);
  double* base = ghat + (int64_t)o * outer_stride + k0;
  // loads issued together before the stores, as in fft3_z_r2c
  constexpr int IT = (L * Q + kThreads - 1) / kThreads;
  double av[IT];
  uint32_t msk = 0;
#pragma unroll
  for (int u = 0; u < IT; u++) {
    const int idx = (int)threadIdx.x + u * kThreads;
    const int p = idx / L, l = idx - p * L;
    const bool ok = idx < L * Q && l < nl;
    av[u] = base[ok ? (int64_t)p * line_stride + l : 0];
    msk |= (ok ? 1u : 0u) << u;
  }
#pragma unroll
  for (int u = 0; u < IT; u++) {
    const int idx = (int)threadIdx.x + u * kThreads;
    if (idx < L * Q) {
      const int p = idx / L, l = idx - p * L;
      const double v = ((msk >> u) & 1u) ? av[u] : 0.0;
      double* d = ((l & 1) ? im : re) + (l >> 1) * ls;
      d[p] = v;
      if (p > 0 && p < N / 2) d[N - p] = v;
    }
  }
  __syncthreads();
  lds_fft<N, C, kThreads, -1>(re, im, (nl + 1) >> 1, ls, tw, nmax);
  for (int idx = threadIdx.x; idx < L * Q; idx += kThreads) {
    const int p = idx / L, l = idx - p * L;
    if (l < nl) base[(int64_t)p * line_stride + l] = ((l & 1) ? im : re)[(l >> 1) * ls + p];
  }
}

}  // namespace
}  // namespace ob

struct ob_sc3d {
  int nx, ny, nz, M0, M1, M2, H2, H2p, H2g, qx, qy, nmax;
  double ratio_limit = 1.2, xy_ratio = 1.0, xz_ratio = 1.0;
  double sMin[3], sMax[3], sd[3];  // solver grid limits and steps (PoissonSolver3D.cc)
  double g[9];                     // rho/phi grid: xMin,xMax,yMin,yMax,zMin,zMax (rest), dx,dy,dz
  int rebuilds = 0;
  double *rho = nullptr, *phi = nullptr, *ghat = nullptr, *partials = nullptr;
  ob::ExtremaLink ext;             // bunch extrema to the host
  double2 *spec = nullptr, *tw = nullptr;
  int nbunches = 0;                // external bunches (PoissonSolverFFT3D nBunches_, even)
  double frequency = 402.5e+6;     // bunch frequency (SpaceChargeCalc3D.cc:58)
  double lambda_cm = 0.0;          // image spacing used by the current Green function
  bool phi_wrap = false;           // phiGrid longWrapping (set by wrapped calls, sticky)
  bool tiled = true;              // tile-sorted LDS deposit (else particle-order atomics)
  int lds_budget = 16 * 1024;     // LDS bytes for one deposit / kick tile
  ob::TileWork work;
  hipEvent_t ev_main = nullptr, ev_green = nullptr;  // Green rebuild on the second stream
  bool green_pending = false;      // ev_green recorded, not yet waited on by the main stream
  bool green_launch = false;       // box changed, Green kernels not yet enqueued
};

namespace ob {
namespace {

// Dynamic LDS above 64 KiB has to be allowed per kernel; done once per instantiation.
template <typename K>
void allow_lds(K* kernel, size_t bytes) {
  if (bytes > 64 * 1024)
    OB_HIP(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)bytes));
}

template <int N>
size_t zpass_lds() { return sizeof(double) * 2 * zpass_pairs(N) * fft_line_stride(N); }
template <int N>
size_t strided_lds() { return sizeof(double) * 2 * strided_lines(N) * strided_line_doubles<N>(); }

template <int MODE, int N>
void launch_strided_n(ob_sc3d* c, int n_in, int n_out, int64_t line_stride,
                      int64_t outer_stride, int outers, const char* label) {
  static bool ready = false;
  const size_t lds = strided_lds<N>();
  if (!ready) {
    allow_lds(fft3_strided<MODE, N>, lds);
    ready = true;
  }
  const int tiles = (c->H2 + strided_lines(N) - 1) / strided_lines(N);
  OB_REQUIRE(MODE == INV || 2 * n_in <= N, "strided FFT: input longer than half the line");
  OB_REQUIRE(MODE == FWD || 2 * n_out <= N, "strided FFT: output longer than half the line");
  // x pass: ky pairs (o, M1-o), o <= M1/2; y passes: row pairs (2o, 2o+1)
  const int blocks_outer = (MODE == CONV) ? c->M1 / 2 + 1 : (outers + 1) / 2;
  ob::ProfScope ps_(label);
  hipLaunchKernelGGL((fft3_strided<MODE, N>), dim3(blocks_outer * tiles), dim3(strided_threads(N)),
                     lds, stream(), c->spec, n_in, n_out, line_stride, outer_stride, outers, c->H2,
                     c->ghat, c->qy, c->H2g, c->M1, c->tw, c->nmax, 0.0, 0.0);
  OB_HIP(hipGetLastError());
}

template <int MODE>
void launch_strided(ob_sc3d* c, int n, int n_in, int n_out, int64_t line_stride,
                    int64_t outer_stride, int outers, const char* label) {
  dispatch_pow2(n, [&](auto NC) {
    constexpr int N = decltype(NC)::value;
    if constexpr (N >= 8 && N <= 2048)
      launch_strided_n<MODE, N>(c, n_in, n_out, line_stride, outer_stride, outers, label);
    else
      throw Error{"SpaceChargeCalc3D: transverse FFT length out of range"};
  });
}

// PoissonSolverFFT3D::setGridXYZ (:136-158): the solver box and steps, one more rebuild
void set_solver_box(ob_sc3d* c, double xMin, double xMax, double yMin, double yMax, double zMin,
                    double zMax) {
  c->sMin[0] = xMin; c->sMax[0] = xMax; c->sd[0] = (xMax - xMin) / (c->nx - 1);
  c->sMin[1] = yMin; c->sMax[1] = yMax; c->sd[1] = (yMax - yMin) / (c->ny - 1);
  c->sMin[2] = zMin; c->sMax[2] = zMax; c->sd[2] = (zMax - zMin) / c->nz;
  c->rebuilds++;
}

// _defineGreenF (:160-247) on the current box, enqueued on s
void build_green(ob_sc3d* c, hipStream_t s) {
  GreenGeom gg{c->M0, c->M1, c->M2, c->sd[0], c->sd[1], c->sd[2], c->nbunches, c->lambda_cm};
  dispatch_pow2(c->M2, [&](auto NC) {
    constexpr int N = decltype(NC)::value;
    if constexpr (N <= 2048) {
      static bool ready = false;
      const size_t lds = zpass_lds<N>();
      if (!ready) {
        allow_lds(green3d_z<N>, lds);
        ready = true;
      }
      const int64_t pairs = ((int64_t)c->qx * c->qy + 1) / 2;
      const int64_t blocks = (pairs + zpass_pairs(N) - 1) / zpass_pairs(N);
      ob::ProfScope ps_("green3d_z", s);
      hipLaunchKernelGGL(green3d_z<N>, dim3((unsigned)blocks), dim3(kThreads), lds, s, c->qx,
                         c->qy, c->H2g, c->ghat, gg, c->tw, c->nmax);
    }
  });
  OB_HIP(hipGetLastError());
  // The y and x axis transforms: complex lines are kz pairs, strides in double2 units
  // (H2g is a multiple of 16, so a pair never straddles a row; the padding is zero).
  auto axis = [&](int n, int64_t line_stride, int64_t outer_stride, int outers) {
    dispatch_pow2(n, [&](auto NC) {
      constexpr int N = decltype(NC)::value;
      if constexpr (N >= 8 && N <= 2048) {
        static bool ready = false;
        const size_t lds = strided_lds<N>();
        if (!ready) {
          allow_lds(fft3_strided<GREEN, N>, lds);
          ready = true;
        }
        const int pairs = (c->H2 + 1) / 2;
        const int tiles = (pairs + strided_lines(N) - 1) / strided_lines(N);
        ob::ProfScope ps_("green3d_axis", s);
        hipLaunchKernelGGL((fft3_strided<GREEN, N>), dim3(((outers + 1) / 2) * tiles),
                           dim3(strided_threads(N)), lds, s, (double2*)c->ghat, c->H2, 0,
                           line_stride, outer_stride, outers, pairs, nullptr, 0, 0, 0, c->tw,
                           c->nmax, 0.0, 0.0);
      }
    });
    OB_HIP(hipGetLastError());
  };
  const int64_t H2q = c->H2g / 2;
  axis(c->M1, H2q, (int64_t)c->qy * H2q, c->qx);  // y: lines (i, kz pair) along j
  axis(c->M0, (int64_t)c->qy * H2q, H2q, c->qy);  // x: lines (j, kz pair) along i
}

// PoissonSolverFFT3D::setGridXYZ + _defineGreenF (:136-247), enqueued on s
void set_solver_grid(ob_sc3d* c, double xMin, double xMax, double yMin, double yMax, double zMin,
                     double zMax, hipStream_t s) {
  c->green_launch = false;  // supersedes a deferred rebuild of an earlier box
  set_solver_box(c, xMin, xMax, yMin, yMax, zMin, zMax);
  build_green(c, s);
}

// Rebuild the Green function on the second stream: it is first read by the x pass of the
// next solve, so it overlaps the deposit and the first passes; the previous solves, which
// read the old one, are enqueued before it starts (ev_main, recorded here).  The box is
// set at once (the step's grid depends on it); the kernels are enqueued by
// launch_pending_green after the step's deposit, so the host time spent enqueuing them
// overlaps the tile sort instead of delaying it.
void set_solver_grid_async(ob_sc3d* c, double xMin, double xMax, double yMin, double yMax,
                           double zMin, double zMax) {
  OB_HIP(hipEventRecord(c->ev_main, stream()));
  set_solver_box(c, xMin, xMax, yMin, yMax, zMin, zMax);
  c->green_launch = true;
}

void launch_pending_green(ob_sc3d* c) {
  if (!c->green_launch) return;
  c->green_launch = false;
  OB_HIP(hipStreamWaitEvent(aux_stream(), c->ev_main, 0));
  build_green(c, aux_stream());
  OB_HIP(hipEventRecord(c->ev_green, aux_stream()));
  c->green_pending = true;
}

void solve(ob_sc3d* c, double denom) {
  OB_REQUIRE((int64_t)c->nx * c->ny < ((int64_t)1 << 31), "SpaceChargeCalc3D: grid too large");
  launch_pending_green(c);
  hipStream_t s = stream();
  const int64_t nl = (int64_t)c->nx * c->ny;
  const int64_t H2p = c->H2p;
  dispatch_pow2(c->M2, [&](auto NC) {
    constexpr int N = decltype(NC)::value;
    if constexpr (N <= 2048) {
      static bool ready = false;
      const size_t lds = zpass_lds<N>();
      if (!ready) {
        allow_lds(fft3_z_r2c<N>, lds);
        ready = true;
      }
      const int64_t blocks = ((nl + 1) / 2 + zpass_pairs(N) - 1) / zpass_pairs(N);
      ob::ProfScope ps_("fft3_z_r2c");
      hipLaunchKernelGGL(fft3_z_r2c<N>, dim3((unsigned)blocks), dim3(kThreads), lds, s, c->rho,
                         nl, c->ny, c->nz, c->M1, c->H2p, c->spec, c->tw, c->nmax);
    }
  });
  OB_HIP(hipGetLastError());
  launch_strided<FWD>(c, c->M1, c->ny, c->M1, H2p, (int64_t)c->M1 * H2p, c->nx,
                      "fft3_strided_fwd");
  if (c->green_pending) {
    OB_HIP(hipStreamWaitEvent(s, c->ev_green, 0));
    c->green_pending = false;
  }
  launch_strided<CONV>(c, c->M0, c->nx, c->nx, (int64_t)c->M1 * H2p, H2p, c->M1,
                       "fft3_strided_conv");
  launch_strided<INV>(c, c->M1, c->M1, c->ny, H2p, (int64_t)c->M1 * H2p, c->nx,
                      "fft3_strided_inv");
  dispatch_pow2(c->M2, [&](auto NC) {
    constexpr int N = decltype(NC)::value;
    if constexpr (N <= 2048) {
      static bool ready = false;
      const size_t lds = zpass_lds<N>();
      if (!ready) {
        allow_lds(fft3_z_c2r<N>, lds);
        ready = true;
      }
      const int64_t blocks = ((nl + 1) / 2 + zpass_pairs(N) - 1) / zpass_pairs(N);
      ob::ProfScope ps_("fft3_z_c2r");
      hipLaunchKernelGGL(fft3_z_c2r<N>, dim3((unsigned)blocks), dim3(kThreads), lds, s, c->spec,
                         nl, c->ny, c->M1, c->H2p, c->nz, denom, c->phi, c->tw, c->nmax);
    }
  });
  OB_HIP(hipGetLastError());
}

}  // namespace
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

OB_API int ob_sc3d_create(int nx, int ny, int nz, ob_sc3d** out) {
  OB_TRY OB_REQUIRE(out != nullptr, "null output pointer");
  OB_REQUIRE(nx >= 3 && ny >= 3, "SpaceChargeCalc3D: grid size too small (should be >= 3)");
  OB_REQUIRE(nz >= 1, "SpaceChargeCalc3D: number of z bins should be >= 1");
  ob::ensure_init();
  ob_sc3d* c = new ob_sc3d();
  c->nx = nx; c->ny = ny; c->nz = nz;
  c->M0 = ob::pad_len(nx);
  c->M1 = ob::pad_len(ny);
  c->M2 = ob::pad_len(nz);
  if (c->M0 > 2048 || c->M1 > 2048 || c->M2 > 2048) {
    delete c;
    throw ob::Error{"SpaceChargeCalc3D: grid larger than 1024 points per axis"};
  }
  c->H2 = c->M2 / 2 + 1;
  c->H2p = (c->H2 + 7) & ~7;  // kz rows padded to 8 complex: 128-byte aligned tile rows
  c->H2g = (c->H2 + 15) & ~15;  // Green rows padded to 16 reals: 128-byte aligned as well
  c->qx = c->M0 / 2 + 1;
  c->qy = c->M1 / 2 + 1;
  c->nmax = std::max(c->M0, std::max(c->M1, c->M2));
  const size_t ng = (size_t)nx * ny * nz;
  OB_HIP(hipMalloc((void**)&c->rho, sizeof(double) * ng));
  OB_HIP(hipMalloc((void**)&c->phi, sizeof(double) * ng));
  // spectrum after the y pass: x < nx only (the x pass reads and writes just those planes)
  OB_HIP(hipMalloc((void**)&c->spec, sizeof(double2) * (size_t)nx * c->M1 * c->H2p));
  OB_HIP(hipMalloc((void**)&c->ghat, sizeof(double) * (size_t)c->qx * c->qy * c->H2g));
  OB_HIP(hipMalloc((void**)&c->partials, sizeof(double) * 6 * 4096));
  ob::extrema_link_create(c->ext);
  OB_HIP(hipMemset(c->rho, 0, sizeof(double) * ng));
  OB_HIP(hipMemset(c->phi, 0, sizeof(double) * ng));
  OB_HIP(hipMemset(c->spec, 0, sizeof(double2) * (size_t)nx * c->M1 * c->H2p));
  OB_HIP(hipMemset(c->ghat, 0, sizeof(double) * (size_t)c->qx * c->qy * c->H2g));
  c->tw = ob::make_twiddles(c->nmax);
  // SpaceChargeCalc3D(:32-44): solver on [-1,1]^3 (xy_ratio = xz_ratio = 1); rho/phi grids
  // take the solver limits.
  OB_HIP(hipEventCreateWithFlags(&c->ev_main, hipEventDisableTiming));
  OB_HIP(hipEventCreateWithFlags(&c->ev_green, hipEventDisableTiming));
  ob::set_solver_grid(c, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0, ob::stream());
  c->rebuilds = 0;
  for (int k = 0; k < 3; k++) {
    c->g[2 * k] = c->sMin[k];
    c->g[2 * k + 1] = c->sMax[k];
    c->g[6 + k] = c->sd[k];
  }
  OB_HIP(hipStreamSynchronize(ob::stream()));
  *out = c;
  OB_CATCH
}

OB_API int ob_sc3d_destroy(ob_sc3d* c) {
  OB_TRY if (!c) return 0;
  // Pending kicks read phi: apply them before the buffers go.
  std::string err;
  try {
    ob::flush_kicks_reading(c->phi);
  } catch (const ob::Error& e) {
    err = e.msg;
  }
  (void)hipStreamSynchronize(ob::stream());
  (void)hipStreamSynchronize(ob::aux_stream());
  if (c->ev_main) (void)hipEventDestroy(c->ev_main);
  if (c->ev_green) (void)hipEventDestroy(c->ev_green);
  (void)hipFree(c->rho); (void)hipFree(c->phi); (void)hipFree(c->spec); (void)hipFree(c->ghat);
  (void)hipFree(c->partials); (void)hipFree(c->tw);
  ob::extrema_link_destroy(c->ext);
  c->work.release();
  delete c;
  OB_REQUIRE(err.empty(), err);
  OB_CATCH
}

OB_API int ob_sc3d_track(ob_sc3d* c, ob_bunch* b, double length) {
  OB_TRY OB_REQUIRE(c != nullptr && b != nullptr, "null handle");
  ob::check_statuses(false);
  const int64_t n = ob::bunch_size(b);
  if (n < 2) return 0;  // :102-103
  hipStream_t s = ob::stream();
  const int nblk = ob::bunch_flush(b, c->partials);
  const double* ext = ob::extrema_to_host(c->ext, c->partials, nblk, s);
  const ob::SyncInfo sy = ob::bunch_sync(b);
  const double gamma = sy.gamma, beta = sy.beta;
  double xMin = ext[0], xMax = ext[1], yMin = ext[2];
  double yMax = ext[3], zMin = ext[4], zMax = ext[5];
  double center, width;
  const bool wrapped = c->nbunches > 0;
  double lambda = 0.0, lambda_cm = 0.0;
  ob::Grid3 g;
  g.nx = c->nx; g.ny = c->ny; g.nz = c->nz;
  if (wrapped) {
    // wrappedBunchAnalysis (:251-328), scalar arithmetic as in the reference
    c->phi_wrap = true;
    center = (zMax + zMin) / 2.0;
    width = (zMax - zMin) / 2.0;
    zMin = center - width;
    zMax = center + width;
    lambda = ob::kSpeedOfLight * beta / c->frequency;
    lambda_cm = ob::kSpeedOfLight * beta * gamma / c->frequency;
    if (fabs(zMax) > lambda / 2 || fabs(zMin) > lambda / 2) {
      const double extension_coeff = 1.0001;
      center = 0.;
      width = extension_coeff * lambda / 2;
      zMin = center - width;
      zMax = center + width;
    }
    OB_REQUIRE(!(xMin >= xMax || yMin >= yMax || zMin >= zMax),
               "SpaceChargeCalc3D::wrappedBunchAnalysis: the bunch min and max sizes are wrong "
               "(zero extent); cannot calculate space charge");
    g.xMin = xMin; g.xMax = xMax; g.dx = (g.xMax - g.xMin) / (c->nx - 1);
    g.yMin = yMin; g.yMax = yMax; g.dy = (g.yMax - g.yMin) / (c->ny - 1);
    g.zMin = zMin; g.zMax = zMax; g.dz = (g.zMax - g.zMin) / c->nz;
  } else {
    // bunchAnalysis (:150-249), scalar arithmetic as in the reference
    center = (zMax + zMin) / 2.0;
    width = (zMax - zMin) * gamma / 2.0;
    zMin = center - width;
    zMax = center + width;
    OB_REQUIRE(!(xMin >= xMax || yMin >= yMax || zMin >= zMax),
               "SpaceChargeCalc3D::bunchAnalysis: the bunch min and max sizes are wrong (zero "
               "extent); cannot calculate space charge");
    const double xy_ratio_beam = (xMax - xMin) / (yMax - yMin);
    const double xz_ratio_beam = (xMax - xMin) / (zMax - zMin);
    if (c->ratio_limit < xy_ratio_beam / c->xy_ratio || 1.0 / c->ratio_limit > xy_ratio_beam / c->xy_ratio ||
        c->ratio_limit < xz_ratio_beam / c->xz_ratio || 1.0 / c->ratio_limit > xz_ratio_beam / c->xz_ratio) {
      c->xy_ratio = xy_ratio_beam;
      c->xz_ratio = xz_ratio_beam;
      ob::set_solver_grid_async(c, -(xMax - xMin) / 2.0, (xMax - xMin) / 2.0, -(yMax - yMin) / 2.0,
                          (yMax - yMin) / 2.0, -(zMax - zMin) / 2.0, (zMax - zMin) / 2.0);
    }
    const double scx = (xMax - xMin) / (c->sMax[0] - c->sMin[0]);
    const double scy = (yMax - yMin) / (c->sMax[1] - c->sMin[1]);
    const double scz = (zMax - zMin) / (c->sMax[2] - c->sMin[2]);
    double sc = scx;
    if (sc < scy) sc = scy;
    if (sc < scz) sc = scz;
    center = (xMax + xMin) / 2.0;
    width = (c->sMax[0] - c->sMin[0]) * sc / 2.0;
    g.xMin = center - width; g.xMax = center + width; g.dx = (g.xMax - g.xMin) / (c->nx - 1);
    center = (yMax + yMin) / 2.0;
    width = (c->sMax[1] - c->sMin[1]) * sc / 2.0;
    g.yMin = center - width; g.yMax = center + width; g.dy = (g.yMax - g.yMin) / (c->ny - 1);
    center = (zMax + zMin) / 2.0;
    width = (c->sMax[2] - c->sMin[2]) * sc / (2.0 * gamma);
    g.zMin = center - width; g.zMax = center + width; g.dz = (g.zMax - g.zMin) / c->nz;
  }
  // deposit in lab z
  const int64_t ng = (int64_t)c->nx * c->ny * c->nz;
  {
    ob::ProfScope ps_("fill_zero");
    hipLaunchKernelGGL(ob::fill_zero, dim3(1024), dim3(256), 0, s, c->rho, ng);
  }
  const ob::TileGeom tg = ob::make_tile_geom(g.nx, g.ny, g.nz, g.xMin, g.xMax, g.yMin, g.yMax,
                                             g.zMin, g.zMax, g.dx, g.dy, g.dz, c->lds_budget);
  const bool tiled = c->tiled && tg.T > 0 && !wrapped;
  if (tiled) {
    ob::sort_and_deposit_3d(b, tg, sy.macrosize, c->rho, c->work);
  } else {  // particle-order deposit with global atomics
    ob::DeviceBunchView v = ob::bunch_device_view(b);
    int dblk = (int)((n + ob::kThreads - 1) / ob::kThreads);
    const int cap = 16 * ob::device_cu_count();
    if (dblk > cap) dblk = cap;
    ob::ProfScope ps_("deposit3d");
    hipLaunchKernelGGL(ob::deposit3d, dim3(dblk), dim3(ob::kThreads), 0, s, v, g, sy.macrosize,
                       lambda, c->rho);
  }
  OB_HIP(hipGetLastError());
  // move the z grid to the rest frame (:242-248)
  center = (g.zMax + g.zMin) / 2.0;
  width = (g.zMax - g.zMin) * gamma / 2.0;
  const double rzMin = center - width, rzMax = center + width;
  const double rdz = (rzMax - rzMin) / c->nz;
  c->g[0] = g.xMin; c->g[1] = g.xMax; c->g[2] = g.yMin; c->g[3] = g.yMax;
  c->g[4] = rzMin; c->g[5] = rzMax; c->g[6] = g.dx; c->g[7] = g.dy; c->g[8] = rdz;
  if (wrapped) {  // the solver takes the rho grid and the image spacing (:326-330)
    c->lambda_cm = lambda_cm;
    ob::set_solver_grid(c, g.xMin, g.xMax, g.yMin, g.yMax, rzMin, rzMax, ob::stream());
  }
  // findPotential (PoissonSolverFFT3D.cc:249-347)
  const double scale_coeff = c->sd[0] / g.dx;
  const double denom = scale_coeff / ((double)c->M0 * c->M1 * c->M2);
  ob::solve(c, denom);
  // kick (:120-147): over the deposit's tiles when the bunch was tile-sorted, else fused
  // into the next map kernel
  if (tiled && !c->phi_wrap) {
    const ob::Kick3D k{c->phi, rzMin, rdz, (rzMax + rzMin) / 2.0, gamma,
                       length * sy.r0 / ((beta * beta) * (gamma * gamma)),  // :121
                       length * sy.r0 * sy.mass};                           // :122
    ob::tile_kick_3d(b, tg, k, c->work);
    return 0;
  }
  ob::Op op;
  std::memset(&op, 0, sizeof(op));
  op.kind = ob::OP_SC3D_KICK;
  op.i0 = c->nx; op.i1 = c->ny; op.i2 = c->nz | (c->phi_wrap ? ob::kSC3DWrapBit : 0);
  op.p[0] = ob::ptr_bits(c->phi);
  op.p[1] = g.xMin; op.p[2] = g.dx; op.p[3] = g.yMin; op.p[4] = g.dy;
  op.p[5] = rzMin; op.p[6] = rdz;
  op.p[7] = (rzMax + rzMin) / 2.0;                                   // :125
  op.p[8] = gamma;
  op.p[9] = length * sy.r0 / ((beta * beta) * (gamma * gamma));     // :121
  op.p[10] = length * sy.r0 * sy.mass;                               // :122
  op.p[11] = wrapped ? ob::kSpeedOfLight * beta * gamma / c->frequency : 0.0;  // :123, 138-140
  ob::bunch_queue(b, op);
  OB_CATCH
}

OB_API int ob_sc3d_get_grid(ob_sc3d* c, int which, double* out, double* lims) {
  OB_TRY OB_REQUIRE(c != nullptr, "null handle");
  OB_REQUIRE(which == 0 || which == 1, "grid selector must be 0 (rho) or 1 (phi)");
  const size_t ng = (size_t)c->nx * c->ny * c->nz;
  if (out) {
    std::vector<double> tmp(ng);
    OB_HIP(hipMemcpyAsync(tmp.data(), which == 0 ? c->rho : c->phi, sizeof(double) * ng,
                          hipMemcpyDeviceToHost, ob::stream()));
    OB_HIP(hipStreamSynchronize(ob::stream()));
    // device [ix][iy][iz] -> reference Arr3D[iz][ix][iy]
    for (int ix = 0; ix < c->nx; ix++)
      for (int iy = 0; iy < c->ny; iy++)
        for (int iz = 0; iz < c->nz; iz++)
          out[((size_t)iz * c->nx + ix) * c->ny + iy] = tmp[((size_t)ix * c->ny + iy) * c->nz + iz];
  }
  if (lims)
    for (int k = 0; k < 6; k++) lims[k] = c->g[k];
  OB_CATCH
}

OB_API int ob_sc3d_set_ratio_limit(ob_sc3d* c, double r) {
  OB_TRY OB_REQUIRE(c != nullptr, "null handle");
  c->ratio_limit = r;
  OB_CATCH
}
OB_API int ob_sc3d_get_ratio_limit(ob_sc3d* c, double* r) {
  OB_TRY OB_REQUIRE(c != nullptr, "null handle");
  *r = c->ratio_limit;
  OB_CATCH
}
OB_API int ob_sc3d_green_rebuilds(ob_sc3d* c, int* n) {
  OB_TRY OB_REQUIRE(c != nullptr, "null handle");
  *n = c->rebuilds;
  OB_CATCH
}

OB_API int ob_sc3d_set_num_ext_bunches(ob_sc3d* c, int n) {
  OB_TRY OB_REQUIRE(c != nullptr, "null handle");
  OB_REQUIRE(n >= 0, "number of external bunches must be >= 0");
  if (n % 2 != 0) n = n + 1;  // PoissonSolverFFT3D.cc:101-106
  c->nbunches = n;
  OB_CATCH
}
OB_API int ob_sc3d_get_num_ext_bunches(ob_sc3d* c, int* n) {
  OB_TRY OB_REQUIRE(c != nullptr && n != nullptr, "null handle");
  *n = c->nbunches;
  OB_CATCH
}
OB_API int ob_sc3d_set_freq(ob_sc3d* c, double f) {
  OB_TRY OB_REQUIRE(c != nullptr, "null handle");
  c->frequency = f;
  OB_CATCH
}
OB_API int ob_sc3d_get_freq(ob_sc3d* c, double* f) {
  OB_TRY OB_REQUIRE(c != nullptr && f != nullptr, "null handle");
  *f = c->frequency;
  OB_CATCH
}

OB_API int ob_sc3d_set_deposit_mode(ob_sc3d* c, int tiled, int lds_budget_bytes) {
  OB_TRY OB_REQUIRE(c != nullptr, "null handle");
  // (the deposit stages its work item's particles beside the tile block: 40 KiB more)
  OB_REQUIRE(lds_budget_bytes == 0 || (lds_budget_bytes >= 1024 && lds_budget_bytes <= 116 * 1024),
             "LDS budget must be 0 (keep) or 1 KiB..116 KiB");
  c->tiled = tiled != 0;
  if (lds_budget_bytes) c->lds_budget = lds_budget_bytes;
  OB_CATCH
}

// ---------------------------------------------------------------------------------------
// SpaceChargeCalcSliceBySlice2D (SpaceChargeCalcSliceBySlice2D.cc): the 3D TSC deposit of
// SpaceChargeCalc3D (lab frame), then one 2D Hockney solve per z slice
// (PoissonSolverFFT2D.cc:174-256), then the transverse kick from the 3D gradient of phi.
// The nz slice solves run as one stack: the real slices 2l and 2l+1 are the real and
// imaginary parts of complex line l (the 2D Green spectrum is real and even, so the two
// solutions stay apart), and the stack goes through the batched y / x / y passes of the
// 3D solver with a 2D Green spectrum.
struct ob_slice2d {
  int nx, ny, nz, M0, M1, L2, H2p, qx, qy, nmax;
  double xy_ratio;
  double sol[4];         // solver box (PoissonSolverFFT2D xMin_, xMax_, yMin_, yMax_)
  double sdx, sdy;       // solver steps
  int rebuilds = 0;
  bool green_valid = false;
  double g[9];           // last grid: xMin, xMax, yMin, yMax, zMin, zMax, dx, dy, dz
  double *rho = nullptr, *phi = nullptr, *ghat = nullptr, *partials = nullptr;
  ob::ExtremaLink ext;   // bunch extrema to the host
  double2 *spec = nullptr, *tw = nullptr;
  ob::SC2DState* geo = nullptr;    // grid geometry for the boundary kernels (device)
  ob::SC2DState* geo_host = nullptr;  // pinned staging of it
  ob::TileWork work;
  int lds_budget = 16 * 1024;
};

namespace ob {
namespace {

// Quadrant i <= M0/2, j <= M1/2 of the 2D Green function -ln(r^2)/2 on the solver steps
// (PoissonSolverFFT2D::_defineGreenF, :115-150), stored ghat[i * qy + j].
__global__ void green2d_quadrant(double* __restrict__ q, int qx, int qy, double dx, double dy) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)qx * qy) return;
  const int i = (int)(t / qy), j = (int)(t - (int64_t)i * qy);
  const double rTransX = i * dx, rTransY = j * dy;
  const double rTot2 = rTransX * rTransX + rTransY * rTransY;
  q[t] = (i != 0 || j != 0) ? -log(rTot2) / 2 : 0.0;
}

// rho slices 2l, 2l+1 -> complex line l of spec[x][y][l] (x < nx, y < ny).
__global__ void slice_pack(const double* __restrict__ rho, int nx, int ny, int nz, int L2,
                           int H2p, int M1, double2* __restrict__ spec) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nx * ny * L2) return;
  const int l = (int)(t % L2);
  const int64_t r = t / L2;
  const int x = (int)(r / ny), y = (int)(r - (int64_t)x * ny);
  const double* src = rho + r * nz;
  const double a = src[2 * l];
  const double b = (2 * l + 1 < nz) ? src[2 * l + 1] : 0.0;
  spec[((int64_t)x * M1 + y) * H2p + l] = make_double2(a, b);
}

// complex line l -> phi slices 2l, 2l+1, scaled by 1 / (M0 M1) (:252-255).
__global__ void slice_unpack(const double2* __restrict__ spec, int nx, int ny, int nz, int L2,
                             int H2p, int M1, double denom, double* __restrict__ phi) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nx * ny * L2) return;
  const int l = (int)(t % L2);
  const int64_t r = t / L2;
  const int x = (int)(r / ny), y = (int)(r - (int64_t)x * ny);
  const double2 v = spec[((int64_t)x * M1 + y) * H2p + l];
  double* dst = phi + r * nz;
  dst[2 * l] = denom * v.x;
  if (2 * l + 1 < nz) dst[2 * l + 1] = denom * v.y;
}

template <int MODE>
void slice_strided(ob_slice2d* c, int n, int n_in, int n_out, int64_t line_stride,
                   int64_t outer_stride, int outers, double s_coeff, const char* label) {
  dispatch_pow2(n, [&](auto NC) {
    constexpr int N = decltype(NC)::value;
    if constexpr (N >= 8 && N <= 2048) {
      static bool ready = false;
      const size_t lds = strided_lds<N>();
      if (!ready) {
        allow_lds(fft3_strided<MODE, N>, lds);
        ready = true;
      }
      const int tiles = (c->L2 + strided_lines(N) - 1) / strided_lines(N);
      const int blocks_outer = (MODE == CONV2) ? c->M1 / 2 + 1 : (outers + 1) / 2;
      ProfScope ps_(label);
      hipLaunchKernelGGL((fft3_strided<MODE, N>), dim3(blocks_outer * tiles),
                         dim3(strided_threads(N)), lds, stream(), c->spec, n_in, n_out,
                         line_stride, outer_stride, outers, c->L2, c->ghat, c->qy, c->H2p, c->M1,
                         c->tw, c->nmax, s_coeff, (double)c->M0 * (double)c->M1);
    } else {
      throw Error{"SpaceChargeCalcSliceBySlice2D: transverse FFT length out of range"};
    }
  });
  OB_HIP(hipGetLastError());
}

// PoissonSolverFFT2D::setGridXY (:103-111): the Green spectrum of the new box, as the
// even-symmetric transform of its quadrant along y, then along x.
void slice_set_solver(ob_slice2d* c, double xMin, double xMax, double yMin, double yMax) {
  c->sol[0] = xMin; c->sol[1] = xMax; c->sol[2] = yMin; c->sol[3] = yMax;
  c->sdx = (xMax - xMin) / (c->nx - 1);
  c->sdy = (yMax - yMin) / (c->ny - 1);
  hipStream_t s = stream();
  const int64_t tot = (int64_t)c->qx * c->qy;
  {
    ProfScope ps_("green2d_quadrant");
    hipLaunchKernelGGL(green2d_quadrant, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s,
                       c->ghat, c->qx, c->qy, c->sdx, c->sdy);
  }
  auto axis = [&](int n, int64_t line_stride, int64_t outer_stride, int outers, int lines) {
    dispatch_pow2(n, [&](auto NC) {
      constexpr int N = decltype(NC)::value;
      if constexpr (N >= 8 && N <= 2048) {
        static bool ready = false;
        const size_t lds = sizeof(double) * 2 * (strided_lines(N) / 2) * fft_line_stride(N);
        if (!ready) {
          allow_lds(green3d_axis<N>, lds);
          ready = true;
        }
        const int tiles = (lines + strided_lines(N) - 1) / strided_lines(N);
        ProfScope ps_("green2d_axis");
        hipLaunchKernelGGL(green3d_axis<N>, dim3(outers * tiles), dim3(kThreads), lds, s,
                           c->ghat, line_stride, outer_stride, lines, c->tw, c->nmax);
      }
    });
    OB_HIP(hipGetLastError());
  };
  axis(c->M1, 1, c->qy, c->qx, 1);    // y: one line per i
  axis(c->M0, c->qy, 0, 1, c->qy);    // x: lines j
  c->green_valid = true;
}

// One trackBunch (SpaceChargeCalcSliceBySlice2D.cc:66-113), boundary bnd or none.
void slice_track(ob_slice2d* c, ob_bunch* b, double length, ob_boundary2d* bnd) {
  check_statuses(false);
  const int64_t n = bunch_size(b);
  if (n < 2) return;  // :69-70
  hipStream_t s = stream();
  const int nblk = bunch_flush(b, c->partials);
  const double* ext = extrema_to_host(c->ext, c->partials, nblk, s);
  const SyncInfo sy = bunch_sync(b);
  double xMin = ext[0], xMax = ext[1], yMin = ext[2];
  double yMax = ext[3], zMin = ext[4], zMax = ext[5];
  double bnd6[6] = {0, 0, 0, 0, 0, 0};
  int bnd_type = 0;
  // bunchAnalysis (:116-218), scalar arithmetic as in the reference
  if (!bnd) {
    OB_REQUIRE(!(xMin >= xMax || yMin >= yMax || zMin >= zMax),
               "SpaceChargeCalcSliceBySlice2D::bunchAnalysis: the bunch min and max sizes are "
               "wrong (zero extent); cannot calculate space charge");
    const double xy_ratio_beam = (xMax - xMin) / (yMax - yMin);
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
      slice_set_solver(c, xMin, xMax, yMin, yMax);
      c->rebuilds++;
    }
  } else {
    bnd_type = boundary_describe(bnd, bnd6);
    xMin = bnd6[0]; xMax = bnd6[1]; yMin = bnd6[2]; yMax = bnd6[3];
    c->xy_ratio = (xMax - xMin) / (yMax - yMin);
    OB_REQUIRE(!(zMin >= zMax),
               "SpaceChargeCalcSliceBySlice2D::bunchAnalysis: the bunch min and max sizes are "
               "wrong (zero extent); cannot calculate space charge");
  }
  Grid3 g;
  g.nx = c->nx; g.ny = c->ny; g.nz = c->nz;
  g.xMin = xMin; g.xMax = xMax; g.dx = (xMax - xMin) / (c->nx - 1);  // Grid3D.cc:141-158
  g.yMin = yMin; g.yMax = yMax; g.dy = (yMax - yMin) / (c->ny - 1);
  g.zMin = zMin; g.zMax = zMax; g.dz = (zMax - zMin) / c->nz;         // :165-170
  if (fabs((c->sol[1] - c->sol[0]) / (c->sol[3] - c->sol[2]) - c->xy_ratio) > 0.00000001) {
    slice_set_solver(c, xMin, xMax, yMin, yMax);  // :203-209
    c->rebuilds++;
  }
  c->g[0] = g.xMin; c->g[1] = g.xMax; c->g[2] = g.yMin; c->g[3] = g.yMax;
  c->g[4] = g.zMin; c->g[5] = g.zMax; c->g[6] = g.dx; c->g[7] = g.dy; c->g[8] = g.dz;
  // Grid3D::binBunch (Grid3D.cc:281-305)
  const int64_t ng = (int64_t)c->nx * c->ny * c->nz;
  {
    ProfScope ps_("fill_zero");
    hipLaunchKernelGGL(fill_zero, dim3(1024), dim3(256), 0, s, c->rho, ng);
  }
  const TileGeom tg = make_tile_geom(g.nx, g.ny, g.nz, g.xMin, g.xMax, g.yMin, g.yMax, g.zMin,
                                     g.zMax, g.dx, g.dy, g.dz, c->lds_budget);
  const bool tiled = tg.T > 0;
  OB_REQUIRE(tiled || !bnd, "SpaceChargeCalcSliceBySlice2D: grid too large for the tiled kick "
                            "with a boundary");
  if (tiled) {
    sort_and_deposit_3d(b, tg, sy.macrosize, c->rho, c->work);
  } else {
    DeviceBunchView v = bunch_device_view(b);
    int dblk = (int)((n + kThreads - 1) / kThreads);
    const int cap = 16 * device_cu_count();
    if (dblk > cap) dblk = cap;
    ProfScope ps_("deposit3d");
    hipLaunchKernelGGL(deposit3d, dim3(dblk), dim3(kThreads), 0, s, v, g, sy.macrosize, 0.0,
                       c->rho);
  }
  OB_HIP(hipGetLastError());
  // one findPotential per slice (:77-90), as a stack
  const int64_t np = (int64_t)c->nx * c->ny * c->L2;
  {
    ProfScope ps_("slice_pack");
    hipLaunchKernelGGL(slice_pack, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s, c->rho,
                       c->nx, c->ny, c->nz, c->L2, c->H2p, c->M1, c->spec);
  }
  const int64_t H2p = c->H2p;
  const double scale_coeff = log(c->sdx / g.dx);  // PoissonSolverFFT2D.cc:216
  slice_strided<FWD>(c, c->M1, c->ny, c->M1, H2p, (int64_t)c->M1 * H2p, c->nx, 0.0,
                     "slice_fft_fwd");
  slice_strided<CONV2>(c, c->M0, c->nx, c->nx, (int64_t)c->M1 * H2p, H2p, c->M1, scale_coeff,
                       "slice_fft_conv");
  slice_strided<INV>(c, c->M1, c->M1, c->ny, H2p, (int64_t)c->M1 * H2p, c->nx, 0.0,
                     "slice_fft_inv");
  {
    ProfScope ps_("slice_unpack");
    hipLaunchKernelGGL(slice_unpack, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s,
                       c->spec, c->nx, c->ny, c->nz, c->L2, c->H2p, c->M1,
                       1.0 / ((double)c->M0 * c->M1), c->phi);
  }
  OB_HIP(hipGetLastError());
  if (bnd) {  // :86-88
    SC2DState& h = *c->geo_host;
    h.xMin = g.xMin; h.dx = g.dx; h.yMin = g.yMin; h.dy = g.dy; h.skip = 0;
    OB_HIP(hipMemcpyAsync(c->geo, c->geo_host, sizeof(SC2DState), hipMemcpyHostToDevice, s));
    boundary_add_potential_slices(bnd, c->phi, c->nx, c->ny, c->nz, c->geo, s);
  }
  // kick (:96-112): xp -= ex factor, yp -= ey factor, factor = 2 L / dz r0 / (beta^2 gamma^3)
  const double factor = 2 * length / g.dz * sy.r0 / (pow(sy.beta, 2) * pow(sy.gamma, 3));
  if (tiled) {
    Kick3D k{c->phi, g.zMin, g.dz, 0.0, 1.0, factor, 0.0};
    k.bnd_type = bnd_type;
    for (int i = 0; i < 6; i++) k.bnd[i] = bnd6[i];
    tile_kick_3d(b, tg, k, c->work);
    return;
  }
  Op op;  // the 3D kick op with gamma 1 and z centre 0 (z unchanged), no dE kick
  std::memset(&op, 0, sizeof(op));
  op.kind = OP_SC3D_KICK;
  op.i0 = c->nx; op.i1 = c->ny; op.i2 = c->nz;
  op.p[0] = ptr_bits(c->phi);
  op.p[1] = g.xMin; op.p[2] = g.dx; op.p[3] = g.yMin; op.p[4] = g.dy;
  op.p[5] = g.zMin; op.p[6] = g.dz;
  op.p[7] = 0.0;
  op.p[8] = 1.0;
  op.p[9] = factor;
  op.p[10] = 0.0;
  op.p[11] = 0.0;
  bunch_queue(b, op);
}
}  // namespace
}  // namespace ob

// SpaceChargeCalcSliceBySlice2D(nx, ny, nz[, xy_ratio]) (:15-32).
OB_API int ob_slice2d_create(int nx, int ny, int nz, double xy_ratio, ob_slice2d** out) {
  OB_TRY OB_REQUIRE(out != nullptr, "null output pointer");
  OB_REQUIRE(nx >= 3 && ny >= 3,
             "SpaceChargeCalcSliceBySlice2D: grid size too small (should be >= 3)");
  OB_REQUIRE(nz >= 1, "SpaceChargeCalcSliceBySlice2D: number of z slices should be >= 1");
  OB_REQUIRE(xy_ratio > 0.0, "SpaceChargeCalcSliceBySlice2D: xy_ratio must be positive");
  OB_REQUIRE(ob::pad_len(nx) <= 2048 && ob::pad_len(ny) <= 2048,
             "SpaceChargeCalcSliceBySlice2D: grid larger than 1024 points per axis");
  ob::ensure_init();
  ob_slice2d* c = new ob_slice2d();
  c->nx = nx; c->ny = ny; c->nz = nz;
  c->M0 = ob::pad_len(nx);
  c->M1 = ob::pad_len(ny);
  c->L2 = (nz + 1) / 2;
  c->H2p = (c->L2 + 7) & ~7;
  c->qx = c->M0 / 2 + 1;
  c->qy = c->M1 / 2 + 1;
  c->nmax = std::max(c->M0, c->M1);
  c->xy_ratio = xy_ratio;
  for (int k = 0; k < 9; k++) c->g[k] = 0.0;  // Grid3D(nx, ny, nz) limits (Grid3D.cc:37-39)
  const size_t ng = (size_t)nx * ny * nz;
  OB_HIP(hipMalloc((void**)&c->rho, sizeof(double) * ng));
  OB_HIP(hipMalloc((void**)&c->phi, sizeof(double) * ng));
  OB_HIP(hipMalloc((void**)&c->spec, sizeof(double2) * (size_t)nx * c->M1 * c->H2p));
  OB_HIP(hipMalloc((void**)&c->ghat, sizeof(double) * (size_t)c->qx * c->qy));
  OB_HIP(hipMalloc((void**)&c->partials, sizeof(double) * 6 * 4096));
  ob::extrema_link_create(c->ext);
  OB_HIP(hipMalloc((void**)&c->geo, sizeof(ob::SC2DState)));
  OB_HIP(hipHostMalloc((void**)&c->geo_host, sizeof(ob::SC2DState), hipHostMallocDefault));
  std::memset(c->geo_host, 0, sizeof(ob::SC2DState));
  OB_HIP(hipMemset(c->rho, 0, sizeof(double) * ng));
  OB_HIP(hipMemset(c->phi, 0, sizeof(double) * ng));
  OB_HIP(hipMemset(c->spec, 0, sizeof(double2) * (size_t)nx * c->M1 * c->H2p));
  c->tw = ob::make_twiddles(c->nmax);
  ob::slice_set_solver(c, -xy_ratio, xy_ratio, -1.0, 1.0);  // PoissonSolverFFT2D(:19)
  OB_HIP(hipStreamSynchronize(ob::stream()));
  *out = c;
  OB_CATCH
}

OB_API int ob_slice2d_destroy(ob_slice2d* c) {
  OB_TRY if (!c) return 0;
  std::string err;
  try {
    ob::flush_kicks_reading(c->phi);
  } catch (const ob::Error& e) {
    err = e.msg;
  }
  (void)hipStreamSynchronize(ob::stream());
  for (void* p : {(void*)c->rho, (void*)c->phi, (void*)c->spec, (void*)c->ghat,
                  (void*)c->partials, (void*)c->tw, (void*)c->geo})
    if (p) (void)hipFree(p);
  ob::extrema_link_destroy(c->ext);
  (void)hipHostFree(c->geo_host);
  c->work.release();
  delete c;
  OB_REQUIRE(err.empty(), err);
  OB_CATCH
}

// trackBunch(bunch, length[, boundary]) (:66-113); bnd NULL = no boundary.
OB_API int ob_slice2d_track(ob_slice2d* c, ob_bunch* b, double length, ob_boundary2d* bnd) {
  OB_TRY OB_REQUIRE(c != nullptr && b != nullptr, "null handle");
  ob::slice_track(c, b, length, bnd);
  OB_CATCH
}

// getRhoGrid / getPhiGrid: which 0 rho, 1 phi as the reference's Arr3D[iz][ix][iy];
// lims[6] = x, y, z limits of the last step.
OB_API int ob_slice2d_get_grid(ob_slice2d* c, int which, double* out, double* lims) {
  OB_TRY OB_REQUIRE(c != nullptr, "null handle");
  OB_REQUIRE(which == 0 || which == 1, "grid selector must be 0 (rho) or 1 (phi)");
  const size_t ng = (size_t)c->nx * c->ny * c->nz;
  if (out) {
    std::vector<double> tmp(ng);
    OB_HIP(hipMemcpyAsync(tmp.data(), which == 0 ? c->rho : c->phi, sizeof(double) * ng,
                          hipMemcpyDeviceToHost, ob::stream()));
    OB_HIP(hipStreamSynchronize(ob::stream()));
    for (int ix = 0; ix < c->nx; ix++)
      for (int iy = 0; iy < c->ny; iy++)
        for (int iz = 0; iz < c->nz; iz++)
          out[((size_t)iz * c->nx + ix) * c->ny + iy] = tmp[((size_t)ix * c->ny + iy) * c->nz + iz];
  }
  if (lims)
    for (int k = 0; k < 6; k++) lims[k] = c->g[k];
  OB_CATCH
}

// out[6] = xy_ratio, solver box (xMin, xMax, yMin, yMax), Green rebuilds.
OB_API int ob_slice2d_get_state(ob_slice2d* c, double* out) {
  OB_TRY OB_REQUIRE(c != nullptr && out != nullptr, "null handle");
  out[0] = c->xy_ratio;
  for (int k = 0; k < 4; k++) out[1 + k] = c->sol[k];
  out[5] = c->rebuilds;
  OB_CATCH
}
