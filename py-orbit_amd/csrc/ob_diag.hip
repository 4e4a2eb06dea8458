// ob_diag.hip -- bunch diagnostics on the device: the moments of BunchTwissAnalysis
// (src/orbit/BunchDiagnostics/BunchTwissAnalysis.cc:49-142), its x-y power moments
// (computeBunchMoments, :145-337), the six coordinate extrema of BunchExtremaCalculator
// (src/utils/bunch/BunchExtremaCalculator.cc:37-120) and Grid1D::binBunch along any axis
// (src/spacecharge/Grid1D.cc:241-275, 328-352), which the profile diagnostics use.
//
// analyzeBunch sums, over the particles, the weight (macrosize attribute or 1), the
// weighted coordinates and the weighted products of coordinate pairs.  Here each
// workgroup sums its particles in registers and writes its 28 partial sums; one workgroup
// adds the partials in a fixed order.  The host derives averages, correlations and
// emittances exactly as the reference does, without downloading the 6 x N coordinates.
#include <algorithm>
#include <cmath>

#include "ob_internal.hpp"
#include "ob_sc_device.hpp"

namespace ob {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxBlocks = 1024;
constexpr int kSums = 28;  // weight, 6 first moments, 21 products (i <= j)

__device__ __forceinline__ double block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += sh[w];
  return s;
}

__global__ __launch_bounds__(kThreads) void moment_partials(DeviceBunchView v,
                                                            double* __restrict__ part) {
  __shared__ double sh[kThreads / 64];
  double s[kSums];
#pragma unroll
  for (int k = 0; k < kSums; k++) s[k] = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < v.n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double w = v.msize ? v.msize[i] : 1.0;
    double c[6];
#pragma unroll
    for (int k = 0; k < 6; k++) c[k] = v.c[k][i];
    s[0] += w;
#pragma unroll
    for (int k = 0; k < 6; k++) s[1 + k] += w * c[k];
    int p = 7;
#pragma unroll
    for (int a = 0; a < 6; a++)
#pragma unroll
      for (int b = a; b < 6; b++) s[p++] += w * c[a] * c[b];
  }
#pragma unroll
  for (int k = 0; k < kSums; k++) {
    const double t = block_sum(s[k], sh);
    if (threadIdx.x == 0) part[(size_t)blockIdx.x * kSums + k] = t;
  }
}

__global__ __launch_bounds__(64) void moment_reduce(const double* __restrict__ part, int nblk,
                                                    double* __restrict__ out) {
  const int k = threadIdx.x;
  if (k >= kSums) return;
  double s = 0.0;
  for (int b = 0; b < nblk; b++) s += part[(size_t)b * kSums + k];
  out[k] = s;
}

// ---- extrema of the six coordinates -------------------------------------------------
// Per workgroup: min and max of x, xp, y, yp, z, dE (12 values, min/max interleaved);
// comparisons as in the reference (a value replaces the extremum only if strictly beyond).
__device__ __forceinline__ double lo(double a, double b) { return b < a ? b : a; }
__device__ __forceinline__ double hi(double a, double b) { return b > a ? b : a; }

__global__ __launch_bounds__(kThreads) void extrema6_partials(DeviceBunchView v,
                                                              double* __restrict__ part) {
  __shared__ double sh[kThreads / 64][12];
  double e[12];
#pragma unroll
  for (int k = 0; k < 6; k++) {
    e[2 * k] = DBL_MAX;
    e[2 * k + 1] = -DBL_MAX;
  }
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < v.n;
       i += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
    for (int k = 0; k < 6; k++) {
      const double c = v.c[k][i];
      e[2 * k] = lo(e[2 * k], c);
      e[2 * k + 1] = hi(e[2 * k + 1], c);
    }
  }
#pragma unroll
  for (int k = 0; k < 6; k++)
    for (int o = 32; o > 0; o >>= 1) {
      e[2 * k] = lo(e[2 * k], __shfl_xor(e[2 * k], o, 64));
      e[2 * k + 1] = hi(e[2 * k + 1], __shfl_xor(e[2 * k + 1], o, 64));
    }
  if ((threadIdx.x & 63) == 0)
    for (int k = 0; k < 12; k++) sh[threadIdx.x >> 6][k] = e[k];
  __syncthreads();
  if (threadIdx.x < 12) {
    const int k = threadIdx.x;
    double r = sh[0][k];
    for (int w = 1; w < kThreads / 64; w++) r = (k & 1) ? hi(r, sh[w][k]) : lo(r, sh[w][k]);
    part[(size_t)blockIdx.x * 12 + k] = r;
  }
}

__global__ __launch_bounds__(64) void extrema6_reduce(const double* __restrict__ part, int nblk,
                                                      double* __restrict__ out) {
  const int k = threadIdx.x;
  if (k >= 12) return;
  double r = (k & 1) ? -DBL_MAX : DBL_MAX;
  for (int b = 0; b < nblk; b++) r = (k & 1) ? hi(r, part[(size_t)b * 12 + k]) : lo(r, part[(size_t)b * 12 + k]);
  out[k] = r;
}

// ---- extremaR: largest distance from the mean transverse position -----------------------
// Pass 1: per-workgroup sums of x and y (unweighted, as the reference's mean); pass 2: each
// workgroup adds the sums in a fixed order, forms the mean and its particles' largest
// distance; pass 3: the largest of the workgroups'.
__global__ __launch_bounds__(kThreads) void xy_sum_partials(DeviceBunchView v,
                                                            double* __restrict__ part) {
  __shared__ double sh[kThreads / 64];
  double sx = 0.0, sy = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < v.n;
       i += (int64_t)gridDim.x * blockDim.x) {
    sx += v.c[0][i];
    sy += v.c[2][i];
  }
  sx = block_sum(sx, sh);
  __syncthreads();
  sy = block_sum(sy, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sx;
    part[2 * blockIdx.x + 1] = sy;
  }
}

__global__ __launch_bounds__(kThreads) void rmax_partials(DeviceBunchView v,
                                                          const double* __restrict__ sums,
                                                          int nblk, double* __restrict__ part) {
  __shared__ double sh[kThreads / 64];
  double sx = 0.0, sy = 0.0;
  for (int b = 0; b < nblk; b++) {
    sx += sums[2 * b];
    sy += sums[2 * b + 1];
  }
  const double xa = sx / (double)v.n, ya = sy / (double)v.n;
  double r = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < v.n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double dx = v.c[0][i] - xa, dy = v.c[2][i] - ya;
    const double ri = sqrt(dx * dx + dy * dy);
    r = ri > r ? ri : r;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const double t = __shfl_xor(r, o, 64);
    r = t > r ? t : r;
  }
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = r;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kThreads / 64; w++) r = sh[w] > r ? sh[w] : r;
    part[blockIdx.x] = r;
  }
}

__global__ __launch_bounds__(64) void max_reduce(const double* __restrict__ part, int nblk,
                                                 double* __restrict__ out) {
  if (threadIdx.x != 0) return;
  double r = 0.0;
  for (int b = 0; b < nblk; b++) r = part[b] > r ? part[b] : r;
  *out = r;
}

// ---- Grid1D::binBunch along one coordinate --------------------------------------------
// Linear (cloud-in-cell) weights of Grid1D::binValue (Grid1D.cc:422-440) with the grid's
// periodic wrap at both ends.  Grids up to kBinLds bins accumulate a private histogram per
// workgroup in LDS and write it out; a second pass adds the workgroups' histograms bin by
// bin in a fixed order.  Larger grids add into one global histogram with atomics.
constexpr int kBinLds = 4096;  // 32 KiB of FP64 bins per workgroup

struct Bin1D {
  int axis, by_particle, nz;
  double value;  // bunch macrosize when the bunch has no macrosize attribute
  double zMin, zMax, dz;
};

template <bool LDS>
__global__ __launch_bounds__(kThreads) void bin1d_partials(DeviceBunchView v, Bin1D g,
                                                           double* __restrict__ part) {
  extern __shared__ double hist[];
  double* h = LDS ? hist : part;
  if (LDS) {
    for (int k = threadIdx.x; k < g.nz; k += blockDim.x) hist[k] = 0.0;
    __syncthreads();
  }
  const double* __restrict__ zc = v.c[g.axis];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < v.n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double z = zc[i];
    if (z < g.zMin || z > g.zMax) continue;
    const double val = g.by_particle ? 1.0 : (v.msize ? v.msize[i] : g.value);
    int i0, ip;
    double w0, wp;
    grid1d_ind_wz(z, g.zMin, g.dz, g.nz, i0, ip, w0, wp);
    if (g.nz > 1) {
      atomicAdd(&h[i0], w0 * val);
      atomicAdd(&h[ip], wp * val);
    } else {
      atomicAdd(&h[i0], val);
    }
  }
  if (LDS) {
    __syncthreads();
    for (int k = threadIdx.x; k < g.nz; k += blockDim.x)
      part[(size_t)blockIdx.x * g.nz + k] = hist[k];
  }
}

__global__ __launch_bounds__(kThreads) void bin1d_reduce(const double* __restrict__ part,
                                                         int nblk, int nz,
                                                         double* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nz) return;
  double s = 0.0;
  for (int b = 0; b < nblk; b++) s += part[(size_t)b * nz + k];
  out[k] = s;
}

// ---- computeBunchMoments: sums of x-y power products --------------------------------
// BunchTwissAnalysis::computeBunchMoments (BunchTwissAnalysis.cc:145-337) accumulates, for
// j < order and i <= order - j, momX[i] / xb^i * momY[j] / yb^j over the particles, with
// momX[i] = prod (m (x - D dE / (E + M) / beta^2 - <x>)) and momY[j] = prod (m (y - <y>))
// (m = 1 without the macrosize attribute); with the attribute each term also adds
// momX[i] momY[j].  Without the attribute the statement after the loops adds
// momX[2] momY[order] to entry [2][order] once per particle (it sits outside the braces).
// Each lane keeps its sums in an LDS column (one wave per workgroup); the lanes' sums are
// added in lane order, then the workgroups' in workgroup order.
constexpr int kMaxMomOrder = 12;
constexpr int kMomLanes = 64;

struct XYMom {
  int order, nterms, disp, stray;  // nterms = order (order + 3) / 2
  double xAvg, yAvg, D, Em, b2;    // Em = E_kin + mass, b2 = beta^2
  double px[kMaxMomOrder + 1], py[kMaxMomOrder + 1];  // xb^i, yb^j
};

__global__ __launch_bounds__(kMomLanes) void xy_moment_partials(DeviceBunchView v, XYMom p,
                                                                double* __restrict__ part) {
  extern __shared__ double acc[];  // [nterms + 2][kMomLanes]
  __shared__ double spx[kMaxMomOrder + 1], spy[kMaxMomOrder + 1];
  const int t = threadIdx.x;
  const int nt = p.nterms + 2;
  for (int k = 0; k < nt; k++) acc[k * kMomLanes + t] = 0.0;
  if (t <= p.order) {
    spx[t] = p.px[t];
    spy[t] = p.py[t];
  }
  __syncthreads();
  const bool ms = v.msize != nullptr;
  for (int64_t i = (int64_t)blockIdx.x * kMomLanes + t; i < v.n;
       i += (int64_t)gridDim.x * kMomLanes) {
    const double m = ms ? v.msize[i] : 1.0;
    const double dispterm = p.disp ? p.D * v.c[5][i] / p.Em / p.b2 : 0.0;
    const double dx = (v.c[0][i] - dispterm) - p.xAvg;
    const double dy = v.c[2][i] - p.yAvg;
    double mY = 1.0;
    int k = 0;
    for (int j = 0; j < p.order; j++) {
      double mX = 1.0;
      for (int a = 0; a <= p.order - j; a++, k++) {
        double s = acc[k * kMomLanes + t];
        s += mX / spx[a] * mY / spy[j];
        if (ms) s += mX * mY;
        acc[k * kMomLanes + t] = s;
        mX = ms ? mX * m * dx : mX * dx;
      }
      mY = ms ? mY * m * dy : mY * dy;
    }
    if (p.stray) acc[k * kMomLanes + t] += dx * dx * mY;  // momX[2] momY[order]
    acc[(k + 1) * kMomLanes + t] += m;
  }
  __syncthreads();
  for (int k = t; k < nt; k += kMomLanes) {
    double s = 0.0;
    for (int l = 0; l < kMomLanes; l++) s += acc[k * kMomLanes + l];
    part[(size_t)blockIdx.x * nt + k] = s;
  }
}

__global__ __launch_bounds__(128) void xy_moment_reduce(const double* __restrict__ part, int nblk,
                                                        int nt, double* __restrict__ out) {
  const int k = threadIdx.x;
  if (k >= nt) return;
  double s = 0.0;
  for (int b = 0; b < nblk; b++) s += part[(size_t)b * nt + k];
  out[k] = s;
}

// Scratch of ob_bunch_moments_noalloc: the per-workgroup partial sums (at most kMaxBlocks
// rows of kSums) and their total, 229,600 bytes whatever the bunch size.
constexpr size_t kScratchDoubles = (size_t)(kMaxBlocks + 1) * kSums;

}  // namespace
}  // namespace ob

// Held by every device the library runs on until the process ends.  In host code this name
// is a host-side stand-in, not a device address: host code uses the address looked up by
// diag_scratch_setup for the current device.
static __device__ __align__(256) double scratch[ob::kScratchDoubles];

namespace ob {
namespace {
constexpr int kMaxDevices = 64;
double* g_scratch_dev[kMaxDevices] = {};

// Partial sums into buf[0 .. nblk * kSums), their total at buf + nblk * kSums, copied to
// the host array out[kSums]; waits for the copy.  Every partial row and the total are
// written before they are read, so buf needs no clearing.  timed: record the partials
// kernel in the kernel profiler (which creates HIP events when it is on).
void moments_into(const DeviceBunchView& v, int nblk, double* buf, double* out, hipStream_t s,
                  bool timed) {
  if (timed) {
    ProfScope ps("moment_partials");
    hipLaunchKernelGGL(moment_partials, dim3(nblk), dim3(kThreads), 0, s, v, buf);
  } else {
    hipLaunchKernelGGL(moment_partials, dim3(nblk), dim3(kThreads), 0, s, v, buf);
  }
  hipLaunchKernelGGL(moment_reduce, dim3(1), dim3(64), 0, s, buf, nblk,
                     buf + (size_t)nblk * kSums);
  OB_HIP(hipGetLastError());
  OB_HIP(hipMemcpyAsync(out, buf + (size_t)nblk * kSums, sizeof(double) * kSums,
                        hipMemcpyDeviceToHost, s));
  OB_HIP(hipStreamSynchronize(s));
}
}  // namespace

void diag_scratch_setup(int device) {
  OB_REQUIRE(device >= 0 && device < kMaxDevices, "device index out of range");
  if (g_scratch_dev[device] != nullptr) return;
  double* p = nullptr;
  OB_HIP(hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch)));
  OB_REQUIRE(p != nullptr, "no device address for the moments scratch");
  g_scratch_dev[device] = p;
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

// out[0] = sum of weights, out[1..6] = sum w c_k, out[7..27] = sum w c_a c_b (a <= b,
// row by row), weight = macrosize attribute or 1 (BunchTwissAnalysis.cc:49-142).
OB_API int ob_bunch_moments(ob_bunch* b, double* out) {
  OB_TRY OB_REQUIRE(b != nullptr && out != nullptr, "null handle");
  ob::bunch_flush(b, nullptr);
  const ob::DeviceBunchView v = ob::bunch_device_view(b);
  hipStream_t s = ob::stream();
  if (v.n == 0) {
    for (int k = 0; k < ob::kSums; k++) out[k] = 0.0;
    return 0;
  }
  const int nblk =
      (int)std::min<int64_t>(ob::kMaxBlocks, (v.n + ob::kThreads - 1) / ob::kThreads);
  double* buf = nullptr;
  OB_HIP(hipMalloc((void**)&buf, sizeof(double) * (nblk + 1) * ob::kSums));
  try {
    ob::moments_into(v, nblk, buf, out, s, true);
  } catch (...) {
    (void)hipFree(buf);
    throw;
  }
  OB_HIP(hipFree(buf));
  OB_CATCH
}

// ob_bunch_moments without any allocation: the partial sums go to the library's fixed
// device scratch (229,600 bytes per device).  Calls that overlap on one device race on
// that scratch.  A batched bunch is refused (its flush may rebuild the batch's records).
OB_API int ob_bunch_moments_noalloc(ob_bunch* b, double* out) {
  OB_TRY OB_REQUIRE(b != nullptr && out != nullptr, "null handle");
  OB_REQUIRE(!ob::bunch_in_batch(b),
             "ob_bunch_moments_noalloc: the bunch belongs to a batch; use ob_bunch_moments");
  int dev = -1;
  OB_HIP(hipGetDevice(&dev));
  OB_REQUIRE(dev >= 0 && dev < ob::kMaxDevices && ob::g_scratch_dev[dev] != nullptr,
             "ob_bunch_moments_noalloc: the library is not initialised on this device");
  ob::bunch_flush(b, nullptr);
  const ob::DeviceBunchView v = ob::bunch_device_view(b);
  if (v.n == 0) {
    for (int k = 0; k < ob::kSums; k++) out[k] = 0.0;
    return 0;
  }
  const int nblk =
      (int)std::min<int64_t>(ob::kMaxBlocks, (v.n + ob::kThreads - 1) / ob::kThreads);
  const size_t need = (size_t)(nblk + 1) * ob::kSums;
  OB_REQUIRE(need <= ob::kScratchDoubles,
             "ob_bunch_moments_noalloc: " + std::to_string(need) +
                 " doubles of scratch needed, the maximum is " +
                 std::to_string(ob::kScratchDoubles));
  ob::moments_into(v, nblk, ob::g_scratch_dev[dev], out, ob::stream(), false);
  OB_CATCH
}

// BunchExtremaCalculator::getExtremaXYZ and getExtremaXpYpdE (BunchExtremaCalculator.cc:37-120)
// in one pass: out[12] = min, max of x, xp, y, yp, z, dE; DBL_MAX, -DBL_MAX for an empty
// bunch, as in the reference.
OB_API int ob_bunch_extrema(ob_bunch* b, double* out) {
  OB_TRY OB_REQUIRE(b != nullptr && out != nullptr, "null handle");
  ob::bunch_flush(b, nullptr);
  const ob::DeviceBunchView v = ob::bunch_device_view(b);
  hipStream_t s = ob::stream();
  if (v.n == 0) {
    for (int k = 0; k < 6; k++) {
      out[2 * k] = DBL_MAX;
      out[2 * k + 1] = -DBL_MAX;
    }
    return 0;
  }
  const int nblk =
      (int)std::min<int64_t>(ob::kMaxBlocks, (v.n + ob::kThreads - 1) / ob::kThreads);
  double* buf = nullptr;
  OB_HIP(hipMalloc((void**)&buf, sizeof(double) * (nblk + 1) * 12));
  try {
    {
      ob::ProfScope ps("extrema6_partials");
      hipLaunchKernelGGL(ob::extrema6_partials, dim3(nblk), dim3(ob::kThreads), 0, s, v, buf);
    }
    hipLaunchKernelGGL(ob::extrema6_reduce, dim3(1), dim3(64), 0, s, buf, nblk,
                       buf + (size_t)nblk * 12);
    OB_HIP(hipGetLastError());
    OB_HIP(hipMemcpyAsync(out, buf + (size_t)nblk * 12, sizeof(double) * 12,
                          hipMemcpyDeviceToHost, s));
    OB_HIP(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipFree(buf);
    throw;
  }
  OB_HIP(hipFree(buf));
  OB_CATCH
}

// Grid1D::binBunch(bunch, axis) / binBunchByParticle(bunch, axis) (Grid1D.cc:241-275,
// 335-352) on a grid of nz cells over [zMin, zMax]: the binned values are ADDED to the
// host array arr[nz], as the reference adds into its grid.  by_particle = 1 gives each
// particle weight 1; otherwise its macrosize attribute or the bunch macrosize.
OB_API int ob_grid1d_bin_bunch(ob_bunch* b, int axis, int by_particle, int nz, double zMin,
                               double zMax, double* arr) {
  OB_TRY OB_REQUIRE(b != nullptr && arr != nullptr, "null handle");
  OB_REQUIRE(axis >= 0 && axis <= 5, "Grid1D::binBunch: the axis index should be between 0 and 5");
  OB_REQUIRE(nz >= 1 && nz <= (1 << 26),
             "Grid1D.binBunch: the number of bins should be between 1 and 2^26");
  ob::bunch_flush(b, nullptr);
  const ob::DeviceBunchView v = ob::bunch_device_view(b);
  if (v.n == 0) return 0;
  const ob::SyncInfo sy = ob::bunch_sync(b);
  hipStream_t s = ob::stream();
  ob::Bin1D g;
  g.axis = axis;
  g.by_particle = by_particle ? 1 : 0;
  g.nz = nz;
  g.value = sy.macrosize;
  g.zMin = zMin;
  g.zMax = zMax;
  g.dz = (zMax - zMin) / nz;  // Grid1D::setGridZ (Grid1D.cc:89-94)
  const bool lds = nz <= ob::kBinLds;
  const int64_t want = (v.n + ob::kThreads - 1) / ob::kThreads;
  const int nblk = (int)std::min<int64_t>(lds ? 512 : ob::kMaxBlocks, want);
  const size_t part_n = lds ? (size_t)nblk * nz : (size_t)nz;
  double* buf = nullptr;
  OB_HIP(hipMalloc((void**)&buf, sizeof(double) * (part_n + nz)));
  std::vector<double> h((size_t)nz);
  try {
    double* res = buf + part_n;
    if (lds) {
      ob::ProfScope ps("bin1d_partials");
      hipLaunchKernelGGL(ob::bin1d_partials<true>, dim3(nblk), dim3(ob::kThreads),
                         sizeof(double) * nz, s, v, g, buf);
      OB_HIP(hipGetLastError());
      hipLaunchKernelGGL(ob::bin1d_reduce, dim3((nz + ob::kThreads - 1) / ob::kThreads),
                         dim3(ob::kThreads), 0, s, buf, nblk, nz, res);
    } else {
      OB_HIP(hipMemsetAsync(buf, 0, sizeof(double) * nz, s));
      ob::ProfScope ps("bin1d_partials");
      hipLaunchKernelGGL(ob::bin1d_partials<false>, dim3(nblk), dim3(ob::kThreads), 0, s, v, g,
                         buf);
      res = buf;
    }
    OB_HIP(hipGetLastError());
    OB_HIP(hipMemcpyAsync(h.data(), res, sizeof(double) * nz, hipMemcpyDeviceToHost, s));
    OB_HIP(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipFree(buf);
    throw;
  }
  OB_HIP(hipFree(buf));
  for (int k = 0; k < nz; k++) arr[k] += h[(size_t)k];
  OB_CATCH
}

// BunchTwissAnalysis::computeBunchMoments (BunchTwissAnalysis.cc:145-337), the particle
// sums.  prm[8] = xAvg, yAvg, D_x, E_kin + mass, beta^2, xb, yb, dispersion flag (> 0:
// subtract D_x dE / (E_kin + mass) / beta^2 from x); xb, yb are sqrt(beta) or
// sqrt(beta emittance) as the caller chose.  out[order (order + 3) / 2 + 2] = the sums
// for j = 0 .. order-1, i = 0 .. order-j (j outer), then the [2][order] entry of a bunch
// without the macrosize attribute (order >= 2; else 0), then the total weight.
OB_API int ob_bunch_xy_moments(ob_bunch* b, int order, const double* prm, double* out) {
  OB_TRY OB_REQUIRE(b != nullptr && prm != nullptr && out != nullptr, "null handle");
  OB_REQUIRE(order >= 1 && order <= ob::kMaxMomOrder,
             "computeBunchMoments: order must be between 1 and 12");
  ob::bunch_flush(b, nullptr);
  const ob::DeviceBunchView v = ob::bunch_device_view(b);
  hipStream_t s = ob::stream();
  ob::XYMom p;
  p.order = order;
  p.nterms = order * (order + 3) / 2;
  p.disp = prm[7] > 0 ? 1 : 0;
  p.stray = (v.msize == nullptr && order >= 2) ? 1 : 0;
  p.xAvg = prm[0]; p.yAvg = prm[1]; p.D = prm[2]; p.Em = prm[3]; p.b2 = prm[4];
  for (int k = 0; k <= ob::kMaxMomOrder; k++) {
    p.px[k] = pow(prm[5], (double)k);
    p.py[k] = pow(prm[6], (double)k);
  }
  const int nt = p.nterms + 2;
  if (v.n == 0) {
    for (int k = 0; k < nt; k++) out[k] = 0.0;
    return 0;
  }
  const int nblk = (int)std::min<int64_t>(1024, (v.n + ob::kMomLanes - 1) / ob::kMomLanes);
  double* buf = nullptr;
  OB_HIP(hipMalloc((void**)&buf, sizeof(double) * (size_t)(nblk + 1) * nt));
  try {
    {
      ob::ProfScope ps("xy_moment_partials");
      hipLaunchKernelGGL(ob::xy_moment_partials, dim3(nblk), dim3(ob::kMomLanes),
                         sizeof(double) * nt * ob::kMomLanes, s, v, p, buf);
    }
    OB_HIP(hipGetLastError());
    hipLaunchKernelGGL(ob::xy_moment_reduce, dim3(1), dim3(128), 0, s, buf, nblk, nt,
                       buf + (size_t)nblk * nt);
    OB_HIP(hipGetLastError());
    OB_HIP(hipMemcpyAsync(out, buf + (size_t)nblk * nt, sizeof(double) * nt,
                          hipMemcpyDeviceToHost, s));
    OB_HIP(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipFree(buf);
    throw;
  }
  OB_HIP(hipFree(buf));
  OB_CATCH
}

// BunchExtremaCalculator::getExtremaR (BunchExtremaCalculator.cc:167-208): the largest
// distance sqrt((x - <x>)^2 + (y - <y>)^2) from the unweighted mean position; 0 for an
// empty bunch.
OB_API int ob_bunch_extrema_r(ob_bunch* b, double* rmax) {
  OB_TRY OB_REQUIRE(b != nullptr && rmax != nullptr, "null handle");
  ob::bunch_flush(b, nullptr);
  const ob::DeviceBunchView v = ob::bunch_device_view(b);
  hipStream_t s = ob::stream();
  if (v.n == 0) {
    *rmax = 0.0;
    return 0;
  }
  const int nblk =
      (int)std::min<int64_t>(ob::kMaxBlocks, (v.n + ob::kThreads - 1) / ob::kThreads);
  double* buf = nullptr;
  OB_HIP(hipMalloc((void**)&buf, sizeof(double) * (3 * (size_t)nblk + 1)));
  try {
    double* sums = buf;
    double* rpart = buf + 2 * (size_t)nblk;
    {
      ob::ProfScope ps("rmax_partials");
      hipLaunchKernelGGL(ob::xy_sum_partials, dim3(nblk), dim3(ob::kThreads), 0, s, v, sums);
      hipLaunchKernelGGL(ob::rmax_partials, dim3(nblk), dim3(ob::kThreads), 0, s, v, sums,
                         nblk, rpart);
    }
    hipLaunchKernelGGL(ob::max_reduce, dim3(1), dim3(64), 0, s, rpart, nblk,
                       buf + 3 * (size_t)nblk);
    OB_HIP(hipGetLastError());
    OB_HIP(hipMemcpyAsync(rmax, buf + 3 * (size_t)nblk, sizeof(double), hipMemcpyDeviceToHost,
                          s));
    OB_HIP(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipFree(buf);
    throw;
  }
  OB_HIP(hipFree(buf));
  OB_CATCH
}
