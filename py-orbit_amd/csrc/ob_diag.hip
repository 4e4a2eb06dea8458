// ob_diag.hip -- bunch moments on the device for BunchTwissAnalysis
// (src/orbit/BunchDiagnostics/BunchTwissAnalysis.cc:49-142).
//
// analyzeBunch sums, over the particles, the weight (macrosize attribute or 1), the
// weighted coordinates and the weighted products of coordinate pairs.  Here each
// workgroup sums its particles in registers and writes its 28 partial sums; one workgroup
// adds the partials in a fixed order.  The host derives averages, correlations and
// emittances exactly as the reference does, without downloading the 6 x N coordinates.
#include "ob_internal.hpp"

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
    {
      ob::ProfScope ps("moment_partials");
      hipLaunchKernelGGL(ob::moment_partials, dim3(nblk), dim3(ob::kThreads), 0, s, v, buf);
    }
    hipLaunchKernelGGL(ob::moment_reduce, dim3(1), dim3(64), 0, s, buf, nblk,
                       buf + (size_t)nblk * ob::kSums);
    OB_HIP(hipGetLastError());
    OB_HIP(hipMemcpyAsync(out, buf + (size_t)nblk * ob::kSums, sizeof(double) * ob::kSums,
                          hipMemcpyDeviceToHost, s));
    OB_HIP(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipFree(buf);
    throw;
  }
  OB_HIP(hipFree(buf));
  OB_CATCH
}
