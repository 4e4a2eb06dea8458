// ob_tiled3d.hip -- tile-sorted 3D charge deposition (Grid3D::binValue, Grid3D.cc:307-432).
//
// Scattering 27 FP64 atomics per particle in random particle order over a 64 MiB grid
// misses L2 on almost every update.  Instead, each SC step:
//   1. tile_hist    key = xy-tile of the particle's TSC centre node; per-block LDS
//                   histogram, merged into global tile counts;
//   2. tile_scan    exclusive scan of the counts; work items of at most kChunk particles
//                   per tile;
//   3. tile_scatter counting-sort scatter of the bunch (6 coordinates, macrosize and the
//                   user index) into the spare buffers, ranges reserved per block;
//   4. tile_deposit one workgroup per work item accumulates its particles into an LDS
//                   block of (T+2) x (T+2) x nz cells, then adds the block to rho with
//                   coalesced atomics (z is the fastest axis).
// The bunch stays sorted for the kick, whose 27-point potential gathers then hit cache.
// Weights and indices are the reference's (same device helpers as the unsorted path);
// only the FP64 summation order differs.
#include "ob_internal.hpp"
#include "ob_sc_device.hpp"
#include "ob_tiled3d.hpp"

namespace ob {
namespace {

constexpr int kThreads = 256;
constexpr int kChunk = 4096;

__device__ __forceinline__ bool inside(const TileGeom& g, double x, double y, double z) {
  return !(x < g.xMin || x > g.xMax || y < g.yMin || y > g.yMax || z < g.zMin || z > g.zMax);
}

__device__ __forceinline__ uint32_t tile_key(const TileGeom& g, double x, double y, double z) {
  if (!inside(g, x, y, z)) return (uint32_t)g.ntiles;
  int iX, iY;
  double xF, yF;
  node_ind_frac(x, g.xMin, g.dx, g.nx, iX, xF);
  node_ind_frac(y, g.yMin, g.dy, g.ny, iY, yF);
  return (uint32_t)((iX / g.T) * g.ty + (iY / g.T));
}

__global__ __launch_bounds__(kThreads) void tile_hist(const double* __restrict__ cur, int64_t cap,
                                                      int64_t n, TileGeom g,
                                                      uint32_t* __restrict__ key,
                                                      uint32_t* __restrict__ counts) {
  extern __shared__ uint32_t lh[];
  const int nb = g.ntiles + 1;
  for (int t = threadIdx.x; t < nb; t += blockDim.x) lh[t] = 0;
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint32_t k = tile_key(g, cur[i], cur[2 * cap + i], cur[4 * cap + i]);
    key[i] = k;
    atomicAdd(&lh[k], 1u);
  }
  __syncthreads();
  for (int t = threadIdx.x; t < nb; t += blockDim.x)
    if (lh[t]) atomicAdd(&counts[t], lh[t]);
}

// One block.  counts[0..ntiles] -> offsets (exclusive), cursor = offsets, work items;
// counts are cleared for the next step.
__global__ __launch_bounds__(1024) void tile_scan(uint32_t* __restrict__ counts, int ntiles,
                                                  uint32_t* __restrict__ offsets,
                                                  uint32_t* __restrict__ cursor,
                                                  int4* __restrict__ items,
                                                  uint32_t* __restrict__ nitems) {
  __shared__ uint32_t part[1024], cpart[1024];
  const int nb = ntiles + 1;
  const int per = (nb + blockDim.x - 1) / blockDim.x;
  const int lo = threadIdx.x * per, hi = min(nb, lo + per);
  uint32_t s = 0, cs = 0;
  for (int t = lo; t < hi; t++) {
    s += counts[t];
    if (t < ntiles) cs += (counts[t] + kChunk - 1) / kChunk;
  }
  part[threadIdx.x] = s;
  cpart[threadIdx.x] = cs;
  __syncthreads();
  if (threadIdx.x == 0) {  // serial scan of 1024 partial sums
    uint32_t a = 0, c = 0;
    for (int k = 0; k < (int)blockDim.x; k++) {
      const uint32_t va = part[k], vc = cpart[k];
      part[k] = a;
      cpart[k] = c;
      a += va;
      c += vc;
    }
    *nitems = c;
  }
  __syncthreads();
  uint32_t off = part[threadIdx.x], item = cpart[threadIdx.x];
  for (int t = lo; t < hi; t++) {
    const uint32_t cnt = counts[t];
    offsets[t] = off;
    cursor[t] = off;
    if (t < ntiles) {
      for (uint32_t c0 = 0; c0 < cnt; c0 += kChunk) {
        items[item++] = make_int4(t, (int)(off + c0), (int)(off + min(cnt, c0 + kChunk)), 0);
      }
    }
    off += cnt;
    counts[t] = 0;
  }
}

__global__ __launch_bounds__(kThreads) void tile_scatter(
    const double* __restrict__ cur, double* __restrict__ alt, int64_t cap, int64_t n,
    const uint32_t* __restrict__ key, int ntiles, uint32_t* __restrict__ cursor,
    const int32_t* __restrict__ perm, int32_t* __restrict__ perm_alt,
    const double* __restrict__ msize, double* __restrict__ msize_alt, int64_t per_block) {
  extern __shared__ uint32_t sh[];
  uint32_t* lc = sh;               // local counts, then running destinations
  const int nb = ntiles + 1;
  for (int t = threadIdx.x; t < nb; t += blockDim.x) lc[t] = 0;
  __syncthreads();
  const int64_t lo = (int64_t)blockIdx.x * per_block;
  const int64_t hi = min(n, lo + per_block);
  for (int64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) atomicAdd(&lc[key[i]], 1u);
  __syncthreads();
  for (int t = threadIdx.x; t < nb; t += blockDim.x) {
    const uint32_t c = lc[t];
    lc[t] = c ? atomicAdd(&cursor[t], c) : 0u;
  }
  __syncthreads();
  for (int64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
    const int64_t d = atomicAdd(&lc[key[i]], 1u);
#pragma unroll
    for (int k = 0; k < 6; k++) alt[k * cap + d] = cur[k * cap + i];
    perm_alt[d] = perm[i];
    if (msize) msize_alt[d] = msize[i];
  }
}

__global__ __launch_bounds__(kThreads) void tile_deposit(const double* __restrict__ cur,
                                                         int64_t cap,
                                                         const double* __restrict__ msize,
                                                         double macrosize, TileGeom g,
                                                         const int4* __restrict__ items,
                                                         const uint32_t* __restrict__ nitems,
                                                         double* __restrict__ rho) {
  if (blockIdx.x >= *nitems) return;
  extern __shared__ double acc[];
  const int4 it = items[blockIdx.x];
  const int t = it.x;
  const int W = g.T + 2, nz = g.nz;
  const int cells = W * W * nz;
  const int x0 = (t / g.ty) * g.T - 1, y0 = (t % g.ty) * g.T - 1;
  for (int c = threadIdx.x; c < cells; c += blockDim.x) acc[c] = 0.0;
  __syncthreads();
  for (int i = it.y + threadIdx.x; i < it.z; i += blockDim.x) {
    const double x = cur[i], y = cur[2 * cap + i], z = cur[4 * cap + i];
    const double m = msize ? msize[i] : macrosize;
    int iX, iY, iZ;
    double xF, yF, zF;
    node_ind_frac(x, g.xMin, g.dx, g.nx, iX, xF);
    node_ind_frac(y, g.yMin, g.dy, g.ny, iY, yF);
    grid3d_z_ind_frac(z, g.zMin, g.dz, nz, iZ, zF);
    const double wx[3] = {0.5 * (0.5 - xF) * (0.5 - xF), 0.75 - xF * xF, 0.5 * (0.5 + xF) * (0.5 + xF)};
    const double wy[3] = {0.5 * (0.5 - yF) * (0.5 - yF), 0.75 - yF * yF, 0.5 * (0.5 + yF) * (0.5 + yF)};
    // local origin of the 3x3 stencil: node (iX-1, iY-1) in tile coordinates
    double* base = acc + ((iX - 1 - x0) * W + (iY - 1 - y0)) * nz;
    if (nz >= 3) {
      int iZm = iZ - 1, iZp = iZ + 1;
      if (iZm < 0) iZm = nz - 1;
      if (iZp >= nz) iZp = 0;
      const double wz[3] = {0.5 * (0.5 - zF) * (0.5 - zF), 0.75 - zF * zF,
                            0.5 * (0.5 + zF) * (0.5 + zF)};
      const int zz[3] = {iZm, iZ, iZp};
#pragma unroll
      for (int a = 0; a < 3; a++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
          const double tmp = wy[c] * wz[a] * m;
#pragma unroll
          for (int d = 0; d < 3; d++) atomicAdd(base + (d * W + c) * nz + zz[a], wx[d] * tmp);
        }
    } else if (nz == 2) {
      const double wz[2] = {1.0 - zF, zF};
      for (int a = 0; a < 2; a++)
        for (int c = 0; c < 3; c++) {
          const double tmp = wy[c] * wz[a] * m;
          for (int d = 0; d < 3; d++) atomicAdd(base + (d * W + c) * nz + a, wx[d] * tmp);
        }
    } else {
      for (int c = 0; c < 3; c++) {
        const double tmp = wy[c] * m;
        for (int d = 0; d < 3; d++) atomicAdd(base + (d * W + c) * nz, wx[d] * tmp);
      }
    }
  }
  __syncthreads();
  const int64_t sx = (int64_t)g.ny * nz;
  for (int c = threadIdx.x; c < cells; c += blockDim.x) {
    const double v = acc[c];
    if (v == 0.0) continue;
    const int lz = c % nz, r = c / nz, ly = r % W, lx = r / W;
    const int gx = x0 + lx, gy = y0 + ly;
    atomicAdd(rho + gx * sx + (int64_t)gy * nz + lz, v);
  }
}

}  // namespace

TileGeom make_tile_geom(int nx, int ny, int nz, double xMin, double xMax, double yMin,
                        double yMax, double zMin, double zMax, double dx, double dy, double dz,
                        int lds_budget) {
  TileGeom g;
  g.nx = nx; g.ny = ny; g.nz = nz;
  g.xMin = xMin; g.xMax = xMax; g.yMin = yMin; g.yMax = yMax; g.zMin = zMin; g.zMax = zMax;
  g.dx = dx; g.dy = dy; g.dz = dz;
  g.T = 0;
  for (int T : {8, 6, 4, 3, 2, 1})
    if ((int64_t)(T + 2) * (T + 2) * nz * 8 <= lds_budget) {
      g.T = T;
      break;
    }
  if (g.T > 0) {
    g.tx = (nx + g.T - 1) / g.T;
    g.ty = (ny + g.T - 1) / g.T;
    g.ntiles = g.tx * g.ty;
  } else {
    g.tx = g.ty = g.ntiles = 0;
  }
  return g;
}

size_t tile_lds_bytes(const TileGeom& g) { return sizeof(double) * (g.T + 2) * (g.T + 2) * g.nz; }

void TileWork::reserve(int64_t n, int ntiles) {
  if (n > n_cap) {
    if (key) OB_HIP(hipFree(key));
    int64_t c = n_cap ? n_cap : 1 << 20;
    while (c < n) c *= 2;
    OB_HIP(hipMalloc((void**)&key, sizeof(uint32_t) * c));
    n_cap = c;
  }
  const int64_t want_items = n_cap / kChunk + ntiles + 2;
  if (ntiles + 1 > t_cap || want_items > i_cap) {
    for (void* p : {(void*)counts, (void*)offsets, (void*)cursor, (void*)items, (void*)nitems})
      if (p) OB_HIP(hipFree(p));
    t_cap = std::max<int64_t>(ntiles + 1, t_cap);
    i_cap = std::max<int64_t>(want_items, i_cap);
    OB_HIP(hipMalloc((void**)&counts, sizeof(uint32_t) * t_cap));
    OB_HIP(hipMalloc((void**)&offsets, sizeof(uint32_t) * t_cap));
    OB_HIP(hipMalloc((void**)&cursor, sizeof(uint32_t) * t_cap));
    OB_HIP(hipMalloc((void**)&items, sizeof(int4) * i_cap));
    OB_HIP(hipMalloc((void**)&nitems, sizeof(uint32_t)));
    OB_HIP(hipMemset(counts, 0, sizeof(uint32_t) * t_cap));
  }
}

void TileWork::release() {
  for (void* p : {(void*)key, (void*)counts, (void*)offsets, (void*)cursor, (void*)items,
                  (void*)nitems})
    if (p) (void)hipFree(p);
  key = nullptr;
  counts = offsets = cursor = nullptr;
  items = nullptr;
  nitems = nullptr;
  n_cap = t_cap = i_cap = 0;
}

void sort_and_deposit_3d(ob_bunch* b, const TileGeom& g, double macrosize, double* rho,
                         TileWork& w) {
  static bool attr_done = false;
  if (!attr_done) {
    OB_HIP(hipFuncSetAttribute((const void*)tile_deposit,
                               hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    OB_HIP(hipFuncSetAttribute((const void*)tile_hist,
                               hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    OB_HIP(hipFuncSetAttribute((const void*)tile_scatter,
                               hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    attr_done = true;
  }
  const int64_t n = bunch_size(b);
  OB_REQUIRE((int64_t)(g.ntiles + 1) * 4 <= 150 * 1024, "tile sort: too many tiles for LDS");
  OB_REQUIRE(n < (int64_t)1 << 31, "tile sort: more than 2^31 particles");
  w.reserve(n, g.ntiles);
  hipStream_t s = stream();
  SortBuffers sb = bunch_sort_begin(b);
  const size_t hl = sizeof(uint32_t) * (g.ntiles + 1);
  int hb = (int)std::min<int64_t>((n + kThreads - 1) / kThreads, 4 * device_cu_count());
  if (hb < 1) hb = 1;
  {
    ProfScope ps("tile_hist");
    hipLaunchKernelGGL(tile_hist, dim3(hb), dim3(kThreads), hl, s, sb.cur, sb.cap, n, g, w.key,
                       w.counts);
  }
  {
    ProfScope ps("tile_scan");
    hipLaunchKernelGGL(tile_scan, dim3(1), dim3(1024), 0, s, w.counts, g.ntiles, w.offsets,
                       w.cursor, w.items, w.nitems);
  }
  const int64_t per = std::max<int64_t>(2048, (n + 1023) / 1024);
  const int sb_blocks = (int)((n + per - 1) / per);
  if (n) {
    ProfScope ps("tile_scatter");
    hipLaunchKernelGGL(tile_scatter, dim3(sb_blocks), dim3(kThreads), hl, s, sb.cur, sb.alt,
                       sb.cap, n, w.key, g.ntiles, w.cursor, sb.perm, sb.perm_alt, sb.msize,
                       sb.msize_alt, per);
  }
  OB_HIP(hipGetLastError());
  bunch_sort_commit(b);
  DeviceBunchView v = bunch_device_view(b);
  const int64_t max_items = n / kChunk + g.ntiles + 1;
  {
    ProfScope ps("tile_deposit");
    hipLaunchKernelGGL(tile_deposit, dim3((unsigned)max_items), dim3(kThreads), tile_lds_bytes(g),
                       s, v.c[0], (int64_t)(v.c[1] - v.c[0]), v.msize, macrosize, g, w.items,
                       w.nitems, rho);
  }
  OB_HIP(hipGetLastError());
}

}  // namespace ob
