// ob_tiled3d.hip -- tile-sorted 3D charge deposition (Grid3D::binValue, Grid3D.cc:307-432)
// and potential gather for the kick (SpaceChargeCalc3D.cc:120-147).
//
// Scattering 27 FP64 atomics per particle in random particle order over a 64 MiB grid
// misses L2 on almost every update, and gathering 27 potential values per particle in
// particle order touches a different cache line per lane.  Instead, each SC step:
//   1. tile_hist    key = T x T x Tz box of the particle's TSC centre node; per-block LDS
//                   histogram, merged into global tile counts;
//   2. tile_scan    exclusive scan of the counts; work items of at most kChunk particles
//                   per tile;
//   3. tile_scatter counting-sort scatter of the bunch (6 coordinates, macrosize and the
//                   user index) into the spare buffers, ranges reserved per block;
//   4. tile_deposit one workgroup per work item accumulates its particles into an LDS
//                   block of the box plus a one-node halo, then adds the block to rho with
//                   atomics;
//   5. tile_kick    after the Poisson solve, one workgroup per work item copies the same
//                   block of phi into LDS and computes its particles' kicks from it.
// The blocks are small (~14 KiB for 8 x 8 x 16 boxes), so many workgroups share a CU.
// Weights, indices and the kick arithmetic are the reference's (same device helpers as
// the unsorted path); only the FP64 summation order of rho differs.
#include "ob_internal.hpp"
#include "ob_sc_device.hpp"
#include "ob_tiled3d.hpp"

namespace ob {
namespace {

constexpr int kThreads = 256;
constexpr int kChunk = 1024;
constexpr int kMaxTiles = 16383;  // (ntiles + 1) 32-bit counters in 64 KiB of LDS (the scan: 2x)

__device__ __forceinline__ bool inside(const TileGeom& g, double x, double y, double z) {
  return !(x < g.xMin || x > g.xMax || y < g.yMin || y > g.yMax || z < g.zMin || z > g.zMax);
}

__device__ __forceinline__ uint32_t tile_key(const TileGeom& g, double x, double y, double z) {
  if (!inside(g, x, y, z)) return (uint32_t)g.ntiles;
  int iX, iY, iZ;
  double xF, yF, zF;
  node_ind_frac(x, g.xMin, g.dx, g.nx, iX, xF);
  node_ind_frac(y, g.yMin, g.dy, g.ny, iY, yF);
  grid3d_z_ind_frac(z, g.zMin, g.dz, g.nz, iZ, zF);
  return (uint32_t)(((iX / g.T) * g.ty + (iY / g.T)) * g.tz + iZ / g.Tz);
}

// Cell c = (lx*W + ly)*Wz + lz of a tile block, advanced by kThreads cells per step with
// carries instead of a division per cell (the divisions are done once per thread).
struct CellWalk {
  int lx, ly, lz, sx, sy, sz, W, Wz;
  __device__ __forceinline__ CellWalk(int c, int W_, int Wz_) : W(W_), Wz(Wz_) {
    lz = c % Wz;
    const int r = c / Wz;
    ly = r % W;
    lx = r / W;
    sz = kThreads % Wz;
    const int sr = kThreads / Wz;
    sy = sr % W;
    sx = sr / W;
  }
  __device__ __forceinline__ void next() {
    lz += sz;
    int carry = lz >= Wz ? 1 : 0;
    lz -= carry * Wz;
    ly += sy + carry;  // < 2W
    carry = ly >= W ? 1 : 0;
    ly -= carry * W;
    lx += sx + carry;
  }
};

// Tile t's LDS block: nodes x0..x0+W-1, y0..y0+W-1, z0..z0+Wz-1, z fastest.
struct TileBox {
  int x0, y0, z0, W, Wz;
  __device__ __forceinline__ TileBox(const TileGeom& g, int t) {
    const int tzi = t % g.tz, r = t / g.tz;
    const int tyi = r % g.ty, txi = r / g.ty;
    W = g.T + 2;
    Wz = g.Tz + 2 * g.zh;
    x0 = txi * g.T - 1;
    y0 = tyi * g.T - 1;
    z0 = tzi * g.Tz - g.zh;
  }
  __device__ __forceinline__ int cells() const { return W * W * Wz; }
  __device__ __forceinline__ int idx(int gx, int gy, int gz) const {
    return ((gx - x0) * W + (gy - y0)) * Wz + (gz - z0);
  }
  // nodes gx-1..gx+1, gy-1..gy+1, zlo..zhi inside the block
  __device__ __forceinline__ bool holds(int gx, int gy, int zlo, int zhi) const {
    return gx - 1 >= x0 && gx + 1 < x0 + W && gy - 1 >= y0 && gy + 1 < y0 + W && zlo >= z0 &&
           zhi < z0 + Wz;
  }
};

__global__ __launch_bounds__(1024) void tile_hist(const double* __restrict__ cur, int64_t cap,
                                                      int64_t n, TileGeom g,
                                                      uint32_t* __restrict__ key,
                                                      uint32_t* __restrict__ counts) {
  extern __shared__ uint32_t lh[];
  const int nb = g.ntiles + 1;
  for (int t = threadIdx.x; t < nb; t += blockDim.x) lh[t] = 0;
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  // four particles per round: their 12 loads are in flight together
  constexpr int U = 4;
  for (; i + (U - 1) * stride < n; i += U * stride) {
    double x[U], y[U], z[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t j = i + u * stride;
      x[u] = cur[j];
      y[u] = cur[2 * cap + j];
      z[u] = cur[4 * cap + j];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const uint32_t k = tile_key(g, x[u], y[u], z[u]);
      key[i + u * stride] = k;
      atomicAdd(&lh[k], 1u);
    }
  }
  for (; i < n; i += stride) {
    const uint32_t k = tile_key(g, cur[i], cur[2 * cap + i], cur[4 * cap + i]);
    key[i] = k;
    atomicAdd(&lh[k], 1u);
  }
  __syncthreads();
  for (int t = threadIdx.x; t < nb; t += blockDim.x)
    if (lh[t]) atomicAdd(&counts[t], lh[t]);
}

// One block.  counts[0..ntiles] (the last = particles outside the grid) -> offsets
// (exclusive) and cursor = offsets; work items of at most kChunk particles per tile, in tile
// order; counts are cleared for the next step.  The counts are staged through LDS: thread i
// loads tiles i, i + 1024, ... (coalesced, all its loads in flight together), the scan runs
// over contiguous ranges read back from LDS, and the stores go out tile-strided again, so
// neighbouring lanes write neighbouring offsets and items.  LDS: 2 * (ntiles + 2) counters.
__global__ __launch_bounds__(1024) void tile_scan(uint32_t* __restrict__ counts, int ntiles,
                                                  uint32_t* __restrict__ offsets,
                                                  uint32_t* __restrict__ cursor,
                                                  int4* __restrict__ items,
                                                  uint32_t* __restrict__ nitems) {
  extern __shared__ uint32_t ls[];
  const int nb = ntiles + 1;
  uint32_t* loff = ls;            // [nb + 1]: counts, then exclusive particle offsets
  uint32_t* litem = ls + nb + 1;  // [nb + 1]: exclusive item offsets
  constexpr int kPer = (kMaxTiles + 1) / 1024;
  uint32_t cv[kPer];
#pragma unroll
  for (int u = 0; u < kPer; u++) {
    const int t = (int)threadIdx.x + u * 1024;
    cv[u] = counts[t < nb ? t : 0];
  }
#pragma unroll
  for (int u = 0; u < kPer; u++) {
    const int t = (int)threadIdx.x + u * 1024;
    if (t < nb) loff[t] = cv[u];
  }
  __syncthreads();
  const int per = (nb + blockDim.x - 1) / blockDim.x;
  const int lo = min(nb, (int)threadIdx.x * per), hi = min(nb, lo + per);
  uint32_t off = 0, item = 0;
  for (int t = lo; t < hi; t++) {
    const uint32_t k = loff[t];
    off += k;
    if (t < ntiles) item += (k + kChunk - 1) / kChunk;
  }
  uint32_t total, total_items;
  block_scan2(off, item, total, total_items);
  // each thread rewrites only the entries it alone has read
  for (int t = lo; t < hi; t++) {
    const uint32_t k = loff[t];
    loff[t] = off;
    litem[t] = item;
    off += k;
    if (t < ntiles) item += (k + kChunk - 1) / kChunk;
  }
  if (threadIdx.x == 0) {
    loff[nb] = total;
    *nitems = total_items;
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < kPer; u++) {
    const int t = (int)threadIdx.x + u * 1024;
    if (t >= nb) break;
    const uint32_t o = loff[t], e = loff[t + 1];
    offsets[t] = o;
    cursor[t] = o;
    counts[t] = 0;
    if (t < ntiles) {
      uint32_t j = litem[t];
      for (uint32_t b0 = o; b0 < e; b0 += kChunk)
        items[j++] = make_int4(t, (int)b0, (int)min(e, b0 + kChunk), 0);
    }
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
  // U particles per thread and round: all their loads are issued before any store
  constexpr int U = 4;
  for (int64_t i0 = lo + threadIdx.x; i0 < hi; i0 += U * (int64_t)blockDim.x) {
    double v[U][6], m[U];
    uint32_t kk[U];
    int32_t pp[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t i = i0 + u * (int64_t)blockDim.x;
      const int64_t j = i < hi ? i : lo;  // a masked lane reloads the first particle
      kk[u] = key[j];
#pragma unroll
      for (int k = 0; k < 6; k++) v[u][k] = cur[k * cap + j];
      pp[u] = perm[j];
      m[u] = msize ? msize[j] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t i = i0 + u * (int64_t)blockDim.x;
      if (i >= hi) break;
      const int64_t d = atomicAdd(&lc[kk[u]], 1u);
#pragma unroll
      for (int k = 0; k < 6; k++) alt[k * cap + d] = v[u][k];
      perm_alt[d] = pp[u];
      if (msize) msize_alt[d] = m[u];
    }
  }
}

// Grid3D::binValue for one particle into a z-fastest window (global rho or an LDS block).
__device__ __forceinline__ void deposit_particle(double* __restrict__ base, int64_t sx,
                                                 int64_t sy, int nz, int zoff, double m, double xF,
                                                 double yF, int iZ, double zF) {
  const double wx[3] = {0.5 * (0.5 - xF) * (0.5 - xF), 0.75 - xF * xF, 0.5 * (0.5 + xF) * (0.5 + xF)};
  const double wy[3] = {0.5 * (0.5 - yF) * (0.5 - yF), 0.75 - yF * yF, 0.5 * (0.5 + yF) * (0.5 + yF)};
  // base: node (iX-1, iY-1, 0) of the window; zoff: window z origin
  if (nz >= 3) {
    int iZm = iZ - 1, iZp = iZ + 1;
    if (iZm < 0) iZm = nz - 1;
    if (iZp >= nz) iZp = 0;
    const double wz[3] = {0.5 * (0.5 - zF) * (0.5 - zF), 0.75 - zF * zF,
                          0.5 * (0.5 + zF) * (0.5 + zF)};
    const int zz[3] = {iZm - zoff, iZ - zoff, iZp - zoff};
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const double tmp = wy[c] * wz[a] * m;
#pragma unroll
        for (int d = 0; d < 3; d++) atomicAdd(base + d * sx + c * sy + zz[a], wx[d] * tmp);
      }
  } else if (nz == 2) {
    const double wz[2] = {1.0 - zF, zF};
    for (int a = 0; a < 2; a++)
      for (int c = 0; c < 3; c++) {
        const double tmp = wy[c] * wz[a] * m;
        for (int d = 0; d < 3; d++) atomicAdd(base + d * sx + c * sy + (a - zoff), wx[d] * tmp);
      }
  } else {
    for (int c = 0; c < 3; c++) {
      const double tmp = wy[c] * m;
      for (int d = 0; d < 3; d++) atomicAdd(base + d * sx + c * sy - zoff, wx[d] * tmp);
    }
  }
}

// LDS bank scheduling of the deposit: the 27 LDS atomics of a particle are at fixed word
// offsets from its first node's word, so lanes whose first words differ modulo kBankWords
// (8-byte words) hit distinct banks in all 27.  Half a work item at a time, the particles
// are ranked within their residue class, staged in LDS grouped by class, and deposited in
// rounds where lane l of each 32-lane group takes the next particle of class l (an idle
// lane when the class is used up).  In particle order, 64 lanes fall on the banks about as
// balls in bins.  Staging half an item keeps 4 workgroups per CU.
constexpr int kBankWords = 32;
constexpr int kDepU = kChunk / kThreads;  // particles per thread of a work item
constexpr int kDepStage = kChunk / 2;     // particles staged at a time
static_assert(kDepU % 2 == 0, "deposit halves");

// staging after the accumulation block: xF, yF, zF, m (doubles), first-node word << 16 |
// z index (int)
__host__ __device__ inline int tile_dep_acc_doubles(int cells) { return (cells + 1) & ~1; }

__global__ __launch_bounds__(kThreads) void tile_deposit(const double* __restrict__ cur,
                                                         int64_t cap,
                                                         const double* __restrict__ msize,
                                                         double macrosize, TileGeom g,
                                                         const int4* __restrict__ items,
                                                         const uint32_t* __restrict__ nitems,
                                                         double* __restrict__ rho) {
  if (blockIdx.x >= *nitems) return;
  extern __shared__ double acc[];
  __shared__ uint32_t rcount[kBankWords], roff[kBankWords], rmax;
  const int4 it = items[blockIdx.x];
  const TileBox box(g, it.x);
  const int nz = g.nz, cells = box.cells();
  double* sxf = acc + tile_dep_acc_doubles(cells);
  double* syf = sxf + kDepStage;
  double* szf = syf + kDepStage;
  double* smm = szf + kDepStage;
  int* sbz = (int*)(smm + kDepStage);
  const int64_t gsx = (int64_t)g.ny * nz;
  const int np = it.z - it.y;  // at most kChunk
  // all of a thread's particle loads go out before the block is cleared
  double px[kDepU], py[kDepU], pz[kDepU], pm[kDepU];
#pragma unroll
  for (int u = 0; u < kDepU; u++) {
    const int j = (int)threadIdx.x + u * kThreads;
    const int64_t i = it.y + (j < np ? j : 0);
    px[u] = cur[i];
    py[u] = cur[2 * cap + i];
    pz[u] = cur[4 * cap + i];
    pm[u] = msize ? msize[i] : macrosize;
  }
  for (int c = threadIdx.x; c < cells; c += blockDim.x) acc[c] = 0.0;
#pragma unroll
  for (int h = 0; h < 2; h++) {
    if (threadIdx.x < kBankWords) rcount[threadIdx.x] = 0u;
    __syncthreads();  // (h = 0: the block is clear)
    // rank each particle within its residue class (particles outside the block: global)
    constexpr int H = kDepU / 2;
    int bz[H];
    double xF[H], yF[H], zF[H];
    uint32_t res[H], rank[H];
#pragma unroll
    for (int v = 0; v < H; v++) {
      const int u = h * H + v;
      const int j = (int)threadIdx.x + u * kThreads;
      res[v] = kBankWords;  // none
      if (j >= np) continue;
      int iX, iY, iZ;
      node_ind_frac(px[u], g.xMin, g.dx, g.nx, iX, xF[v]);
      node_ind_frac(py[u], g.yMin, g.dy, g.ny, iY, yF[v]);
      grid3d_z_ind_frac(pz[u], g.zMin, g.dz, nz, iZ, zF[v]);
      const int zlo = nz >= 3 ? iZ - 1 : 0, zhi = nz >= 3 ? iZ + 1 : nz - 1;
      if (!box.holds(iX, iY, zlo, zhi)) {  // not in this tile's block (cannot happen for a
        // consistent sort): global
        deposit_particle(rho + (iX - 1) * gsx + (int64_t)(iY - 1) * nz, gsx, nz, nz, 0, pm[u],
                         xF[v], yF[v], iZ, zF[v]);
        continue;
      }
      const int b0 = box.idx(iX - 1, iY - 1, box.z0);  // < 2^15 (tile blocks <= 116 KiB)
      bz[v] = (b0 << 16) | iZ;
      const int w0 = b0 + (nz >= 3 ? iZ - 1 - box.z0 : -box.z0);
      res[v] = (uint32_t)w0 & (kBankWords - 1);
      rank[v] = atomicAdd(&rcount[res[v]], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 64) {  // one wave: exclusive scan and maximum of the class counts
      const int l = threadIdx.x;
      const uint32_t c = l < kBankWords ? rcount[l] : 0u;
      uint32_t sc = c, mx = c;
      for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(sc, o, 64);
        if (l >= o) sc += t;
        mx = max(mx, __shfl_xor(mx, o, 64));
      }
      if (l < kBankWords) roff[l] = sc - c;
      if (l == 0) rmax = mx;
    }
    __syncthreads();
#pragma unroll
    for (int v = 0; v < H; v++) {
      if (res[v] >= (uint32_t)kBankWords) continue;
      const uint32_t q = roff[res[v]] + rank[v];
      sxf[q] = xF[v];
      syf[q] = yF[v];
      szf[q] = zF[v];
      smm[q] = pm[h * H + v];
      sbz[q] = bz[v];
    }
    __syncthreads();
    const int l = (int)threadIdx.x & (kBankWords - 1);
    const int grp = (int)threadIdx.x / kBankWords, ngrp = kThreads / kBankWords;
    const uint32_t cl = rcount[l], ol = roff[l], rm = rmax;
    for (uint32_t k = grp; k < rm; k += ngrp) {
      if (k >= cl) continue;
      const uint32_t q = ol + k;
      const int b = sbz[q];
      deposit_particle(acc + (b >> 16), box.W * box.Wz, box.Wz, nz, box.z0, smm[q], sxf[q],
                       syf[q], b & 0xFFFF, szf[q]);
    }
    __syncthreads();  // every wave is done with the counters and the staging of this half
  }
  CellWalk w((int)threadIdx.x, box.W, box.Wz);
  for (int c = threadIdx.x; c < cells; c += kThreads, w.next()) {
    const double v = acc[c];
    if (v == 0.0) continue;
    const int gx = box.x0 + w.lx, gy = box.y0 + w.ly, gz = box.z0 + w.lz;
    if (gx < 0 || gx >= g.nx || gy < 0 || gy >= g.ny || gz < 0 || gz >= nz) continue;
    atomicAdd(rho + gx * gsx + (int64_t)gy * nz + gz, v);
  }
}

// SpaceChargeCalc3D kick of particle i, gathering phi through view v (which must hold the
// 27 nodes) or the whole grid.
// BND: particles outside the boundary in k are not kicked.
// The kick of particle i at (x, y, z) with momenta (xp, yp, dE) already read: writes
// xp, yp, dE (the rest of kick_particle).
template <bool BND>
__device__ __forceinline__ void kick_particle_all(double* __restrict__ cur, int64_t cap,
                                                  int64_t i, double x, double y, double z,
                                                  double xp, double yp, double dE,
                                                  const TileGeom& g, const Kick3D& k,
                                                  const double* lds, const TileBox* box);

template <bool BND>
__device__ __forceinline__ void kick_particle_xyz(double* __restrict__ cur, int64_t cap,
                                                  int64_t i, double x, double y, double z,
                                                  const TileGeom& g, const Kick3D& k,
                                                  const double* lds, const TileBox* box) {
  if (BND && !boundary_inside(k.bnd_type, k.bnd, x, y)) return;
  const double zr = (z - k.zc) * k.gamma + k.zc;
  int iX, iY, iZ;
  double xF, yF, zF;
  node_ind_frac(x, g.xMin, g.dx, g.nx, iX, xF);
  node_ind_frac(y, g.yMin, g.dy, g.ny, iY, yF);
  grid3d_z_ind_frac(zr, k.rzMin, k.rdz, g.nz, iZ, zF);
  const int zlo = g.nz >= 3 ? iZ - 1 : 0, zhi = g.nz >= 3 ? iZ + 1 : g.nz - 1;
  double ex, ey, ez;
  if (box && box->holds(iX, iY, zlo, zhi)) {
    const Grid3View v{lds, (int64_t)box->W * box->Wz, box->Wz, box->x0, box->y0, box->z0};
    grid3d_gradient_at(v, g.nz, iX, iY, iZ, xF, yF, zF, g.dx, g.dy, k.rdz, ex, ey, ez);
  } else {
    const Grid3View v{k.phi, (int64_t)g.ny * g.nz, g.nz, 0, 0, 0};
    grid3d_gradient_at(v, g.nz, iX, iY, iZ, xF, yF, zF, g.dx, g.dy, k.rdz, ex, ey, ez);
  }
  cur[cap + i] += -ex * k.trans_factor;
  cur[3 * cap + i] += -ey * k.trans_factor;
  cur[5 * cap + i] += -ez * k.long_factor;
}

template <bool BND>
__device__ __forceinline__ void kick_particle_all(double* __restrict__ cur, int64_t cap,
                                                  int64_t i, double x, double y, double z,
                                                  double xp, double yp, double dE,
                                                  const TileGeom& g, const Kick3D& k,
                                                  const double* lds, const TileBox* box) {
  if (BND && !boundary_inside(k.bnd_type, k.bnd, x, y)) return;
  const double zr = (z - k.zc) * k.gamma + k.zc;
  int iX, iY, iZ;
  double xF, yF, zF;
  node_ind_frac(x, g.xMin, g.dx, g.nx, iX, xF);
  node_ind_frac(y, g.yMin, g.dy, g.ny, iY, yF);
  grid3d_z_ind_frac(zr, k.rzMin, k.rdz, g.nz, iZ, zF);
  const int zlo = g.nz >= 3 ? iZ - 1 : 0, zhi = g.nz >= 3 ? iZ + 1 : g.nz - 1;
  double ex, ey, ez;
  if (box->holds(iX, iY, zlo, zhi)) {
    const Grid3View v{lds, (int64_t)box->W * box->Wz, box->Wz, box->x0, box->y0, box->z0};
    grid3d_gradient_at(v, g.nz, iX, iY, iZ, xF, yF, zF, g.dx, g.dy, k.rdz, ex, ey, ez);
  } else {
    const Grid3View v{k.phi, (int64_t)g.ny * g.nz, g.nz, 0, 0, 0};
    grid3d_gradient_at(v, g.nz, iX, iY, iZ, xF, yF, zF, g.dx, g.dy, k.rdz, ex, ey, ez);
  }
  cur[cap + i] = xp + -ex * k.trans_factor;
  cur[3 * cap + i] = yp + -ey * k.trans_factor;
  cur[5 * cap + i] = dE + -ez * k.long_factor;
}

template <bool BND>
__device__ __forceinline__ void kick_particle(double* __restrict__ cur, int64_t cap, int64_t i,
                                              const TileGeom& g, const Kick3D& k,
                                              const double* lds, const TileBox* box) {
  kick_particle_xyz<BND>(cur, cap, i, cur[i], cur[2 * cap + i], cur[4 * cap + i], g, k, lds,
                         box);
}

// Blocks 0..nitems-1: work items; block nitems: the particles outside the grid (last in the
// sorted order, from offsets[ntiles]), kicked from the whole grid like the reference.
// At least 3 waves per SIMD: the kick with the next particle's prefetch takes 157 VGPRs
// without spills.  Bounded to 4 waves (128 VGPRs) the prefetch spills and the kick ran
// 133 us against 113; without the prefetch, 119.
template <bool BND>
__global__ __launch_bounds__(kThreads, 3) void tile_kick(double* __restrict__ cur, int64_t cap,
                                                      int64_t n, TileGeom g, Kick3D k,
                                                      const int4* __restrict__ items,
                                                      const uint32_t* __restrict__ nitems,
                                                      const uint32_t* __restrict__ offsets) {
  const uint32_t ni = *nitems;
  if (blockIdx.x > ni) return;
  if (blockIdx.x == ni) {
    for (int64_t i = (int64_t)offsets[g.ntiles] + threadIdx.x; i < n; i += blockDim.x)
      kick_particle<BND>(cur, cap, i, g, k, nullptr, nullptr);
    return;
  }
  extern __shared__ double sphi[];
  const int4 it = items[blockIdx.x];
  const TileBox box(g, it.x);
  const int cells = box.cells();
  const int64_t gsx = (int64_t)g.ny * g.nz;
  // The first particle's coordinates are requested with the block of phi, and all of a
  // thread's block loads are issued before any is stored (a masked lane loads phi[0] and
  // stores 0), so these latencies overlap instead of adding up.
  const int i0 = it.y + (int)threadIdx.x;
  const int ia = i0 < it.z ? i0 : it.y;
  const double px = cur[ia], py = cur[2 * cap + ia], pz = cur[4 * cap + ia];
  const double pxp = cur[cap + ia], pyp = cur[3 * cap + ia], pdE = cur[5 * cap + ia];
  constexpr int KB = 8;  // cells <= KB * kThreads for blocks of up to 16 KiB
  double bv[KB];
  uint32_t inmask = 0;
  CellWalk w((int)threadIdx.x, box.W, box.Wz);
#pragma unroll
  for (int u = 0; u < KB; u++, w.next()) {
    const int c = (int)threadIdx.x + u * kThreads;
    const int gx = box.x0 + w.lx, gy = box.y0 + w.ly, gz = box.z0 + w.lz;
    const bool in = c < cells && gx >= 0 && gx < g.nx && gy >= 0 && gy < g.ny && gz >= 0 &&
                    gz < g.nz;
    bv[u] = k.phi[in ? gx * gsx + (int64_t)gy * g.nz + gz : 0];
    inmask |= (in ? 1u : 0u) << u;
  }
#pragma unroll
  for (int u = 0; u < KB; u++) {
    const int c = (int)threadIdx.x + u * kThreads;
    if (c < cells) sphi[c] = ((inmask >> u) & 1u) ? bv[u] : 0.0;
  }
  for (int c = (int)threadIdx.x + KB * kThreads; c < cells; c += kThreads, w.next()) {
    // larger blocks: the walk continues from cell threadIdx.x + KB * kThreads
    const int gx = box.x0 + w.lx, gy = box.y0 + w.ly, gz = box.z0 + w.lz;
    const bool in = gx >= 0 && gx < g.nx && gy >= 0 && gy < g.ny && gz >= 0 && gz < g.nz;
    sphi[c] = in ? k.phi[gx * gsx + (int64_t)gy * g.nz + gz] : 0.0;
  }
  __syncthreads();
  // The next particle's coordinates are loaded before this one is kicked.
  double q[6] = {px, py, pz, pxp, pyp, pdE};
  for (int i = i0; i < it.z; i += kThreads) {
    const int in = i + kThreads < it.z ? i + kThreads : i;
    const double nq[6] = {cur[in], cur[2 * cap + in], cur[4 * cap + in],
                          cur[cap + in], cur[3 * cap + in], cur[5 * cap + in]};
    kick_particle_all<BND>(cur, cap, i, q[0], q[1], q[2], q[3], q[4], q[5], g, k, sphi, &box);
#pragma unroll
    for (int c = 0; c < 6; c++) q[c] = nq[c];
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
  g.T = g.Tz = g.zh = g.tx = g.ty = g.tz = g.ntiles = 0;
  // Small boxes first (many workgroups per CU); larger ones only to bound the tile count.
  static const int shapes[][2] = {{8, 16}, {8, 8},  {16, 16}, {8, 32}, {16, 32},
                                  {32, 32}, {6, 8}, {4, 8},   {4, 4},  {2, 2},  {1, 1}};
  for (const auto& sh : shapes) {
    const int T = sh[0];
    const int Tz = nz <= 2 ? nz : std::min(sh[1], nz);
    const int zh = nz >= 3 ? 1 : 0;
    const int64_t tx = (nx + T - 1) / T, ty = (ny + T - 1) / T, tz = (nz + Tz - 1) / Tz;
    const int64_t ntiles = tx * ty * tz;
    const int64_t lds = (int64_t)(T + 2) * (T + 2) * (Tz + 2 * zh) * 8;
    if (lds <= lds_budget && ntiles <= kMaxTiles) {
      g.T = T; g.Tz = Tz; g.zh = zh;
      g.tx = (int)tx; g.ty = (int)ty; g.tz = (int)tz; g.ntiles = (int)ntiles;
      break;
    }
  }
  return g;
}

size_t tile_deposit_lds_bytes(const TileGeom& g) {
  const int cells = (g.T + 2) * (g.T + 2) * (g.Tz + 2 * g.zh);
  return sizeof(double) * tile_dep_acc_doubles(cells) +
         (size_t)kDepStage * (4 * sizeof(double) + sizeof(int));
}

size_t tile_lds_bytes(const TileGeom& g) {
  return sizeof(double) * (g.T + 2) * (g.T + 2) * (g.Tz + 2 * g.zh);
}

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
    OB_HIP(hipFuncSetAttribute((const void*)tile_deposit,  // (plus its static class counters)
                               hipFuncAttributeMaxDynamicSharedMemorySize, 159 * 1024));
    OB_HIP(hipFuncSetAttribute((const void*)tile_hist,
                               hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    OB_HIP(hipFuncSetAttribute((const void*)tile_scatter,
                               hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    OB_HIP(hipFuncSetAttribute((const void*)tile_scan,
                               hipFuncAttributeMaxDynamicSharedMemorySize, 136 * 1024));
    OB_HIP(hipFuncSetAttribute((const void*)tile_kick<false>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    OB_HIP(hipFuncSetAttribute((const void*)tile_kick<true>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    attr_done = true;
  }
  const int64_t n = bunch_size(b);
  OB_REQUIRE((int64_t)(g.ntiles + 1) * 4 <= 150 * 1024, "tile sort: too many tiles for LDS");
  OB_REQUIRE(tile_deposit_lds_bytes(g) <= 159 * 1024, "tile deposit: tile block too large for LDS");
  OB_REQUIRE(g.nz < 65536, "tile deposit: more than 65535 z nodes");
  OB_REQUIRE(n < (int64_t)1 << 31, "tile sort: more than 2^31 particles");
  w.reserve(n, g.ntiles);
  hipStream_t s = stream();
  SortBuffers sb = bunch_sort_begin(b);
  const size_t hl = sizeof(uint32_t) * (g.ntiles + 1);
  // one workgroup of 1024 threads per CU: each merges up to ntiles + 1 counters into the
  // global counts, so fewer, larger workgroups mean fewer global atomics (measured 38 -> 32 us)
  int hb = (int)std::min<int64_t>((n + 1023) / 1024, device_cu_count());
  if (hb < 1) hb = 1;
  {
    ProfScope ps("tile_hist");
    hipLaunchKernelGGL(tile_hist, dim3(hb), dim3(1024), hl, s, sb.cur, sb.cap, n, g, w.key,
                       w.counts);
  }
  {
    ProfScope ps("tile_scan");
    hipLaunchKernelGGL(tile_scan, dim3(1), dim3(1024), sizeof(uint32_t) * 2 * (g.ntiles + 2), s,
                       w.counts, g.ntiles, w.offsets,
                       w.cursor, w.items, w.nitems);
  }
  // about 4096 workgroups: measured fastest for the scatter and for the deposit after it
  const int64_t per = std::max<int64_t>(512, (n + 4095) / 4096);
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
    hipLaunchKernelGGL(tile_deposit, dim3((unsigned)max_items), dim3(kThreads),
                       tile_deposit_lds_bytes(g),
                       s, v.c[0], (int64_t)(v.c[1] - v.c[0]), v.msize, macrosize, g, w.items,
                       w.nitems, rho);
  }
  OB_HIP(hipGetLastError());
}

void tile_kick_3d(ob_bunch* b, const TileGeom& g, const Kick3D& k, TileWork& w) {
  const int64_t n = bunch_size(b);
  if (n == 0) return;
  DeviceBunchView v = bunch_device_view(b);
  const int64_t max_items = n / kChunk + g.ntiles + 1;
  {
    ProfScope ps("tile_kick");
    if (k.bnd_type)
      hipLaunchKernelGGL(tile_kick<true>, dim3((unsigned)(max_items + 1)), dim3(kThreads),
                         tile_lds_bytes(g), stream(), v.c[0], (int64_t)(v.c[1] - v.c[0]), n, g, k,
                         w.items, w.nitems, w.offsets);
    else
      hipLaunchKernelGGL(tile_kick<false>, dim3((unsigned)(max_items + 1)), dim3(kThreads),
                         tile_lds_bytes(g), stream(), v.c[0], (int64_t)(v.c[1] - v.c[0]), n, g, k,
                         w.items, w.nitems, w.offsets);
  }
  OB_HIP(hipGetLastError());
  bunch_device_written(b);
}

}  // namespace ob
