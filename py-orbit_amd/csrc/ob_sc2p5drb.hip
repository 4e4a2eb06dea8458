// ob_sc2p5drb.hip -- SpaceChargeCalc2p5Drb on the device.
//
// Per trackBunch (SpaceChargeCalc2p5Drb.cc:110-168), all on the library stream with no
// host round trip:
//   1. fused map kernel flushes the bunch's queued maps and emits per-block extrema;
//   2. bin     (every block reduces the extrema and sizes the grids, bunchAnalysis
//               :170-213): clears rho, counts particles per tile and writes per-block rows
//               of the periodic 2-point z line density (Grid1D.cc:241-275, 568-626);
//   3. deposit : 9-point TSC charge (Grid2D.cc:120-166).  Grids up to 128x128 nodes: each
//               workgroup deposits its particles into a private copy of the grid in LDS
//               and adds it to rho with FP64 atomics.  Larger grids: counting-sort scatter
//               of x, y by tile, then one LDS tile block per work item;
//   4. post    (1 block): dlambda/dz by 3-point LSQ (:257-310), rho moments (:229-254),
//                         kick factors;
//   5. Poisson (3 kernels, PoissonSolverFFT2D.cc:174-256): row r2c of two rows per complex
//      FFT, column FFT x Green x inverse column FFT (zero-padding pruned: only the nx
//      non-zero rows are read and only the nx needed rows written), row c2r;
//   6. the kick (:146-167) is queued on the bunch and runs fused into the next map kernel.
#include <algorithm>
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
constexpr int kUnroll = 4;  // particles per thread whose loads are issued together
constexpr int kFFTThreads = 256;
// Private-grid single-bunch deposit: grids up to kPrivNodes nodes (128 KiB of LDS),
// bunches up to kPrivMaxParticles, kPrivPerBlock particles per workgroup of kPrivThreads
// (at 100k particles on 128x128: 12.0 us against 18.6 us for the tile sort's scatter and
// deposit; 2048 particles per workgroup 12.6 us, 4096 15.7 us, 512 11.7 us).
constexpr int kPrivNodes = 128 * 128;
constexpr int64_t kPrivMaxParticles = 1 << 18;
constexpr int64_t kPrivPerBlock = 1024;
constexpr int kPrivThreads = 1024;

struct SC2DParams {
  int nx, ny, nz, nblk;       // grid sizes, number of extrema partial blocks
  double xy_ratio;
  double dx_solver;           // solver grid step (fixed at construction, :33)
  double macrosize;           // bunch macroSize (used when no per-particle attribute)
  double factor0;             // 2 L r0 / (beta^2 gamma^3)            (:129)
  double pipe_radius;
  int n_long_avg;
  int plain;                  // 1: SpaceChargeCalc2p5D (ratio rule, transverse kick only)
  int par;                    // SpaceChargeCalc2p5D: state entry this step reads
  int bnd_type;               // SpaceChargeCalc2p5D boundary (SC2DState::bnd_type), 0: none
  double bnd[6];              // its box and sizes (SC2DState::bnd)
};

// ---- 2: grid of the step --------------------------------------------------------
// bunchAnalysis grid sizing from the extrema partials: SpaceChargeCalc2p5Drb.cc:170-213
// (the box always takes the calculator's x/y ratio) or SpaceChargeCalc2p5D.cc:122-216
// (the ratio is kept only while the beam's is within 25% of it; otherwise it follows the
// beam and the solver's Green function is rebuilt on the beam's box).
struct Grid2Step {
  double ext[6];
  double xMin, xMax, yMin, yMax, dx, dy, zMin, zMax, dz;
  int32_t skip;     // zero extent: no deposit, no kick (the reference aborts, :176-189)
  int32_t rebuild;  // SpaceChargeCalc2p5D: rebuild the Green function on sol
  double ratio;     // SpaceChargeCalc2p5D: the calculator's ratio after this step
  double sol[4];    // SpaceChargeCalc2p5D: the solver box after this step
};

// Every thread of the block gets the grid of partials [pb, pe) (blockDim.x <= 1024).
// st is read (entry P.par) only for SpaceChargeCalc2p5D.
__device__ __forceinline__ void sc2d_grid_body(const SC2DParams& P,
                                               const double* __restrict__ partials, int pb,
                                               int pe, const SC2DState* st, Grid2Step& out) {
  __shared__ double red[16][6];
  __shared__ Grid2Step gs;
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
    for (int w = 1; w < (int)(blockDim.x >> 6); w++)
      for (int k = 0; k < 6; k++) r[k] = (k & 1) ? fmax(r[k], red[w][k]) : fmin(r[k], red[w][k]);
    Grid2Step g;
    for (int k = 0; k < 6; k++) g.ext[k] = r[k];
    double xMin = r[0], xMax = r[1], yMin = r[2], yMax = r[3];
    const double zMin = r[4], zMax = r[5];
    g.skip = (xMin >= xMax || yMin >= yMax || zMin >= zMax) ? 1 : 0;
    g.rebuild = 0;
    double ratio = P.xy_ratio;
    if (P.plain) {
      ratio = st->ratio[P.par];
      for (int k = 0; k < 4; k++) g.sol[k] = st->sol[P.par][k];
    }
    if (P.plain && P.bnd_type) {
      // with a boundary the grid is its box; only z comes from the bunch (:174-195)
      xMin = P.bnd[0];
      xMax = P.bnd[1];
      yMin = P.bnd[2];
      yMax = P.bnd[3];
      ratio = (xMax - xMin) / (yMax - yMin);
      g.skip = zMin >= zMax ? 1 : 0;
      if (!g.skip && fabs((g.sol[1] - g.sol[0]) / (g.sol[3] - g.sol[2]) - ratio) > 0.00000001) {
        g.sol[0] = xMin; g.sol[1] = xMax; g.sol[2] = yMin; g.sol[3] = yMax;
        g.rebuild = 1;
      }
    } else if (!g.skip) {
      const double xy_ratio_beam = (xMax - xMin) / (yMax - yMin);
      // 2p5Drb: always (:191-203); 2p5D: within 25% of the ratio (SpaceChargeCalc2p5D.cc:152-167)
      if (!P.plain || fabs(xy_ratio_beam - ratio) < 0.25 * ratio) {
        double width, center;
        if (xy_ratio_beam > ratio) {
          center = (yMax + yMin) / 2.0;
          width = ((yMax - yMin) * (xy_ratio_beam / ratio)) / 2.0;
          yMin = center - width;
          yMax = center + width;
        } else {
          center = (xMax + xMin) / 2.0;
          width = ((xMax - xMin) / (xy_ratio_beam / ratio)) / 2.0;
          xMin = center - width;
          xMax = center + width;
        }
        if (P.plain) ratio = (xMax - xMin) / (yMax - yMin);
      } else {  // SpaceChargeCalc2p5D.cc:168-172
        ratio = xy_ratio_beam;
        g.sol[0] = xMin; g.sol[1] = xMax; g.sol[2] = yMin; g.sol[3] = yMax;
        g.rebuild = 1;
      }
      // SpaceChargeCalc2p5D.cc:206-216: a solver box of another shape is rebuilt on this step's box
      if (P.plain && fabs((g.sol[1] - g.sol[0]) / (g.sol[3] - g.sol[2]) - ratio) > 0.00000001) {
        g.sol[0] = xMin; g.sol[1] = xMax; g.sol[2] = yMin; g.sol[3] = yMax;
        g.rebuild = 1;
      }
    }
    g.ratio = ratio;
    g.xMin = xMin;
    g.xMax = xMax;
    g.dx = (xMax - xMin) / (P.nx - 1);  // Grid2D.cc:409-413
    g.yMin = yMin;
    g.yMax = yMax;
    g.dy = (yMax - yMin) / (P.ny - 1);
    g.zMin = zMin;
    g.zMax = zMax;
    g.dz = (zMax - zMin) / P.nz;        // Grid1D.cc:88-93
    gs = g;
  }
  __syncthreads();
  out = gs;
}

// The step's grid into the state; on a zero-extent bunch the grid of the last good step
// stays (it is not used) and the status word reports it.
__device__ __forceinline__ void sc2d_write_state(SC2DState* st, const Grid2Step& g) {
  for (int k = 0; k < 6; k++) st->ext[k] = g.ext[k];
  st->status = g.skip;
  st->skip = g.skip;
  if (!g.skip) {
    st->xMin = g.xMin;
    st->xMax = g.xMax;
    st->dx = g.dx;
    st->yMin = g.yMin;
    st->yMax = g.yMax;
    st->dy = g.dy;
    st->zMin = g.zMin;
    st->zMax = g.zMax;
    st->dz = g.dz;
  }
}

// SpaceChargeCalc2p5D: the ratio and solver box for the next step (entry par ^ 1), the
// solver steps (PoissonSolverFFT2D::setGridXY, PoissonSolverFFT2D.cc:103-111) and whether
// the Green function passes of this step run (gate->skip = 0).
__device__ __forceinline__ void sc2d_write_solver(const SC2DParams& P, SC2DState* st,
                                                  SC2DState* gate, const Grid2Step& g) {
  const int nxt = P.par ^ 1;
  st->ratio[nxt] = g.ratio;
  for (int k = 0; k < 4; k++) st->sol[nxt][k] = g.sol[k];
  if (g.rebuild) {
    st->sdx = (g.sol[1] - g.sol[0]) / (P.nx - 1);
    st->sdy = (g.sol[3] - g.sol[2]) / (P.ny - 1);
    st->rebuilds += 1;
  }
  gate->skip = g.rebuild ? 0 : 1;
  st->bnd_type = P.bnd_type;
  for (int k = 0; k < 6; k++) st->bnd[k] = P.bnd[k];
}

__device__ __forceinline__ Grid2Step sc2d_load_grid(const SC2DState* state) {
  const gptr<const SC2DState> st = to_global(state);
  Grid2Step g;
  g.xMin = st->xMin; g.xMax = st->xMax; g.dx = st->dx;
  g.yMin = st->yMin; g.yMax = st->yMax; g.dy = st->dy;
  g.zMin = st->zMin; g.zMax = st->zMax; g.dz = st->dz;
  g.skip = st->skip;
  return g;
}

// Batch: one block per bunch writes the grid and clears its grids and scatter cursors.
__device__ __forceinline__ void sc2d_setup_body(const SC2DParams& P,
                                                const double* __restrict__ partials, int pb,
                                                int pe, SC2DState* st, const Tile2D& tg,
                                                double* rho, double* zg, double* zd) {
  Grid2Step g;
  sc2d_grid_body(P, partials, pb, pe, st, g);
  if (threadIdx.x == 0) sc2d_write_state(st, g);
  const int ng = P.nx * P.ny;
  for (int i = threadIdx.x; i < ng; i += blockDim.x) rho[i] = 0.0;
  for (int i = threadIdx.x; i < P.nz; i += blockDim.x) {
    zg[i] = 0.0;
    zd[i] = 0.0;
  }
  for (int t = threadIdx.x; t < tg.ntiles; t += blockDim.x) tg.cursor[t] = 0u;
}

// ---- 3: deposit, tile-sorted ----------------------------------------------------
// Binning particle by particle into the whole grid (Grid2D::binValue, Grid2D.cc:142-166)
// piles every wave's 9 FP64 atomics onto the few nodes of the beam core.  Instead:
//   a. bin     per block: LDS counts of the particles per T x T tile of TSC centre nodes
//              (merged with atomics), and the z line density (Grid1D::binValue,
//              Grid1D.cc:422-440) in LDS, written out per block;
//   b. scatter every block scans the counts, then copies (x, y[, macrosize]) of its
//              in-grid particles into scratch arrays, tile by tile (counting sort; the
//              bunch itself is not reordered); block 0 also writes the work items;
//   c. deposit one workgroup per work item (at most `item` particles of one tile)
//              accumulates its particles in an LDS block of the tile plus a one-node halo,
//              adds the block to rho with atomics and writes the block's rho moments.
// Each contribution is the reference's product; only the FP64 summation order differs.
// The single-bunch path computes the grid inside the bin pass (every block from the
// extrema partials, block 0 writes the state); the batch path has a setup launch.

// Tile of the TSC centre node of (x, y), or -1 outside the grid (no charge deposited).
__device__ __forceinline__ int tile2d_key(const Tile2D& tg, int nx, int ny, const Grid2Step& g,
                                          double x, double y) {
  if (x < g.xMin || x > g.xMax || y < g.yMin || y > g.yMax) return -1;
  int iX, iY;
  double xF, yF;
  node_ind_frac(x, g.xMin, g.dx, nx, iX, xF);
  node_ind_frac(y, g.yMin, g.dy, ny, iY, yF);
  return (iX >> tg.shift) * tg.ty + (iY >> tg.shift);
}

// Particles first, first + stride, ... < end: tile counts (atomics into tg.counts) and the
// z line density, added atomically into zg or written as this block's row zrow[nz].
// LDS: zl[nz] doubles, lh[ntiles] counters.
// kUnroll particles of a hist pass (x, y, z, macro size), loaded together.
struct HistPass {
  double x[kUnroll], y[kUnroll], z[kUnroll], m[kUnroll];
};
__device__ __forceinline__ void hist_load(const DeviceBunchView& v, double macrosize, int64_t i0,
                                          int64_t end, int64_t stride, HistPass& q) {
  const gptr<const double> px = to_global((const double*)v.c[0]);
  const gptr<const double> py = to_global((const double*)v.c[2]);
  const gptr<const double> pz = to_global((const double*)v.c[4]);
  const gptr<const double> pm = to_global((const double*)v.msize);
#pragma unroll
  for (int u = 0; u < kUnroll; u++) {
    const int64_t i = i0 + u * stride;
    const bool ok = i < end;
    q.x[u] = ok ? px[i] : 0.0;
    q.y[u] = ok ? py[i] : 0.0;
    q.z[u] = ok ? pz[i] : 0.0;
    q.m[u] = pm ? (ok ? pm[i] : 0.0) : macrosize;
  }
}

// PIPE: each pass's loads are issued before the previous pass is binned (the single-bunch
// kernel, whose workgroups make one or two passes; the batch's, with many, keep their
// registers for occupancy instead); pre: the first pass's particles, already loaded (by
// hist_load at `first`), or null.  COUNT: count particles per tile (not needed by the
// private-grid deposit).
template <bool PIPE, bool COUNT = true>
__device__ __forceinline__ void sc2d_hist_body(const SC2DParams& P, const DeviceBunchView& v,
                                               const Grid2Step& g, const Tile2D& tg,
                                               double* __restrict__ zg, double* __restrict__ zrow,
                                               int64_t first, int64_t end, int64_t stride,
                                               double* zl, uint32_t* lh,
                                               uint32_t* __restrict__ crow = nullptr,
                                               const HistPass* pre = nullptr) {
  const int nz = P.nz;
  for (int i = threadIdx.x; i < nz; i += blockDim.x) zl[i] = 0.0;
  for (int t = threadIdx.x; t < tg.ntiles; t += blockDim.x) lh[t] = 0u;
  __syncthreads();
  if (!g.skip) {
    // kUnroll particles per thread per pass, their loads all in flight together; the next
    // pass's loads are issued before this one is binned
    HistPass q;
    if (PIPE) {
      if (pre)
        q = *pre;
      else if (first < end)
        hist_load(v, P.macrosize, first, end, stride, q);
    }
    for (int64_t i0 = first; i0 < end; i0 += kUnroll * stride) {
      HistPass c;
      if (PIPE) {
        c = q;
        if (i0 + kUnroll * stride < end)
          hist_load(v, P.macrosize, i0 + kUnroll * stride, end, stride, q);
      } else {
        hist_load(v, P.macrosize, i0, end, stride, c);
      }
#pragma unroll
      for (int u = 0; u < kUnroll; u++) {
        if (i0 + u * stride >= end) break;
        if (COUNT) {
          const int k = tile2d_key(tg, P.nx, P.ny, g, c.x[u], c.y[u]);
          if (k >= 0) atomicAdd(&lh[k], 1u);
        }
        // Grid1D::binValue (Grid1D.cc:422-440)
        if (!(c.z[u] < g.zMin || c.z[u] > g.zMax)) {
          int i0z, ip;
          double w0, wp;
          grid1d_ind_wz(c.z[u], g.zMin, g.dz, nz, i0z, ip, w0, wp);
          if (nz > 1) {
            atomicAdd(&zl[i0z], w0 * c.m[u]);
            atomicAdd(&zl[ip], wp * c.m[u]);
          } else {
            atomicAdd(&zl[0], c.m[u]);
          }
        }
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nz; i += blockDim.x) {
    if (zrow)
      zrow[i] = zl[i];
    else if (zl[i] != 0.0)
      atomicAdd(&zg[i], zl[i]);
  }
  for (int t = threadIdx.x; t < tg.ntiles; t += blockDim.x) {
    if (crow) crow[t] = lh[t];
    if (lh[t]) atomicAdd(&tg.counts[t], lh[t]);
  }
}

// Exclusive scan of the tile counts into loff[ntiles] (LDS) by every block; the writer
// block also writes the work items and their number.
__device__ __forceinline__ void tile2d_scan_body(const Tile2D& tg, uint32_t* loff, bool writer) {
  const int per = (tg.ntiles + blockDim.x - 1) / blockDim.x;
  const int lo = min(tg.ntiles, (int)threadIdx.x * per), hi = min(tg.ntiles, lo + per);
  uint32_t s = 0, c = 0;
  for (int t = lo; t < hi; t++) {
    const uint32_t k = tg.counts[t];
    s += k;
    c += (k + tg.item - 1) / tg.item;
  }
  uint32_t ts, tc;
  block_scan2(s, c, ts, tc);
  for (int t = lo; t < hi; t++) {
    const uint32_t k = tg.counts[t];
    loff[t] = s;
    if (writer)
      for (uint32_t c0 = 0; c0 < k; c0 += tg.item)
        tg.items[c++] = make_int4(t, (int)(s + c0), (int)(s + min(k, c0 + (uint32_t)tg.item)), 0);
    s += k;
  }
  if (writer && threadIdx.x == 0) *tg.nitems = tc;
  __syncthreads();
}

// Particles [lo, hi) of one bunch.  LDS: loff[ntiles] (tile offsets), lc[ntiles].
// crow: the range's tile counts, kept by the bin pass over the same range (batch path);
// without it the range is counted here first.
// PIPE: as for sc2d_hist_body.
template <bool PIPE>
__device__ __forceinline__ void sc2d_scatter_body(const SC2DParams& P, const DeviceBunchView& v,
                                                  const Grid2Step& g, const Tile2D& tg,
                                                  int64_t lo, int64_t hi, uint32_t* loff,
                                                  uint32_t* lc, bool writer,
                                                  const uint32_t* __restrict__ crow = nullptr) {
  const bool run = !g.skip;
  const gptr<const double> px = to_global((const double*)v.c[0]);
  const gptr<const double> py = to_global((const double*)v.c[2]);
  const gptr<const double> pm = to_global((const double*)v.msize);
  const gptr<dpair> sxy = to_global((dpair*)tg.sxy);
  const gptr<double> sm = to_global(tg.sm);
  const int64_t stride = blockDim.x;
  // kUnroll particles per thread per pass, loaded together; the first pass's loads go out
  // before the tile scan, each next pass's before this one is processed
  auto load = [&](int64_t i0, double (&x)[kUnroll], double (&y)[kUnroll], double (&m)[kUnroll]) {
#pragma unroll
    for (int u = 0; u < kUnroll; u++) {
      const int64_t i = i0 + u * stride;
      x[u] = i < hi ? px[i] : 0.0;
      y[u] = i < hi ? py[i] : 0.0;
      m[u] = (pm && i < hi) ? pm[i] : 0.0;
    }
  };
  double nx_[kUnroll], ny_[kUnroll], nm_[kUnroll];
  if (PIPE) load(lo + threadIdx.x, nx_, ny_, nm_);
  tile2d_scan_body(tg, loff, writer);
  for (int t = threadIdx.x; t < tg.ntiles; t += blockDim.x) lc[t] = crow ? crow[t] : 0u;
  __syncthreads();
  if (run && !crow)
    for (int64_t i0 = lo + threadIdx.x; i0 < hi; i0 += kUnroll * stride) {
      double x[kUnroll], y[kUnroll], m[kUnroll];
      if (PIPE && i0 == lo + threadIdx.x) {
#pragma unroll
        for (int u = 0; u < kUnroll; u++) { x[u] = nx_[u]; y[u] = ny_[u]; }
      } else {
        load(i0, x, y, m);
      }
#pragma unroll
      for (int u = 0; u < kUnroll; u++) {
        if (i0 + u * stride >= hi) break;
        const int k = tile2d_key(tg, P.nx, P.ny, g, x[u], y[u]);
        if (k >= 0) atomicAdd(&lc[k], 1u);
      }
    }
  __syncthreads();
  for (int t = threadIdx.x; t < tg.ntiles; t += blockDim.x) {
    const uint32_t c = lc[t];
    lc[t] = c ? loff[t] + atomicAdd(&tg.cursor[t], c) : 0u;
  }
  __syncthreads();
  if (run)
    for (int64_t i0 = lo + threadIdx.x; i0 < hi; i0 += kUnroll * stride) {
      double x[kUnroll], y[kUnroll], m[kUnroll];
      if (PIPE) {
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
          x[u] = nx_[u];
          y[u] = ny_[u];
          m[u] = nm_[u];
        }
        if (i0 + kUnroll * stride < hi) load(i0 + kUnroll * stride, nx_, ny_, nm_);
      } else {
        load(i0, x, y, m);
      }
#pragma unroll
      for (int u = 0; u < kUnroll; u++) {
        if (i0 + u * stride >= hi) break;
        const int k = tile2d_key(tg, P.nx, P.ny, g, x[u], y[u]);
        if (k < 0) continue;
        const uint32_t d = atomicAdd(&lc[k], 1u);
        sxy[d] = dpair{x[u], y[u]};  // one 16-byte store per particle
        if (pm) sm[d] = m[u];
      }
    }
}

// rho moments of one node (SpaceChargeCalc2p5Drb.cc:229-238); quirk Q1: the y position
// is getGridX(iy).
__device__ __forceinline__ void moment_add(double s[5], double val, int gx, int gy, double xMin,
                                           double dx) {
  const double x = xMin + gx * dx;
  const double y = xMin + gy * dx;
  s[0] += val;
  s[1] += x * val;
  s[2] += x * x * val;
  s[3] += y * val;
  s[4] += y * y * val;
}

// Grid2D::binValue (Grid2D.cc:142-166) of the particles of work item `it` (tile it.x,
// scratch range [it.y, it.z)) into an LDS block of the tile plus halo, then into rho; the
// block's share of the rho moments goes to mrow[5] (if given).
// PIPE: as for sc2d_hist_body.
template <bool PIPE>
__device__ __forceinline__ void sc2d_tdeposit_body(const SC2DParams& P, const Grid2Step& g,
                                                   const Tile2D& tg, int4 it, bool per_particle_m,
                                                   double* rho_ptr, double* acc,
                                                   double* mrow) {
  const gptr<double> rho = to_global(rho_ptr);
  const int W = tg.T + 2, cells = W * W;
  const int x0 = (it.x / tg.ty) * tg.T - 1, y0 = (it.x % tg.ty) * tg.T - 1;
  const gptr<const dpair> sxy = to_global((const dpair*)tg.sxy);
  const gptr<const double> sm = to_global((const double*)tg.sm);
  const int stride = blockDim.x;
  // kUnroll particles per thread per pass, loaded together; the first pass's loads go out
  // before the LDS block is cleared, each next pass's before this one is deposited
  auto load = [&](int i0, double (&xs)[kUnroll], double (&ys)[kUnroll], double (&ms)[kUnroll]) {
#pragma unroll
    for (int u = 0; u < kUnroll; u++) {
      const int i = i0 + u * stride;
      const dpair v = sxy[i < it.z ? i : it.y];
      xs[u] = v.x;
      ys[u] = v.y;
      ms[u] = per_particle_m ? (i < it.z ? sm[i] : 0.0) : P.macrosize;
    }
  };
  double nxs[kUnroll], nys[kUnroll], nms[kUnroll];
  if (PIPE && it.y < it.z) load(it.y + (int)threadIdx.x, nxs, nys, nms);
  for (int c = threadIdx.x; c < cells; c += blockDim.x) acc[c] = 0.0;
  __syncthreads();
  const double xMin = g.xMin, yMin = g.yMin, dx = g.dx, dy = g.dy;
  const int nx = P.nx, ny = P.ny;
  double mom[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int i0 = it.y + threadIdx.x; i0 < it.z; i0 += kUnroll * stride) {
    double xs[kUnroll], ys[kUnroll], ms[kUnroll];
    if (PIPE) {
#pragma unroll
      for (int u = 0; u < kUnroll; u++) {
        xs[u] = nxs[u];
        ys[u] = nys[u];
        ms[u] = nms[u];
      }
      if (i0 + kUnroll * stride < it.z) load(i0 + kUnroll * stride, nxs, nys, nms);
    } else {
      load(i0, xs, ys, ms);
    }
#pragma unroll
    for (int u = 0; u < kUnroll; u++) {
      if (i0 + u * stride >= it.z) break;
      const double x = xs[u], y = ys[u], m = ms[u];
      int iX, iY;
      double xF, yF;
      node_ind_frac(x, xMin, dx, nx, iX, xF);
      node_ind_frac(y, yMin, dy, ny, iY, yF);
      const double xF2 = xF * xF, yF2 = yF * yF;
      const double wx[3] = {0.5 * (0.25 - xF + xF2), 0.75 - xF2, 0.5 * (0.25 + xF + xF2)};
      const double wy[3] = {0.5 * (0.25 - yF + yF2), 0.75 - yF2, 0.5 * (0.25 + yF + yF2)};
      const int lx = iX - 1 - x0, ly = iY - 1 - y0;
      if (lx >= 0 && lx + 2 < W && ly >= 0 && ly + 2 < W) {
        double* base = acc + lx * W + ly;
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
          for (int c = 0; c < 3; c++) atomicAdd(base + a * W + c, wx[a] * wy[c] * m);
      } else {  // not in this tile's block (cannot happen for a consistent sort): global
        const gptr<double> base = rho + (int64_t)(iX - 1) * ny + (iY - 1);
        for (int a = 0; a < 3; a++)
          for (int c = 0; c < 3; c++) {
            const double val = wx[a] * wy[c] * m;
            global_add(base + a * ny + c, val);
            moment_add(mom, val, iX - 1 + a, iY - 1 + c, xMin, dx);
          }
      }
    }
  }
  __syncthreads();
  for (int c = threadIdx.x; c < cells; c += blockDim.x) {
    const double val = acc[c];
    if (val == 0.0) continue;
    const int gx = x0 + c / W, gy = y0 + c % W;
    if (gx < 0 || gx >= nx || gy < 0 || gy >= ny) continue;
    global_add(rho + (int64_t)gx * ny + gy, val);
    moment_add(mom, val, gx, gy, xMin, dx);
  }
  if (mrow) {
    __shared__ double red[kDepositThreads / 64][5];
    for (int k = 0; k < 5; k++)
      for (int o = 32; o > 0; o >>= 1) mom[k] += __shfl_xor(mom[k], o, 64);
    if ((threadIdx.x & 63) == 0)
      for (int k = 0; k < 5; k++) red[threadIdx.x >> 6][k] = mom[k];
    __syncthreads();
    if (threadIdx.x < 5) {
      double s = 0.0;
      for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += red[w][threadIdx.x];
      mrow[threadIdx.x] = s;
    }
  }
}

// Single bunch: the grid from the extrema partials in every block (block 0 writes the
// state and clears the scatter cursors); blocks clear rho in turn (it is next written by
// the deposit, two launches later) and write their z line density rows.
template <bool COUNT>
__global__ __launch_bounds__(kDepositThreads) void sc2d_bin(SC2DParams P, DeviceBunchView v,
                                                            const double* __restrict__ partials,
                                                            SC2DState* st, SC2DState* gate,
                                                            Tile2D tg, double* __restrict__ rho,
                                                            int* status) {
  extern __shared__ double dyn[];
  // the first pass's particle loads go out before the grid is computed from the extrema
  const int64_t first = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  HistPass pre;
  hist_load(v, P.macrosize, first, v.n, stride, pre);
  Grid2Step g;
  sc2d_grid_body(P, partials, 0, P.nblk, st, g);
  if (blockIdx.x == 0) {
    if (threadIdx.x == 0) {
      sc2d_write_state(st, g);
      if (P.plain) sc2d_write_solver(P, st, gate, g);
      if (g.skip) *status = 1;  // zero extent: the reference aborts (:176-189)
    }
    for (int t = threadIdx.x; t < tg.ntiles; t += blockDim.x) tg.cursor[t] = 0u;
  }
  const int ng = P.nx * P.ny;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ng; i += gridDim.x * blockDim.x)
    rho[i] = 0.0;
  sc2d_hist_body<true, COUNT>(P, v, g, tg, nullptr, tg.zpart + (size_t)blockIdx.x * P.nz, first, v.n, stride,
                 dyn, (uint32_t*)(dyn + P.nz), nullptr, &pre);
}

__global__ __launch_bounds__(kDepositThreads) void sc2d_scatter(SC2DParams P, DeviceBunchView v,
                                                                const SC2DState* st, Tile2D tg,
                                                                int64_t per_block) {
  extern __shared__ uint32_t lds[];
  const int64_t lo = (int64_t)blockIdx.x * per_block;
  sc2d_scatter_body<true>(P, v, sc2d_load_grid(st), tg, lo, min(v.n, lo + per_block), lds,
                    lds + tg.ntiles, blockIdx.x == 0);
}

// Block 0 also clears the tile counts for the next step (every scatter block has read them).
__global__ __launch_bounds__(kDepositThreads) void sc2d_tdeposit(SC2DParams P,
                                                                 const SC2DState* st, Tile2D tg,
                                                                 int per_particle_m,
                                                                 double* __restrict__ rho,
                                                                 int moments) {
  if (blockIdx.x == 0)
    for (int t = threadIdx.x; t < tg.ntiles; t += blockDim.x) tg.counts[t] = 0u;
  if (blockIdx.x >= *tg.nitems) return;
  extern __shared__ double acc[];
  sc2d_tdeposit_body<true>(P, sc2d_load_grid(st), tg, tg.items[blockIdx.x], per_particle_m != 0, rho,
                     acc, moments ? tg.mpart + (size_t)blockIdx.x * 5 : nullptr);
}

// Single bunch, grids that fit in LDS (nx*ny <= kPrivNodes): deposit without the tile
// sort.  Workgroup b takes particles [b*per, min(n, (b+1)*per)) in particle order and
// deposits them (Grid2D::binValue, Grid2D.cc:142-166, the same weights as the tiled
// deposit) into its own copy of the whole grid in LDS, then adds the non-zero nodes to
// rho with global atomics and its share of the rho moments to mpart row b.  Block 0 sets
// the row count the post pass reads and keeps the tile counts zero for the tiled path.
__global__ __launch_bounds__(kPrivThreads) void sc2d_pdeposit(SC2DParams P, DeviceBunchView v,
                                                                 const SC2DState* st, Tile2D tg,
                                                                 int64_t per,
                                                                 double* __restrict__ rho_ptr,
                                                                 int moments) {
  extern __shared__ double grid[];
  const Grid2Step g = sc2d_load_grid(st);
  if (blockIdx.x == 0) {
    for (int t = threadIdx.x; t < tg.ntiles; t += blockDim.x) tg.counts[t] = 0u;
    if (threadIdx.x == 0) *tg.nitems = g.skip ? 0u : gridDim.x;
  }
  if (g.skip) return;
  const gptr<const double> px = to_global((const double*)v.c[0]);
  const gptr<const double> py = to_global((const double*)v.c[2]);
  const gptr<const double> pm = to_global((const double*)v.msize);
  const gptr<double> rho = to_global(rho_ptr);
  const int nx = P.nx, ny = P.ny, nodes = nx * ny;
  const int64_t lo = (int64_t)blockIdx.x * per, hi = min(v.n, lo + per);
  const int64_t stride = blockDim.x;
  // kUnroll particles per thread per pass, loaded together; the first pass's loads go out
  // before the LDS grid is cleared
  double x[kUnroll], y[kUnroll], m[kUnroll];
  auto load = [&](int64_t i0) {
#pragma unroll
    for (int u = 0; u < kUnroll; u++) {
      const int64_t i = i0 + u * stride;
      x[u] = i < hi ? px[i] : 0.0;
      y[u] = i < hi ? py[i] : 0.0;
      m[u] = pm ? (i < hi ? pm[i] : 0.0) : P.macrosize;
    }
  };
  load(lo + threadIdx.x);
  for (int c = threadIdx.x; c < nodes; c += blockDim.x) grid[c] = 0.0;
  __syncthreads();
  const double xMin = g.xMin, yMin = g.yMin, dx = g.dx, dy = g.dy;
  for (int64_t i0 = lo + threadIdx.x; i0 < hi; i0 += kUnroll * stride) {
    if (i0 != lo + (int64_t)threadIdx.x) load(i0);
#pragma unroll
    for (int u = 0; u < kUnroll; u++) {
      if (i0 + u * stride >= hi) break;
      if (x[u] < g.xMin || x[u] > g.xMax || y[u] < g.yMin || y[u] > g.yMax) continue;
      int iX, iY;
      double xF, yF;
      node_ind_frac(x[u], xMin, dx, nx, iX, xF);
      node_ind_frac(y[u], yMin, dy, ny, iY, yF);
      const double xF2 = xF * xF, yF2 = yF * yF;
      const double wx[3] = {0.5 * (0.25 - xF + xF2), 0.75 - xF2, 0.5 * (0.25 + xF + xF2)};
      const double wy[3] = {0.5 * (0.25 - yF + yF2), 0.75 - yF2, 0.5 * (0.25 + yF + yF2)};
      double* base = grid + (iX - 1) * ny + (iY - 1);  // 1 <= iX <= nx-2, 1 <= iY <= ny-2
#pragma unroll
      for (int a = 0; a < 3; a++)
#pragma unroll
        for (int c = 0; c < 3; c++) atomicAdd(base + a * ny + c, wx[a] * wy[c] * m[u]);
    }
  }
  __syncthreads();
  double mom[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  // node c = gx*ny + gy, stepped by blockDim.x without a division per node
  const int step_x = (int)blockDim.x / ny, step_y = (int)blockDim.x - step_x * ny;
  int gx = (int)threadIdx.x / ny, gy = (int)threadIdx.x - gx * ny;
  for (int c = threadIdx.x; c < nodes; c += blockDim.x) {
    const double val = grid[c];
    if (val != 0.0) {
      global_add(rho + c, val);
      moment_add(mom, val, gx, gy, xMin, dx);
    }
    gx += step_x;
    gy += step_y;
    if (gy >= ny) {
      gy -= ny;
      gx++;
    }
  }
  if (moments) {
    __shared__ double red[kPrivThreads / 64][5];
    for (int k = 0; k < 5; k++)
      for (int o = 32; o > 0; o >>= 1) mom[k] += __shfl_xor(mom[k], o, 64);
    if ((threadIdx.x & 63) == 0)
      for (int k = 0; k < 5; k++) red[threadIdx.x >> 6][k] = mom[k];
    __syncthreads();
    if (threadIdx.x < 5) {
      double sum = 0.0;
      for (int w = 0; w < (int)(blockDim.x >> 6); w++) sum += red[w][threadIdx.x];
      tg.mpart[(size_t)blockIdx.x * 5 + threadIdx.x] = sum;
    }
  }
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

// zg: when zpart is given, the sum of its zrows rows of per-block line densities.
// Moments: the sum of the deposit items' partial moments mpart[nitems][5], or (mpart
// nullptr, when rho was changed after the deposit) a pass over rho.
__device__ __forceinline__ void sc2d_post_body(const SC2DParams& P, SC2DState* st,
                                               const double* __restrict__ rho,
                                               double* __restrict__ zg,
                                               double* __restrict__ zd,
                                               const double* __restrict__ zpart, int zrows,
                                               const double* __restrict__ mpart,
                                               const uint32_t* nitems, int part = 3) {
  __shared__ double sh[kSetupThreads / 64];
  // the z line density, read below many times per bin, is kept in LDS
  __shared__ double zl[kMaxZBinsLDS];
  const int nz = P.nz;
  // part 1: z line density and its derivative; part 2: rho moments and kick factors (two
  // workgroups may run them side by side: neither reads what the other writes); 3: both
  if (!(part & 1)) {
  } else if (zpart) {
    // g = blockDim / nz threads per bin (when nz < blockDim) sum the rows in turn, then
    // their partial sums meet in LDS
    __shared__ double zs[kSetupThreads];
    const int g = nz < (int)blockDim.x ? (int)blockDim.x / nz : 1;
    for (int i0 = 0; i0 < nz; i0 += (int)blockDim.x / g) {
      const int i = i0 + (int)threadIdx.x / g, r = (int)threadIdx.x % g;
      double acc = 0.0;
      if (i < nz && (int)threadIdx.x < g * ((int)blockDim.x / g)) {
        // eight rows' loads in flight per round, added in the same order as one by one
        int b = r;
        for (; b + 7 * g < zrows; b += 8 * g) {
          double v[8];
#pragma unroll
          for (int u = 0; u < 8; u++) v[u] = zpart[(size_t)(b + u * g) * nz + i];
#pragma unroll
          for (int u = 0; u < 8; u++) acc += v[u];
        }
        for (; b < zrows; b += g) acc += zpart[(size_t)b * nz + i];
      }
      zs[threadIdx.x] = acc;
      __syncthreads();
      if (r == 0 && i < nz && (int)threadIdx.x < g * ((int)blockDim.x / g)) {
        double t = 0.0;
        for (int k = 0; k < g; k++) t += zs[threadIdx.x + k];
        zg[i] = t;
        zl[i] = t;
      }
      __syncthreads();
    }
  } else {
    for (int i = threadIdx.x; i < nz; i += blockDim.x) zl[i] = zg[i];
    __syncthreads();
  }
  if (st->skip) {
    if (part & 1)
      for (int i = threadIdx.x; i < nz; i += blockDim.x) zd[i] = 0.0;
    return;
  }
  const double z_step = st->dz;
  const double zMin = st->zMin;
  if (P.plain) {  // (part 3 only: the total comes from the line density)
    // SpaceChargeCalc2p5D.cc:229-233, 100: the total macrosize is the sum of the z line
    // density; no derivative and no rho moments are used
    double t = 0.0;
    for (int i = threadIdx.x; i < nz; i += blockDim.x) t += zl[i];
    t = block_sum(t, sh);
    if (threadIdx.x == 0) {
      st->total = t;
      st->factor = P.factor0 / (z_step * t);
      st->scale_coeff = log(st->sdx / st->dx);  // PoissonSolverFFT2D.cc:216
    }
    return;
  }
  // calculateLongDerivative (:257-310)
  if (!(part & 1)) {
  } else if (nz == 1) {
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
        zDeriv = (zl[ip] - zl[i0]) / z_step;
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
            const double y = zl[i];
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
  if (!(part & 2)) return;
  // rho moments (:229-254); quirk Q1: the y position uses getGridX(iy) (:238)
  const int nx = P.nx, ny = P.ny;
  const double gx0 = st->xMin, gdx = st->dx;
  double s0 = 0., s1 = 0., s2 = 0., s3 = 0., s4 = 0.;
  if (mpart) {
    const int ni = (int)*nitems;
    for (int j = threadIdx.x; j < ni; j += blockDim.x) {
      const double* m = mpart + (size_t)j * 5;
      s0 += m[0];
      s1 += m[1];
      s2 += m[2];
      s3 += m[3];
      s4 += m[4];
    }
  } else {
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
                                                           double* __restrict__ zg,
                                                           double* __restrict__ zd, Tile2D tg,
                                                           int zrows, int moments) {
  sc2d_post_body(P, st, rho, zg, zd, tg.zpart, zrows, moments ? tg.mpart : nullptr, tg.nitems);
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
  constexpr int LINE = fft_line(N);  // evaluated at compile time
  double* im = sm + LINE;
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

// The single-bunch step's row transforms of rho and its post pass in one launch:
// workgroups 0 .. rows/2 - 1 transform row pairs, the last two run the post pass's two
// independent halves side by side (z line density and derivative; rho moments and kick
// factors).  All only read rho; the post pass writes the state's kick factors, z line
// density and derivative, which the column pass after this launch (scale_coeff) and the
// next map kernel read.  One launch fewer per step.
template <int N>
__global__ __launch_bounds__(kFFTThreads) void fft_rows_r2c_post(
    const double* __restrict__ rho, int nrows_in, int ncols_in, double2* __restrict__ out,
    const double2* __restrict__ tw, int nmax, SC2DParams P, SC2DState* st,
    double* __restrict__ zg, double* __restrict__ zd, Tile2D tg, int zrows, int moments) {
  const int rb = (nrows_in + 1) / 2;
  if ((int)blockIdx.x < rb) {
    rows_r2c_body<N>(rho, nrows_in, ncols_in, ncols_in, out, tw, nmax, st, blockIdx.x);
    return;
  }
  // two more workgroups: the line density part and the moments part of the post pass
  sc2d_post_body(P, st, rho, zg, zd, tg.zpart, zrows, moments ? tg.mpart : nullptr, tg.nitems,
                 (int)blockIdx.x == rb ? 1 : 2);
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
  constexpr int LINE = fft_line(N);  // evaluated at compile time
  double* im = sm + LINE;
  // The column and (for a solve) its Green values are requested together before the
  // transform; a masked lane loads element 0 and drops it.
  constexpr int CI = (N + kFFTThreads - 1) / kFFTThreads;
  double2 cv[CI], gv[CI];
#pragma unroll
  for (int u = 0; u < CI; u++) {
    const int i = (int)threadIdx.x + u * kFFTThreads;
    cv[u] = spec[(i < nrows_in && i < N) ? (int64_t)i * H + k : (int64_t)k];
    if constexpr (!GREEN) gv[u] = ghat_t[(int64_t)k * N + (i < N ? i : 0)];
  }
#pragma unroll
  for (int u = 0; u < CI; u++) {
    const int i = (int)threadIdx.x + u * kFFTThreads;
    if (i < N) {
      re[i] = i < nrows_in ? cv[u].x : 0.0;
      im[i] = i < nrows_in ? cv[u].y : 0.0;
    }
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
#pragma unroll
    for (int u = 0; u < CI; u++) {
      const int i = (int)threadIdx.x + u * kFFTThreads;
      if (i >= N) continue;
      const double2 g = gv[u];
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
  constexpr int LINE = fft_line(N);  // evaluated at compile time
  double* im = sm + LINE;
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
  P.plain = 0;
  P.par = 0;
  P.bnd_type = 0;
  for (int k = 0; k < 6; k++) P.bnd[k] = 0.0;
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
  sc2d_setup_body(P, partials, B.item_begin, B.item_end, B.st, B.tg, B.rho, B.zg, B.zd);
}

// Tile-sorted deposit of every bunch: bin and scatter take one block per particle work
// item (every scatter block scans its bunch's counts; the bunch's first item writes the
// work items).  The bin block keeps its item's tile counts in a row of tg.icounts, so the
// scatter block of the same item reserves its destinations from that row and reads the
// particles once instead of twice.
__global__ __launch_bounds__(kDepositThreads) void sc2d_hist_batch(const BatchBunch* bb,
                                                                   const BatchItem* items) {
  extern __shared__ double dyn[];
  const BatchItem it = items[blockIdx.x];
  const BatchBunch& B = bb[it.bunch];
  const SC2DParams P = batch_params(B, 0.0, 1.0);
  sc2d_hist_body<false>(P, B.v, sc2d_load_grid(B.st), B.tg, B.zg, nullptr, it.begin + threadIdx.x,
                 it.end, blockDim.x, dyn, (uint32_t*)(dyn + P.nz),
                 B.tg.icounts + (size_t)(blockIdx.x - B.item_begin) * B.tg.ntiles);
}

__global__ __launch_bounds__(kDepositThreads) void sc2d_scatter_batch(const BatchBunch* bb,
                                                                      const BatchItem* items) {
  extern __shared__ uint32_t lds[];
  const BatchItem it = items[blockIdx.x];
  const BatchBunch& B = bb[it.bunch];
  const SC2DParams P = batch_params(B, 0.0, 1.0);
  sc2d_scatter_body<false>(P, B.v, sc2d_load_grid(B.st), B.tg, it.begin, it.end, lds,
                    lds + B.tg.ntiles, (int)blockIdx.x == B.item_begin,
                    B.tg.icounts + (size_t)(blockIdx.x - B.item_begin) * B.tg.ntiles);
}

// Bunch whose deposit block range [slot_begin, slot_end) holds block blk.
__device__ __forceinline__ int batch_find_slot(const BatchBunch* bb, int nb, int blk) {
  int lo = 0, hi = nb - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (bb[mid].slot_begin <= blk) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__global__ __launch_bounds__(kDepositThreads) void sc2d_tdeposit_batch(const BatchBunch* bb,
                                                                       int nb) {
  extern __shared__ double acc[];
  const int b = batch_find_slot(bb, nb, blockIdx.x);
  const BatchBunch& B = bb[b];
  const int j = blockIdx.x - B.slot_begin;
  if (j == 0)  // every scatter block of the bunch has read the counts: clear them
    for (int t = threadIdx.x; t < B.tg.ntiles; t += blockDim.x) B.tg.counts[t] = 0u;
  if (j >= (int)*B.tg.nitems) return;
  const SC2DParams P = batch_params(B, 0.0, 1.0);
  sc2d_tdeposit_body<false>(P, sc2d_load_grid(B.st), B.tg, B.tg.items[j], B.v.msize != nullptr, B.rho,
                     acc, B.moments_from_rho ? nullptr : B.tg.mpart + (size_t)j * 5);
}

__global__ __launch_bounds__(kSetupThreads) void sc2d_post_batch(const BatchBunch* bb,
                                                                 double length,
                                                                 double pipe_radius) {
  const BatchBunch& B = bb[blockIdx.x];
  const SC2DParams P = batch_params(B, length, pipe_radius);
  sc2d_post_body(P, B.st, B.rho, B.zg, B.zd, nullptr, 0,
                 B.moments_from_rho ? nullptr : B.tg.mpart, B.tg.nitems);
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
__device__ __forceinline__ void green2d_fill_body(double* __restrict__ G, int M0, int M1,
                                                  double dx, double dy) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)M0 * M1) return;
  const int i = (int)(t / M1), j = (int)(t - (int64_t)i * M1);
  const int ii = (i <= M0 / 2) ? i : M0 - i;
  const int jj = (j <= M1 / 2) ? j : M1 - j;
  const double rTransX = ii * dx, rTransY = jj * dy;
  const double rTot2 = rTransX * rTransX + rTransY * rTransY;
  G[t] = (ii != 0 || jj != 0) ? -log(rTot2) / 2 : 0.0;
}

__global__ void green2d_fill(double* __restrict__ G, int M0, int M1, double dx, double dy) {
  green2d_fill_body(G, M0, M1, dx, dy);
}

// SpaceChargeCalc2p5D: the same on the solver steps of the state, when this step's grid
// pass asked for a rebuild (gate->skip 0).
__global__ void green2d_fill_state(double* __restrict__ G, int M0, int M1,
                                   const SC2DState* st, const SC2DState* gate) {
  if (gate->skip) return;
  green2d_fill_body(G, M0, M1, st->sdx, st->sdy);
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
  // tile-sorted deposit
  ob::Tile2D tg{};
  int64_t tile_ncap = 0;  // scratch particles allocated
  bool tile_msize = false;
  int tile_item = 2048;   // particles per deposit work item
  ob_denoiser* denoiser = nullptr;  // applied to rho after each deposit when set
  // SpaceChargeCalc2p5D: ratio rule and transverse-only kick; the Green function is
  // rebuilt on the device (gate->skip 0) into ghat_t from gbuf when the grid pass asks
  int plain = 0;
  int par = 0;                      // state entry the next step reads
  ob::SC2DState* gate = nullptr;
  double* gbuf = nullptr;           // M0 x M1 Green function values
  ob::StatusBox status{nullptr, nullptr};  // set when a step finds a zero-extent bunch
};

// SpaceChargeCalc2p5D (SpaceChargeCalc2p5D.cc): the same device object in plain mode.
struct ob_sc2p5d : ob_sc2p5drb {};

namespace ob {
namespace {
template <int N>
size_t fft_lds() { return sizeof(double) * 2 * fft_line(N); }

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

// post: the post pass's arguments (single-bunch SpaceChargeCalc2p5Drb step), run with the
// row transforms in one launch; nullptr when it already ran.
struct PostArgs {
  SC2DParams P;
  Tile2D tg;
  int zrows, moments;
};

void launch_fft_pipeline(ob_sc2p5drb* c, SC2DState* st, const PostArgs* post = nullptr) {
  if (post) {
    dispatch_pow2(c->M1, [&](auto NC) {
      constexpr int N = decltype(NC)::value;
      if constexpr (N >= 8) {
        static bool ready = false;
        if (!ready) {
          allow_lds(fft_rows_r2c_post<N>, fft_lds<N>());
          ready = true;
        }
        ob::ProfScope ps_("fft_rows_r2c_post");
        OB_REQUIRE(!post->P.plain, "2.5D post pass: the split form is for SpaceChargeCalc2p5Drb");
        hipLaunchKernelGGL(fft_rows_r2c_post<N>, dim3((c->nx + 1) / 2 + 2), dim3(kFFTThreads),
                           fft_lds<N>(), stream(), c->rho, c->nx, c->ny, c->spec, c->tw, c->nmax,
                           post->P, st, c->zg, c->zd, post->tg, post->zrows, post->moments);
      }
    });
    OB_HIP(hipGetLastError());
  } else {
    launch_rows_r2c(c, c->rho, c->nx, c->ny, c->ny, st);
  }
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
  OB_REQUIRE(!c->plain, "batches track SpaceChargeCalc2p5Drb calculators only");
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

Tile2D sc2d_tile_reserve(ob_sc2p5drb* c, int64_t n, bool per_particle_m, int item,
                         int zrows, int hist_rows) {
  Tile2D& g = c->tg;
  if (g.T == 0) {
    // 32 x 32-node tiles (a 34 x 34 LDS block): measured faster than 16 (fewer destination
    // segments for the scatter, fewer same-node collisions of a wave's deposit atomics) and
    // than 64 (batch64 1.62e10 against 1.59e10 and 1.57e10)
    int T = 32;
    while ((int64_t)((c->nx + T - 1) / T) * ((c->ny + T - 1) / T) > 4096) T *= 2;
    g.T = T;
    g.shift = __builtin_ctz((unsigned)T);
    g.tx = (c->nx + T - 1) / T;
    g.ty = (c->ny + T - 1) / T;
    g.ntiles = g.tx * g.ty;
    OB_HIP(hipMalloc((void**)&g.counts, sizeof(uint32_t) * (2 * g.ntiles + 1)));
    g.cursor = g.counts + g.ntiles;
    g.nitems = g.cursor + g.ntiles;
    OB_HIP(hipMemset(g.counts, 0, sizeof(uint32_t) * (2 * g.ntiles + 1)));
  }
  const int64_t want_items = n / item + g.ntiles + 1;
  OB_REQUIRE(want_items < (int64_t)1 << 31, "tile sort: too many work items");
  if (n > c->tile_ncap || (per_particle_m && !c->tile_msize) || item != g.item ||
      want_items > g.max_items) {
    OB_HIP(hipStreamSynchronize(stream()));  // buffers may be in use by queued work
    for (void* p : {(void*)g.sxy, (void*)g.sm, (void*)g.items, (void*)g.mpart})
      if (p) OB_HIP(hipFree(p));
    g.sxy = nullptr;
    g.sm = g.mpart = nullptr;
    int64_t cap = std::max<int64_t>(n, std::max<int64_t>(c->tile_ncap, 1024));
    OB_HIP(hipMalloc((void**)&g.sxy, sizeof(double2) * cap));
    c->tile_msize = per_particle_m || c->tile_msize;
    if (c->tile_msize) OB_HIP(hipMalloc((void**)&g.sm, sizeof(double) * cap));
    c->tile_ncap = cap;
    g.item = item;
    g.max_items = (int)(cap / item + g.ntiles + 1);
    OB_HIP(hipMalloc((void**)&g.items, sizeof(int4) * g.max_items));
    OB_HIP(hipMalloc((void**)&g.mpart, sizeof(double) * 5 * g.max_items));
  }
  if (zrows > g.zrows) {
    OB_HIP(hipStreamSynchronize(stream()));
    if (g.zpart) OB_HIP(hipFree(g.zpart));
    OB_HIP(hipMalloc((void**)&g.zpart, sizeof(double) * zrows * c->nz));
    g.zrows = zrows;
  }
  if (hist_rows > g.hist_rows) {
    OB_HIP(hipStreamSynchronize(stream()));
    if (g.icounts) OB_HIP(hipFree(g.icounts));
    OB_HIP(hipMalloc((void**)&g.icounts, sizeof(uint32_t) * (size_t)hist_rows * g.ntiles));
    g.hist_rows = hist_rows;
  }
  return g;
}

bool sc2d_tile_same(ob_sc2p5drb* c, const Tile2D& t, int64_t n, bool per_particle_m) {
  const Tile2D& g = c->tg;
  return g.sxy == t.sxy && g.items == t.items && g.item == t.item && n <= c->tile_ncap &&
         g.icounts == t.icounts && g.hist_rows == t.hist_rows &&
         (!per_particle_m || g.sm != nullptr);
}

void sc2d_mark_tracked(ob_sc2p5drb* c) { c->tracked = true; }

int sc2d_read_status(ob_sc2p5drb* c) {
  int v = 0;
  OB_HIP(hipMemcpy(&v, &c->st->status, sizeof(int), hipMemcpyDeviceToHost));
  return v;
}

bool sc2d_denoise_job(ob_sc2p5drb* c, DenoiseJob* job) {
  if (!c->denoiser) return false;
  *job = DenoiseJob{c->denoiser, c->rho, c->nx, c->ny, &c->st->skip};
  return true;
}

void launch_sc2d_batch(const BatchBunch* bb, const BatchItem* items, const BatchShape& sh,
                       const double* partials, double length, double pipe_radius,
                       hipStream_t s, const std::vector<DenoiseJob>& jobs) {
  OB_REQUIRE(sh.hist_lds <= 64 * 1024 && sh.dep_lds <= 64 * 1024,
             "batch deposit: LDS block too large");
  OB_REQUIRE(sh.max_M <= 1024, "batch: grids up to 512 points per axis");
  {
    ProfScope ps("sc2d_setup_batch");
    hipLaunchKernelGGL(sc2d_setup_batch, dim3(sh.nb), dim3(kSetupThreads), 0, s, bb, partials,
                       length, pipe_radius);
  }
  {
    ProfScope ps("sc2d_hist_batch");
    hipLaunchKernelGGL(sc2d_hist_batch, dim3(sh.nitems), dim3(kDepositThreads), sh.hist_lds, s,
                       bb, items);
  }
  {
    ProfScope ps("sc2d_scatter_batch");
    hipLaunchKernelGGL(sc2d_scatter_batch, dim3(sh.nitems), dim3(kDepositThreads), sh.hist_lds,
                       s, bb, items);
  }
  {
    ProfScope ps("sc2d_tdeposit_batch");
    hipLaunchKernelGGL(sc2d_tdeposit_batch, dim3(sh.dep_slots), dim3(kDepositThreads),
                       sh.dep_lds, s, bb, sh.nb);
  }
  if (!jobs.empty()) dn_denoise_rhos(jobs);
  {
    ProfScope ps("sc2d_post_batch");
    hipLaunchKernelGGL(sc2d_post_batch, dim3(sh.nb), dim3(kSetupThreads), 0, s, bb, length,
                       pipe_radius);
  }
  const size_t fl = sizeof(double) * 2 * fft_line(sh.max_M);  // fft_line grows with N
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

namespace ob {
namespace {

// Calculator object of either kind: grids, FFT buffers, state, the solver Green function
// on [-xy_ratio, xy_ratio] x [-1, 1] (SpaceChargeCalc2p5Drb.cc:31-46,
// SpaceChargeCalc2p5D.cc:25-43, PoissonSolver2D.cc:31-32).
template <typename Calc>
Calc* sc2d_create(int nx, int ny, int nz, double xy_ratio, bool plain, const std::string& who) {
  OB_REQUIRE(nx >= 3 && ny >= 3, who + ": grid size too small (should be >= 3)");
  OB_REQUIRE(nz >= 1, who + ": number of z bins should be >= 1");
  OB_REQUIRE(nz <= kMaxZBinsLDS, who + ": more than 4096 z bins");
  OB_REQUIRE(xy_ratio > 0.0, who + ": xy_ratio must be positive");
  OB_REQUIRE(pad_len(nx) <= 4096 && pad_len(ny) <= 4096,
             who + ": grid larger than 2048 points per axis");
  ensure_init();
  Calc* c = new Calc();
  c->plain = plain ? 1 : 0;
  c->nx = nx;
  c->ny = ny;
  c->nz = nz;
  c->M0 = pad_len(nx);
  c->M1 = pad_len(ny);
  c->H = c->M1 / 2 + 1;
  c->nmax = c->M0 > c->M1 ? c->M0 : c->M1;
  c->xy_ratio = xy_ratio;
  c->dx_solver = (xy_ratio - -xy_ratio) / (nx - 1);
  c->dy_solver = (1.0 - -1.0) / (ny - 1);
  OB_HIP(hipMalloc((void**)&c->rho, sizeof(double) * nx * ny));
  OB_HIP(hipMalloc((void**)&c->phi, sizeof(double) * nx * ny));
  OB_HIP(hipMalloc((void**)&c->zg, sizeof(double) * nz));
  OB_HIP(hipMalloc((void**)&c->zd, sizeof(double) * nz));
  OB_HIP(hipMalloc((void**)&c->spec, sizeof(double2) * (size_t)c->M0 * c->H));
  OB_HIP(hipMalloc((void**)&c->ghat_t, sizeof(double2) * (size_t)c->M0 * c->H));
  OB_HIP(hipMalloc((void**)&c->partials, sizeof(double) * 6 * 4096));
  OB_HIP(hipMalloc((void**)&c->st, sizeof(SC2DState)));
  OB_HIP(hipMemset(c->rho, 0, sizeof(double) * nx * ny));
  OB_HIP(hipMemset(c->phi, 0, sizeof(double) * nx * ny));
  OB_HIP(hipMemset(c->zg, 0, sizeof(double) * nz));
  OB_HIP(hipMemset(c->zd, 0, sizeof(double) * nz));
  SC2DState init;
  std::memset(&init, 0, sizeof(init));
  init.xMin = -1.0; init.xMax = 1.0; init.yMin = -1.0; init.yMax = 1.0;  // Grid2D(nx, ny)
  init.dx = 2.0 / (nx - 1); init.dy = 2.0 / (ny - 1);
  init.zMin = -0.5; init.zMax = 0.5; init.dz = 1.0 / nz;                // Grid1D(nz)
  for (int k = 0; k < 2; k++) {
    init.ratio[k] = xy_ratio;
    init.sol[k][0] = -xy_ratio; init.sol[k][1] = xy_ratio;
    init.sol[k][2] = -1.0; init.sol[k][3] = 1.0;
  }
  init.sdx = c->dx_solver;
  init.sdy = c->dy_solver;
  init.long_kick = plain ? 0 : 1;
  OB_HIP(hipMemcpy(c->st, &init, sizeof(init), hipMemcpyHostToDevice));
  if (plain) {
    OB_HIP(hipMalloc((void**)&c->gate, sizeof(SC2DState)));
    SC2DState g;
    std::memset(&g, 0, sizeof(g));
    g.skip = 1;
    OB_HIP(hipMemcpy(c->gate, &g, sizeof(g), hipMemcpyHostToDevice));
    OB_HIP(hipMalloc((void**)&c->gbuf, sizeof(double) * c->M0 * c->M1));
  }
  c->tw = make_twiddles(c->nmax);
  define_green(c);
  c->status = status_box_create(
      plain ? "SpaceChargeCalc2p5D::bunchAnalysis: the bunch min and max sizes are wrong "
              "(zero extent); cannot calculate space charge"
            : "SpaceChargeCalc2p5Drb::bunchAnalysis: the bunch min and max sizes are wrong "
              "(zero extent); cannot calculate space charge");
  return c;
}

void sc2d_destroy(ob_sc2p5drb* c) {
  // Pending kicks read st/phi/zg/zd: apply them before the buffers go.
  std::string err;
  try {
    batch_forget_calc(c);
    flush_kicks_reading(c->st);
  } catch (const Error& e) {
    err = e.msg;
  }
  (void)hipStreamSynchronize(stream());
  if (c->status.host) status_box_destroy(c->status);
  for (void* p : {(void*)c->rho, (void*)c->phi, (void*)c->zg, (void*)c->zd, (void*)c->spec,
                  (void*)c->ghat_t, (void*)c->tw, (void*)c->partials, (void*)c->st,
                  (void*)c->gate, (void*)c->gbuf, (void*)c->tg.counts, (void*)c->tg.items,
                  (void*)c->tg.sxy, (void*)c->tg.sm, (void*)c->tg.zpart,
                  (void*)c->tg.mpart, (void*)c->tg.icounts})
    if (p) (void)hipFree(p);
  delete c;
  OB_REQUIRE(err.empty(), err);
}

// One trackBunch: SpaceChargeCalc2p5Drb.cc:110-168, or SpaceChargeCalc2p5D.cc:77-120
// (with the boundary bnd, if not null) when c->plain.
void sc2d_track(ob_sc2p5drb* c, ob_bunch* b, double length, double pipe_radius,
                ob_boundary2d* bnd = nullptr) {
  check_statuses(false);
  const int64_t n = bunch_size(b);
  if (n < 2) return;  // :112-113 (2p5D :79-80)
  hipStream_t s = stream();
  const int nblk = bunch_flush(b, c->partials);
  DeviceBunchView v = bunch_device_view(b);
  const SyncInfo sy = bunch_sync(b);
  SC2DParams P;
  P.nx = c->nx; P.ny = c->ny; P.nz = c->nz; P.nblk = nblk;
  P.xy_ratio = c->xy_ratio;
  P.dx_solver = c->dx_solver;
  P.macrosize = sy.macrosize;
  P.factor0 = 2 * length * sy.r0 / ((sy.beta * sy.beta) * pow(sy.gamma, 3));  // :129
  P.pipe_radius = pipe_radius;
  P.n_long_avg = c->n_long_avg;
  P.plain = c->plain;
  P.par = c->par;
  for (int k = 0; k < 6; k++) P.bnd[k] = 0.0;
  P.bnd_type = bnd ? boundary_describe(bnd, P.bnd) : 0;
  // bin (grid + tile counts + z line density), scatter (scan + counting sort), deposit
  // the whole grid fits in LDS: private-grid deposit in particle order, no tile sort
  const bool private_grid = (int64_t)c->nx * c->ny <= kPrivNodes && n <= kPrivMaxParticles;
  // bin workgroups of 2048 particles; 1024 with the private-grid deposit, whose bin pass
  // counts no tiles (100k on 128x128: 0.8% faster per step than 2048, 512 4% slower)
  const int64_t bin_per = private_grid ? 1024 : 2048;
  const int hb = (int)std::max<int64_t>(
      1, std::min<int64_t>({(n + bin_per - 1) / bin_per, (int64_t)4 * device_cu_count(),
                            (int64_t)std::max(1, (1 << 18) / c->nz)}));
  const Tile2D tg = sc2d_tile_reserve(c, n, v.msize != nullptr, c->tile_item, hb);
  {
    ProfScope ps_("sc2d_bin");
    if (private_grid)
      hipLaunchKernelGGL(sc2d_bin<false>, dim3(hb), dim3(kDepositThreads),
                         sizeof(double) * c->nz + sizeof(uint32_t) * tg.ntiles, s, P, v,
                         c->partials, c->st, c->gate, tg, c->rho, c->status.dev);
    else
      hipLaunchKernelGGL(sc2d_bin<true>, dim3(hb), dim3(kDepositThreads),
                         sizeof(double) * c->nz + sizeof(uint32_t) * tg.ntiles, s, P, v,
                         c->partials, c->st, c->gate, tg, c->rho, c->status.dev);
  }
  if (c->plain) {
    // PoissonSolverFFT2D::setGridXY (PoissonSolverFFT2D.cc:103-111) when the grid pass
    // asked for it; each pass returns at once otherwise.  spec is free until the solve.
    const int64_t tot = (int64_t)c->M0 * c->M1;
    {
      ProfScope ps_("green2d_fill_state");
      hipLaunchKernelGGL(green2d_fill_state, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                         s, c->gbuf, c->M0, c->M1, c->st, c->gate);
    }
    launch_rows_r2c(c, c->gbuf, c->M0, c->M1, c->M1, c->gate);
    launch_cols<true>(c, c->M0, c->gate, 0.0);
    c->par ^= 1;
  }
  // rb: the rho moments come from the deposit unless the denoiser changes rho after it
  const int moments = (c->denoiser || c->plain) ? 0 : 1;
  if (private_grid) {
    static bool attr_done = false;
    if (!attr_done) {
      OB_HIP(hipFuncSetAttribute((const void*)sc2d_pdeposit,
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)(sizeof(double) * kPrivNodes)));
      attr_done = true;
    }
    const int64_t blocks = std::max<int64_t>(
        1, std::min<int64_t>({(n + kPrivPerBlock - 1) / kPrivPerBlock,
                              (int64_t)device_cu_count(), (int64_t)tg.max_items}));
    const int64_t per = (n + blocks - 1) / blocks;
    ProfScope ps_("sc2d_pdeposit");
    hipLaunchKernelGGL(sc2d_pdeposit, dim3((unsigned)blocks), dim3(kPrivThreads),
                       sizeof(double) * c->nx * c->ny, s, P, v, c->st, tg, per, c->rho,
                       moments);
  } else {
    {
      ProfScope ps_("sc2d_scatter");
      const int64_t per = std::max<int64_t>(2048, (n + 1023) / 1024);
      hipLaunchKernelGGL(sc2d_scatter, dim3((unsigned)((n + per - 1) / per)),
                         dim3(kDepositThreads), 2 * sizeof(uint32_t) * tg.ntiles, s, P, v,
                         c->st, tg, per);
    }
    {
      ProfScope ps_("sc2d_tdeposit");
      hipLaunchKernelGGL(sc2d_tdeposit, dim3((unsigned)(n / tg.item + tg.ntiles + 1)),
                         dim3(kDepositThreads), sizeof(double) * (tg.T + 2) * (tg.T + 2), s,
                         P, c->st, tg, v.msize != nullptr ? 1 : 0, c->rho, moments);
    }
  }
  if (c->denoiser) dn_denoise_rho(c->denoiser, c->rho, c->nx, c->ny, &c->st->skip);
  if (c->plain || bnd) {  // their solve reads the post pass's results: it runs first
    {
      ProfScope ps_("sc2d_post");
      hipLaunchKernelGGL(sc2d_post, dim3(1), dim3(kSetupThreads), 0, s, P, c->st, c->rho,
                         c->zg, c->zd, tg, hb, moments);
    }
    launch_fft_pipeline(c, c->st);
  } else {
    const PostArgs pa{P, tg, hb, moments};
    launch_fft_pipeline(c, c->st, &pa);
  }
  if (bnd) boundary_add_potential(bnd, c->phi, c->nx, c->ny, c->st, s);  // :90-94
  OB_HIP(hipGetLastError());
  // kick (:146-167; 2p5D :105-119), fused into the next map kernel
  Op op;
  std::memset(&op, 0, sizeof(op));
  op.kind = OP_SC2P5DRB_KICK;
  op.i0 = c->nx;
  op.i1 = c->ny;
  op.i2 = c->nz;
  op.p[0] = ptr_bits(c->st);
  op.p[1] = ptr_bits(c->phi);
  op.p[2] = ptr_bits(c->zg);
  op.p[3] = ptr_bits(c->zd);
  op.p[4] = c->plain ? 0.0 : 2 * log(pipe_radius);                            // :139
  op.p[5] = c->plain ? 0.0 : -length * sy.r0 * sy.mass / (sy.gamma * sy.gamma);  // :141
  bunch_queue(b, op);
  c->tracked = true;
}

void sc2d_get_grid(ob_sc2p5drb* c, int which, double* out, double* lims) {
  hipStream_t s = stream();
  SC2DState h;
  OB_HIP(hipMemcpyAsync(&h, c->st, sizeof(h), hipMemcpyDeviceToHost, s));
  const double* src = which == 0 ? c->rho : which == 1 ? c->phi : which == 2 ? c->zg : c->zd;
  const size_t cnt = which <= 1 ? (size_t)c->nx * c->ny : (size_t)c->nz;
  if (out) OB_HIP(hipMemcpyAsync(out, src, sizeof(double) * cnt, hipMemcpyDeviceToHost, s));
  OB_HIP(hipStreamSynchronize(s));
  check_statuses(true);
  if (lims) {
    if (which <= 1) {
      lims[0] = h.xMin; lims[1] = h.xMax; lims[2] = h.yMin; lims[3] = h.yMax;
    } else {
      lims[0] = h.zMin; lims[1] = h.zMax;
    }
  }
}
}  // namespace
}  // namespace ob

OB_API int ob_sc2p5drb_create(int nx, int ny, int nz, double xy_ratio, ob_sc2p5drb** out) {
  OB_TRY OB_REQUIRE(out != nullptr, "null output pointer");
  *out = ob::sc2d_create<ob_sc2p5drb>(nx, ny, nz, xy_ratio, false, "SpaceChargeCalc2p5Drb");
  OB_CATCH
}

OB_API int ob_sc2p5drb_destroy(ob_sc2p5drb* c) {
  OB_TRY if (!c) return 0;
  OB_REQUIRE(!c->plain, "not a SpaceChargeCalc2p5Drb handle");
  ob::sc2d_destroy(c);
  OB_CATCH
}

OB_API int ob_sc2p5drb_track(ob_sc2p5drb* c, ob_bunch* b, double length, double pipe_radius) {
  OB_TRY OB_REQUIRE(c != nullptr && b != nullptr, "null handle");
  OB_REQUIRE(!c->plain, "not a SpaceChargeCalc2p5Drb handle");
  ob::sc2d_track(c, b, length, pipe_radius);
  OB_CATCH
}

OB_API int ob_sc2p5d_create(int nx, int ny, int nz, double xy_ratio, ob_sc2p5d** out) {
  OB_TRY OB_REQUIRE(out != nullptr, "null output pointer");
  *out = ob::sc2d_create<ob_sc2p5d>(nx, ny, nz, xy_ratio, true, "SpaceChargeCalc2p5D");
  OB_CATCH
}

OB_API int ob_sc2p5d_destroy(ob_sc2p5d* c) {
  OB_TRY if (!c) return 0;
  OB_REQUIRE(c->plain, "not a SpaceChargeCalc2p5D handle");
  ob::sc2d_destroy(c);
  OB_CATCH
}

OB_API int ob_sc2p5d_track(ob_sc2p5d* c, ob_bunch* b, double length) {
  OB_TRY OB_REQUIRE(c != nullptr && b != nullptr, "null handle");
  OB_REQUIRE(c->plain, "not a SpaceChargeCalc2p5D handle");
  ob::sc2d_track(c, b, length, 0.0);
  OB_CATCH
}

OB_API int ob_sc2p5d_track_boundary(ob_sc2p5d* c, ob_bunch* b, double length,
                                   ob_boundary2d* bnd) {
  OB_TRY OB_REQUIRE(c != nullptr && b != nullptr, "null handle");
  OB_REQUIRE(c->plain, "not a SpaceChargeCalc2p5D handle");
  ob::sc2d_track(c, b, length, 0.0, bnd);
  OB_CATCH
}

OB_API int ob_sc2p5d_get_grid(ob_sc2p5d* c, int which, double* out, double* lims) {
  OB_TRY OB_REQUIRE(c != nullptr, "null handle");
  OB_REQUIRE(which >= 0 && which <= 2, "grid selector must be 0..2");
  ob::sc2d_get_grid(c, which, out, lims);
  OB_CATCH
}

// out: xy_ratio, solver box (xMin, xMax, yMin, yMax), totalMacrosize of the last step,
// number of Green function rebuilds.
OB_API int ob_sc2p5d_get_state(ob_sc2p5d* c, double* out) {
  OB_TRY OB_REQUIRE(c != nullptr && out != nullptr, "null handle");
  ob::SC2DState h;
  OB_HIP(hipMemcpyAsync(&h, c->st, sizeof(h), hipMemcpyDeviceToHost, ob::stream()));
  OB_HIP(hipStreamSynchronize(ob::stream()));
  ob::check_statuses(true);
  out[0] = h.ratio[c->par];
  for (int k = 0; k < 4; k++) out[1 + k] = h.sol[c->par][k];
  out[5] = h.total;
  out[6] = h.rebuilds;
  OB_CATCH
}

// PoissonSolverFFT2D::findPotential (PoissonSolverFFT2D.cc:174-256) on a host rho grid
// [nx][ny] spanning [xMin, xMax] x [yMin, yMax]: the calculator's solver (box
// [-xy_ratio, xy_ratio] x [-1, 1]) rescaled to the rho grid's step.
OB_API int ob_sc2p5drb_find_potential(ob_sc2p5drb* c, const double* rho, double xMin,
                                      double xMax, double yMin, double yMax, double* phi) {
  OB_TRY OB_REQUIRE(c != nullptr && rho != nullptr && phi != nullptr, "null handle");
  const double dx = (xMax - xMin) / (c->nx - 1), dy = (yMax - yMin) / (c->ny - 1);
  OB_REQUIRE(dx > 0.0 && dy > 0.0, "PoissonSolverFFT2D: empty grid");
  OB_REQUIRE(fabs(c->dx_solver / c->dy_solver - dx / dy) <= 1e-8,  // :176-210
             "PoissonSolverFFT2D: the grid sizes or shape are different");
  hipStream_t s = ob::stream();
  OB_HIP(hipStreamSynchronize(s));
  ob::SC2DState h;
  OB_HIP(hipMemcpy(&h, c->st, sizeof(h), hipMemcpyDeviceToHost));
  h.xMin = xMin;
  h.xMax = xMax;
  h.yMin = yMin;
  h.yMax = yMax;
  h.dx = dx;
  h.dy = dy;
  h.skip = 0;
  h.status = 0;
  h.scale_coeff = log(c->dx_solver / dx);  // :216
  OB_HIP(hipMemcpy(c->st, &h, sizeof(h), hipMemcpyHostToDevice));
  OB_HIP(hipMemcpyAsync(c->rho, rho, sizeof(double) * c->nx * c->ny, hipMemcpyHostToDevice, s));
  ob::launch_fft_pipeline(c, c->st);
  OB_HIP(hipMemcpyAsync(phi, c->phi, sizeof(double) * c->nx * c->ny, hipMemcpyDeviceToHost, s));
  OB_HIP(hipStreamSynchronize(s));
  OB_CATCH
}

// Denoise rho after every deposit (nullptr: off).  The denoiser must outlive its use here.
OB_API int ob_sc2p5drb_set_denoiser(ob_sc2p5drb* c, ob_denoiser* dn) {
  OB_TRY OB_REQUIRE(c != nullptr, "null handle");
  if (dn) ob::dn_check_shape(dn, c->nx, c->ny);
  c->denoiser = dn;
  OB_CATCH
}

OB_API int ob_sc2p5drb_get_grid(ob_sc2p5drb* c, int which, double* out, double* lims) {
  OB_TRY OB_REQUIRE(c != nullptr, "null handle");
  OB_REQUIRE(which >= 0 && which <= 3, "grid selector must be 0..3");
  ob::sc2d_get_grid(c, which, out, lims);
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
