// ob_denoise.hip -- learned denoiser of deposited charge density (no reference counterpart).
//
// A 3-level U-Net maps a rho grid deposited from few macroparticles to an estimate of the
// grid that many more particles of the same distribution would deposit:
//
//   e1 = c1b(c1a(x))                 16 ch, H x W      (conv3x3 + bias + ReLU each)
//   e2 = c2b(c2a(pool(e1)))          32 ch, H/2
//   e3 = c3d(c3c(c3b(c3a(pool(e2)))))  64 ch, H/4
//   d2 = c4b(c4a([up(e3), e2]))      32 ch, H/2
//   d1 = c5b(c5a([up(d2), e1]))      16 ch, H
//   out = x + head(d1)               1x1 conv, FP32
//
// 191,841 parameters.  Weights and activations are BF16 (activations NHWC), products
// accumulate in FP32 on the matrix cores (v_mfma_f32_16x16x32_bf16); master weights, the
// head and the optimizer (Adam) are FP32.  Every pass is hand-written here: forward, the
// data and weight gradients, the gradient plumbing of pool / upsample / concat, the loss,
// Adam, and the generator of training pairs (seeded particles of four distribution
// families deposited with the TSC weights of Grid2D::binValue, Grid2D.cc:142-166).
//
// Convolutions are implicit GEMMs over 16 x 16 output-pixel tiles: a workgroup of 4 waves
// stages an 18 x 18 halo tile of 16 input channels and the matching weights in LDS; wave w
// owns output rows 4w..4w+3 (one 16-pixel MFMA row tile each) and up to 64 output
// channels.  Pooling, nearest upsampling, channel concatenation and the ReLU mask of the
// backward pass are applied while the halo tile is loaded, so they cost no pass of their
// own.  Weight gradients use the same tiles with pixels as the reduction dimension.
#include <cmath>
#include <cstring>

#include "ob_internal.hpp"

namespace ob {
namespace dn {

using bf16 = uint16_t;
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

// v_mfma_f32_16x16x32_bf16: D(16x16) += A(16x32) B(32x16).  Lane l holds row (A) or
// column (B) l % 16 and eight consecutive k of group l / 16; which k a slot stands for is
// the caller's choice, as long as A and B use the same one.
__device__ __forceinline__ f4 mfma32(uint4 a, uint4 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, a),
                                                 __builtin_bit_cast(bf8, b), c, 0, 0, 0);
}

constexpr int kTile = 16;
constexpr int kHalo = kTile + 2;
constexpr int kThreads = 256;
constexpr int kWgradBlocks = 1024;                         // weight-gradient workgroups
// the largest layer need: 3 chunks per workgroup x 32 output channels (or 1 x 64, 2 x 32)
constexpr size_t kWpartFloats = (size_t)kWgradBlocks * 9 * 16 * 96 + 6 * 9 * 16 * 64;
constexpr int kHeadBlocks = 512;                           // head / loss workgroups

__device__ __forceinline__ float bf2f(bf16 h) { return __uint_as_float((uint32_t)h << 16); }
// Round to nearest, ties to even (finite inputs).
__device__ __forceinline__ bf16 f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (bf16)(u >> 16);
}
__device__ __forceinline__ bool positive(bf16 h) { return (int16_t)h > 0; }

enum SrcMode : int32_t { SRC_SAME = 0, SRC_POOL = 1, SRC_UP = 2 };

// One operand of the logical conv input: channels [coff, coff + C) of an NHWC tensor with
// channel stride cs.  POOL: the tensor is 2H x 2W (2x2 mean); UP: H/2 x W/2 (nearest).
// mask: a tensor of the same layout; an element counts only where the mask is > 0.
struct Src {
  const bf16* p;
  const bf16* mask;
  int32_t C, cs, coff, mode;
};

struct ConvArgs {
  Src s0, s1;          // logical input = [s0, s1] along channels (s1.C = 0: none)
  const float* xin;    // single-channel FP32 input in place of s0/s1 (first layer)
  const bf16* w;       // packed weights (pack_layer)
  const float* bias;   // nullptr: none
  int32_t Cin, Cout, H, W;
  int32_t relu;
  bf16* out;           // output view: channel stride ocs, first channel ocoff
  int32_t ocs, ocoff;
  // weight-gradient pass
  const bf16* dy;      // gradient of the layer output, channel stride Cout
  const bf16* y;       // the layer output (ReLU mask), channel stride Cout
  float* dw;           // [9][Cin][Cout]
  float* db;           // [Cout]
  float* wpart;        // per-workgroup partial sums of the weight-gradient pass
  float* bpart;
};

__device__ __forceinline__ void unpack8(const uint4& v, float f[8]) {
  const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
  for (int k = 0; k < 4; k++) {
    f[2 * k] = __uint_as_float(w[k] << 16);
    f[2 * k + 1] = __uint_as_float(w[k] & 0xFFFF0000u);
  }
}
__device__ __forceinline__ uint4 pack8(const float f[8]) {
  uint32_t w[4];
#pragma unroll
  for (int k = 0; k < 4; k++) w[k] = (uint32_t)f2bf(f[2 * k]) | ((uint32_t)f2bf(f[2 * k + 1]) << 16);
  return make_uint4(w[0], w[1], w[2], w[3]);
}
__device__ __forceinline__ uint4 ld16(const bf16* p) { return *(const uint4*)p; }

// Channels [c8, c8 + 8) of logical input pixel (y, x) of image b (zero outside the image).
__device__ __forceinline__ uint4 load8(const ConvArgs& a, int b, int y, int x, int c8) {
  uint4 r = make_uint4(0, 0, 0, 0);
  if (y < 0 || y >= a.H || x < 0 || x >= a.W) return r;
  if (a.xin) {
    if (c8 == 0) r.x = f2bf(a.xin[((size_t)b * a.H + y) * a.W + x]);
    return r;
  }
  if (c8 >= a.Cin) return r;
  const bool first = c8 < a.s0.C;
  const Src& s = first ? a.s0 : a.s1;
  const int c = (first ? c8 : c8 - a.s0.C) + s.coff;
  if (s.mode == SRC_POOL) {
    const int SW = 2 * a.W;
    const size_t base = ((size_t)b * 2 * a.H + 2 * y) * SW + 2 * x;
    float f[8] = {0, 0, 0, 0, 0, 0, 0, 0}, g[8];
    const size_t off[4] = {base, base + 1, base + SW, base + SW + 1};
#pragma unroll
    for (int k = 0; k < 4; k++) {
      unpack8(ld16(s.p + off[k] * s.cs + c), g);
#pragma unroll
      for (int j = 0; j < 8; j++) f[j] += g[j];
    }
#pragma unroll
    for (int j = 0; j < 8; j++) f[j] *= 0.25f;
    return pack8(f);
  }
  size_t pix;
  if (s.mode == SRC_UP)
    pix = ((size_t)b * (a.H >> 1) + (y >> 1)) * (a.W >> 1) + (x >> 1);
  else
    pix = ((size_t)b * a.H + y) * a.W + x;
  r = ld16(s.p + pix * s.cs + c);
  if (s.mask) {
    const uint4 m = ld16(s.mask + pix * s.cs + c);
    const uint32_t v[4] = {r.x, r.y, r.z, r.w}, mm[4] = {m.x, m.y, m.z, m.w};
    uint32_t o[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const uint32_t lo = positive((bf16)(mm[k] & 0xFFFFu)) ? (v[k] & 0xFFFFu) : 0u;
      const uint32_t hi = positive((bf16)(mm[k] >> 16)) ? (v[k] & 0xFFFF0000u) : 0u;
      o[k] = lo | hi;
    }
    r = make_uint4(o[0], o[1], o[2], o[3]);
  }
  return r;
}

// ---- forward / data-gradient convolution -------------------------------------------
// grid: (tiles of the image, Cout / (16 NT), images).  Packed weights: element
// ((ci / 16 * 9 + tap) * Cout + co) * 16 + ci % 16, tap = 3 * ky + kx.
// One MFMA takes two taps of a 16-channel chunk (k = 16 * (tap & 1) + ci % 16): taps
// (0, 1), (2, 3), (4, 5), (6, 7) and (8, zeros), 5 MFMAs per output tile and chunk.
// In LDS, halo pixels and weight rows are 32-byte rows of 16 channels whose two 16-byte
// halves are swapped in rows with bit 3 set: an operand read (16 consecutive rows, one half
// each) then covers all 64 banks once.
__device__ __forceinline__ int swz(int row, int h) {
  return row * 16 + 8 * (h ^ ((row >> 3) & 1));
}
template <int NT>
__global__ __launch_bounds__(kThreads) void conv3x3_fwd(ConvArgs a) {
  constexpr int NCO = NT * 16;
  constexpr int XS = kHalo * kHalo * 16;  // bf16 elements of the halo tile
  constexpr int WS = 9 * NCO * 16;        // weights of one 16-channel chunk
  // staged output: pixel rows of NCO + 8 channels, so that the four pixel groups of an
  // accumulator write (4 pixels apart) start 16 banks apart
  constexpr int OP = NCO + 8;
  constexpr int OS = kTile * kTile * OP;
  constexpr int SMEM = (XS + WS) > OS ? (XS + WS) : OS;
  // a thread's share of one chunk: NH halo pixel-halves and NW weight runs of 8 channels
  constexpr int NH = (kHalo * kHalo * 2 + kThreads - 1) / kThreads;
  constexpr int NW = (9 * NCO * 2 + kThreads - 1) / kThreads;
  __shared__ __align__(16) bf16 smem[SMEM];
  bf16* sx = smem;
  bf16* sw = smem + XS;
  const int tiles_x = a.W / kTile;
  const int ty0 = (blockIdx.x / tiles_x) * kTile, tx0 = (blockIdx.x % tiles_x) * kTile;
  const int co0 = blockIdx.y * NCO;
  const int b = blockIdx.z;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, q = lane >> 4, l16 = lane & 15;
  f4 acc[4][NT];
#pragma unroll
  for (int m = 0; m < 4; m++)
#pragma unroll
    for (int n = 0; n < NT; n++) acc[m][n] = f4{0.f, 0.f, 0.f, 0.f};
  const int nchunks = (a.Cin + 15) / 16;
  // With two or more output tiles, a chunk's global loads go to registers first, all in
  // flight together, and the next chunk's are issued before this chunk's MFMAs, so their
  // latency is hidden behind them.  With one, the registers that takes are worth more as
  // occupancy (120 against 64 VGPRs, measured slower): its loads go straight to LDS.
  constexpr bool kPrefetch = NT >= 2;
  uint4 hv[kPrefetch ? NH : 1], wvv[kPrefetch ? NW : 1];
  auto fetch = [&](int cc) {
    if constexpr (!kPrefetch) return;
#pragma unroll
    for (int u = 0; u < NH; u++) {
      const int e = (int)threadIdx.x + u * kThreads;
      const int pix = e >> 1, h = e & 1;
      const int hy = pix / kHalo, hx = pix - hy * kHalo;
      hv[u] = e < kHalo * kHalo * 2 ? load8(a, b, ty0 + hy - 1, tx0 + hx - 1, cc * 16 + h * 8)
                                    : make_uint4(0, 0, 0, 0);
    }
    const bf16* wsrc = a.w + (size_t)cc * 9 * a.Cout * 16;
#pragma unroll
    for (int u = 0; u < NW; u++) {
      const int e = (int)threadIdx.x + u * kThreads;
      const int tap = e / (NCO * 2), r = e - tap * (NCO * 2);
      const int n = r >> 1, h = r & 1;
      wvv[u] = e < 9 * NCO * 2 ? ld16(wsrc + ((size_t)tap * a.Cout + co0 + n) * 16 + h * 8)
                               : make_uint4(0, 0, 0, 0);
    }
  };
  fetch(0);
  for (int cc = 0; cc < nchunks; cc++) {
    __syncthreads();  // the previous chunk's MFMAs are done with sx and sw
    if constexpr (kPrefetch) {
#pragma unroll
      for (int u = 0; u < NH; u++) {
        const int e = (int)threadIdx.x + u * kThreads;
        if (e < kHalo * kHalo * 2) *(uint4*)&sx[swz(e >> 1, e & 1)] = hv[u];
      }
#pragma unroll
      for (int u = 0; u < NW; u++) {
        const int e = (int)threadIdx.x + u * kThreads;
        if (e < 9 * NCO * 2) *(uint4*)&sw[swz(e >> 1, e & 1)] = wvv[u];
      }
    } else {
      for (int e = threadIdx.x; e < kHalo * kHalo * 2; e += kThreads) {
        const int pix = e >> 1, h = e & 1;
        const int hy = pix / kHalo, hx = pix - hy * kHalo;
        *(uint4*)&sx[swz(pix, h)] = load8(a, b, ty0 + hy - 1, tx0 + hx - 1, cc * 16 + h * 8);
      }
      const bf16* wsrc = a.w + (size_t)cc * 9 * a.Cout * 16;
      for (int e = threadIdx.x; e < 9 * NCO * 2; e += kThreads) {
        const int tap = e / (NCO * 2), r = e - tap * (NCO * 2);
        const int n = r >> 1, h = r & 1;
        *(uint4*)&sw[swz(e >> 1, h)] = ld16(wsrc + ((size_t)tap * a.Cout + co0 + n) * 16 + h * 8);
      }
    }
    __syncthreads();
    if (cc + 1 < nchunks) fetch(cc + 1);
    // lane group q: tap 2 tp + q / 2, channels 8 (q % 2) .. 8 (q % 2) + 7
    const int h = q & 1;
    auto step = [&](int tap, bool live) {
      const int ky = tap / 3, kx = tap - 3 * ky;
      uint4 bfr[NT];
#pragma unroll
      for (int n = 0; n < NT; n++)
        bfr[n] = live ? *(const uint4*)&sw[swz(tap * NCO + n * 16 + l16, h)] : make_uint4(0, 0, 0, 0);
#pragma unroll
      for (int m = 0; m < 4; m++) {
        const int r = wv * 4 + m;
        const uint4 afr = live ? *(const uint4*)&sx[swz((r + ky) * kHalo + l16 + kx, h)]
                               : make_uint4(0, 0, 0, 0);
#pragma unroll
        for (int n = 0; n < NT; n++) acc[m][n] = mfma32(afr, bfr[n], acc[m][n]);
      }
    };
#pragma unroll 1
    for (int tp = 0; tp < 4; tp++) step(2 * tp + (q >> 1), true);
    step(8, q < 2);
  }
  __syncthreads();
  bf16* so = smem;  // [pixel][OP]
#pragma unroll
  for (int m = 0; m < 4; m++)
#pragma unroll
    for (int n = 0; n < NT; n++) {
      const int ch = n * 16 + l16;
      const float bias = a.bias ? a.bias[co0 + ch] : 0.f;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        float v = acc[m][n][j] + bias;
        if (a.relu) v = fmaxf(v, 0.f);
        so[((wv * 4 + m) * kTile + 4 * q + j) * OP + ch] = f2bf(v);
      }
    }
  __syncthreads();
  for (int e = threadIdx.x; e < kTile * kTile * NCO / 8; e += kThreads) {
    const int pix = e / (NCO / 8), u = e - pix * (NCO / 8);
    const int y = ty0 + pix / kTile, x = tx0 + pix % kTile;
    *(uint4*)&a.out[(((size_t)b * a.H + y) * a.W + x) * a.ocs + a.ocoff + co0 + u * 8] =
        *(const uint4*)&so[pix * OP + u * 8];
  }
}

// ---- weight gradient ------------------------------------------------------------------
// dW[tap][ci][co] += sum_p X[p + off(tap)][ci] * dY'[p][co], dY' = dY where Y > 0.
// grid: (blocks over all tiles of the batch, groups of CH Cin chunks of 16).  Pixels are the
// reduction dimension: the halo tile is stored channel-major in three copies shifted by kx,
// so every MFMA operand is an aligned 8-pixel run (an MFMA takes two tile rows, 32
// pixels).  Wave w accumulates the (tap, co-tile)
// pairs w, w + 4, ... of each of its CH chunks; dY' is staged once per tile for all of them
// (a workgroup per chunk would read dY and Y once per chunk).  Chunk 0 also sums the bias
// gradient.
template <int MT, int CH>
__global__ __launch_bounds__(kThreads) void conv3x3_wgrad(ConvArgs a, int ntiles) {
  constexpr int COUT = MT * 16;
  constexpr int NP = (9 * MT + 3) / 4;
  // a thread's share of one tile: NH halo pixel-halves per chunk and ND 8-channel runs of
  // dY and Y
  constexpr int NH = (kHalo * kHalo * 2 + kThreads - 1) / kThreads;
  constexpr int ND = (kTile * kTile * COUT / 8) / kThreads;
  static_assert(ND * kThreads == kTile * kTile * COUT / 8, "dY tile split");
  // Channel planes padded to 148 (xs) and 132 (ds) dwords, both 4 mod 8: the 16 channels of
  // an MFMA operand read start on 16 distinct multiples of 4 banks, so each 16 lanes'
  // 16-byte reads cover the 64 banks once, and the two 8-channel halves of a staging write
  // are 32 banks apart.
  constexpr int SX = kHalo * kTile + 8, SD = kTile * kTile + 8;
  __shared__ __align__(16) bf16 xs[3][16][SX];
  __shared__ __align__(16) bf16 ds[COUT][SD];
  const int cc0 = blockIdx.y * CH;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, q = lane >> 4, l16 = lane & 15;
  const int tiles_x = a.W / kTile, tiles_img = tiles_x * (a.H / kTile);
  f4 acc[CH][NP];
#pragma unroll
  for (int h = 0; h < CH; h++)
#pragma unroll
    for (int k = 0; k < NP; k++) acc[h][k] = f4{0.f, 0.f, 0.f, 0.f};
  // bias gradient (chunk 0): a thread's dY runs are always channels w8 * 8 .. w8 * 8 + 7
  // (kThreads is a multiple of COUT / 8), summed in registers as they are staged
  float bsum[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  // The tile's global loads go to registers first, all in flight together; the next tile's
  // are issued before this tile's last chunk of MFMAs, so their latency is hidden behind it.
  uint4 hv[CH][NH], dv[ND], yv[ND];
  auto fetch = [&](int t) {
    const int b = t / tiles_img, ti = t - b * tiles_img;
    const int ty0 = (ti / tiles_x) * kTile, tx0 = (ti % tiles_x) * kTile;
#pragma unroll
    for (int h = 0; h < CH; h++)
#pragma unroll
      for (int u = 0; u < NH; u++) {
        const int e = (int)threadIdx.x + u * kThreads;
        const int pix = e >> 1, hf = e & 1;
        const int hy = pix / kHalo, hx = pix - hy * kHalo;
        hv[h][u] = e < kHalo * kHalo * 2
                       ? load8(a, b, ty0 + hy - 1, tx0 + hx - 1, (cc0 + h) * 16 + hf * 8)
                       : make_uint4(0, 0, 0, 0);
      }
#pragma unroll
    for (int u = 0; u < ND; u++) {
      const int e = (int)threadIdx.x + u * kThreads;
      const int pix = e / (COUT / 8), w8 = e - pix * (COUT / 8);
      const int py = pix / kTile, px = pix % kTile;
      const size_t g = (((size_t)b * a.H + ty0 + py) * a.W + tx0 + px) * COUT + w8 * 8;
      dv[u] = ld16(a.dy + g);
      yv[u] = ld16(a.y + g);
    }
  };
  int t = blockIdx.x;
  if (t < ntiles) fetch(t);
  for (; t < ntiles; t += gridDim.x) {
    __syncthreads();  // the previous tile's MFMAs are done with xs and ds
#pragma unroll
    for (int u = 0; u < ND; u++) {
      const int e = (int)threadIdx.x + u * kThreads;
      const int pix = e / (COUT / 8), w8 = e - pix * (COUT / 8);
      const int py = pix / kTile, px = pix % kTile;
      const uint32_t d4[4] = {dv[u].x, dv[u].y, dv[u].z, dv[u].w};
      const uint32_t y4[4] = {yv[u].x, yv[u].y, yv[u].z, yv[u].w};
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const uint32_t sh = (j & 1) ? 16 : 0;
        const bf16 d = (bf16)((d4[j >> 1] >> sh) & 0xFFFFu);
        const bf16 yy = (bf16)((y4[j >> 1] >> sh) & 0xFFFFu);
        const bf16 dm = positive(yy) ? d : (bf16)0;
        ds[w8 * 8 + j][py * kTile + px] = dm;
        bsum[j] += bf2f(dm);
      }
    }
#pragma unroll
    for (int h = 0; h < CH; h++) {
      if (h > 0) __syncthreads();  // the previous chunk's MFMAs are done with xs
#pragma unroll
      for (int u = 0; u < NH; u++) {
        const int e = (int)threadIdx.x + u * kThreads;
        if (e >= kHalo * kHalo * 2) break;
        const int pix = e >> 1, hf = e & 1;
        const int hy = pix / kHalo, hx = pix - hy * kHalo;
        const uint32_t w4[4] = {hv[h][u].x, hv[h][u].y, hv[h][u].z, hv[h][u].w};
#pragma unroll
        for (int j = 0; j < 8; j++) {
          const bf16 val = (bf16)((j & 1) ? (w4[j >> 1] >> 16) : (w4[j >> 1] & 0xFFFFu));
#pragma unroll
          for (int sh = 0; sh < 3; sh++) {
            const int c = hx - sh;
            if (c >= 0 && c < kTile) xs[sh][hf * 8 + j][hy * kTile + c] = val;
          }
        }
      }
      __syncthreads();
      if (h == CH - 1 && t + (int)gridDim.x < ntiles) fetch(t + gridDim.x);
#pragma unroll
      for (int k = 0; k < NP; k++) {
        const int p = wv + 4 * k;
        if (p < 9 * MT) {
          const int tap = p / MT, mt = p - tap * MT;
          const int ky = tap / 3, kx = tap - 3 * ky;
          // k = 32 pixels: tile rows r and r + 1, lane group q the 8-pixel run 8 q .. 8 q + 7
#pragma unroll 2
          for (int r = 0; r < kTile; r += 2) {
            const uint4 af = *(const uint4*)&ds[mt * 16 + l16][r * kTile + 8 * q];
            const uint4 bfr = *(const uint4*)&xs[kx][l16][(r + ky) * kTile + 8 * q];
            acc[h][k] = mfma32(af, bfr, acc[h][k]);
          }
        }
      }
    }
  }
  // Partial sums of this workgroup (every element written once: no atomics, which would
  // pile up on the same few thousand addresses from every workgroup); conv3x3_wgrad_reduce
  // adds them up.  Layout [chunk][workgroup][tap][16 ci][COUT].
#pragma unroll
  for (int h = 0; h < CH; h++) {
    float* wp = a.wpart + ((size_t)(cc0 + h) * gridDim.x + blockIdx.x) * (9 * 16 * COUT);
#pragma unroll
    for (int k = 0; k < NP; k++) {
      const int p = wv + 4 * k;
      if (p < 9 * MT) {
        const int tap = p / MT, mt = p - tap * MT;
#pragma unroll
        for (int j = 0; j < 4; j++)
          wp[(tap * 16 + l16) * COUT + mt * 16 + 4 * q + j] = acc[h][k][j];
      }
    }
  }
  if (cc0 == 0) {
    // threads with equal threadIdx % (COUT / 8) hold the same 8 channels: sum them over
    // the wave by shuffles, then over the 4 waves in LDS (xs is free by now)
    constexpr int G = COUT / 8;
#pragma unroll
    for (int j = 0; j < 8; j++)
      for (int o = G; o < 64; o <<= 1) bsum[j] += __shfl_xor(bsum[j], o, 64);
    __syncthreads();
    float* red = (float*)&xs[0][0][0];  // [4 waves][COUT]
    if (lane < G)
#pragma unroll
      for (int j = 0; j < 8; j++) red[wv * COUT + lane * 8 + j] = bsum[j];
    __syncthreads();
    if (threadIdx.x < COUT) {
      const float v = red[threadIdx.x] + red[COUT + threadIdx.x] + red[2 * COUT + threadIdx.x] +
                      red[3 * COUT + threadIdx.x];
      a.bpart[(size_t)blockIdx.x * COUT + threadIdx.x] = v;
    }
  }
}

// dW[tap][ci][co] = sum over workgroups of the partials; db[co] likewise (chunk 0).  A
// workgroup of 1024 threads takes 64 consecutive outputs: its 16 waves sum the
// workgroups' partials in 16 interleaved groups (lanes read 64 consecutive outputs, eight
// partials in flight per lane), and wave 0 adds the 16 group sums in a fixed order.
constexpr int kReduceGroups = 16;
__global__ __launch_bounds__(1024) void conv3x3_wgrad_reduce(const float* __restrict__ wpart,
                                                             const float* __restrict__ bpart,
                                                             int nblk, int Cin, int Cout,
                                                             float* __restrict__ dw,
                                                             float* __restrict__ db) {
  __shared__ float red[kReduceGroups][64];
  const int nw = 9 * Cin * Cout;
  const int e = blockIdx.x * 64 + (threadIdx.x & 63), grp = threadIdx.x >> 6;
  float s = 0.f;
  if (e < nw + Cout) {
    const float* p;
    size_t stride;
    if (e < nw) {
      const int co = e % Cout, r = e / Cout;
      const int ci = r % Cin, tap = r / Cin;
      const int cc = ci / 16, l = ci % 16;
      p = wpart + (size_t)cc * nblk * (9 * 16 * Cout) + (tap * 16 + l) * Cout + co;
      stride = (size_t)9 * 16 * Cout;
    } else {
      p = bpart + (e - nw);
      stride = (size_t)Cout;
    }
    constexpr int U = 8;
    int b = grp;
    for (; b + (U - 1) * kReduceGroups < nblk; b += U * kReduceGroups) {
      float v[U];
#pragma unroll
      for (int u = 0; u < U; u++) v[u] = p[(size_t)(b + u * kReduceGroups) * stride];
#pragma unroll
      for (int u = 0; u < U; u++) s += v[u];
    }
    for (; b < nblk; b += kReduceGroups) s += p[(size_t)b * stride];
  }
  red[grp][threadIdx.x & 63] = s;
  __syncthreads();
  if (grp == 0 && e < nw + Cout) {
    float t = 0.f;
#pragma unroll
    for (int g = 0; g < kReduceGroups; g++) t += red[g][threadIdx.x];
    if (e < nw)
      dw[e] = t;
    else
      db[e - nw] = t;
  }
}

// ---- head, loss, gradient plumbing ----------------------------------------------------
// out = x + hb + sum_c hw[c] d1[c]   (hw: 16 weights then the bias, FP32)
__global__ __launch_bounds__(kThreads) void head_fwd(const bf16* __restrict__ d1,
                                                     const float* __restrict__ x,
                                                     const float* __restrict__ hw,
                                                     float* __restrict__ out, int64_t npix) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npix;
       p += (int64_t)gridDim.x * blockDim.x) {
    float f[16];
    unpack8(ld16(d1 + p * 16), f);
    unpack8(ld16(d1 + p * 16 + 8), f + 8);
    float s = hw[16];
#pragma unroll
    for (int c = 0; c < 16; c++) s += hw[c] * f[c];
    out[p] = x[p] + s;
  }
}

__device__ __forceinline__ float block_sum(float v, float* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[w] = v;
  __syncthreads();
  float s = 0.f;
  for (int k = 0; k < (int)(blockDim.x >> 6); k++) s += sh[k];
  return s;
}

// Multi-scale error pyramid of the loss: m_s = mean of e = out - y over 2^s x 2^s blocks,
// s = 1..S.  Low spatial frequencies of the density error dominate the field error, so
// the loss adds lam_s * mean(m_s^2) for each scale.
constexpr int kMaxScales = 8;
struct Pyramid {
  float* m[kMaxScales + 1];  // m[s]: [B][H >> s][W >> s]  (m[0] unused)
  float lam[kMaxScales + 1];
  int32_t S, H, W;
};

// Level s from level s - 1 (s = 1: from out - y); loss += lam_s * mean over blocks of m_s^2.
__global__ __launch_bounds__(kThreads) void pyramid_level(const float* __restrict__ out,
                                                          const float* __restrict__ y,
                                                          Pyramid P, int s, int B,
                                                          float* __restrict__ loss) {
  __shared__ float sh[kThreads / 64];
  const int h = P.H >> s, w = P.W >> s, W2 = P.W >> (s - 1);
  const int64_t n = (int64_t)B * h * w;
  float l = 0.f;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int x = (int)(e % w);
    const int64_t r = e / w;
    const int yy = (int)(r % h);
    const int64_t b = r / h;
    const int64_t i0 = (b * (P.H >> (s - 1)) + 2 * yy) * W2 + 2 * x;
    const int64_t idx[4] = {i0, i0 + 1, i0 + W2, i0 + W2 + 1};
    float v = 0.f;
#pragma unroll
    for (int k = 0; k < 4; k++) v += s == 1 ? out[idx[k]] - y[idx[k]] : P.m[s - 1][idx[k]];
    v *= 0.25f;
    P.m[s][e] = v;
    l += v * v;
  }
  const float t = block_sum(l, sh);
  if (threadIdx.x == 0) atomicAdd(loss, P.lam[s] * t / (float)n);
}

// Loss over npix pixels and its gradient: dd1 = g hw, dhw += sum g d1, dhb += sum g, with
// g = 2 (e + sum_s lam_s m_s(block of p)) / npix, e = out - y; loss += sum e^2 / npix
// (the pyramid terms are added by pyramid_level).
__global__ __launch_bounds__(kThreads) void head_bwd(const bf16* __restrict__ d1,
                                                     const float* __restrict__ out,
                                                     const float* __restrict__ y,
                                                     const float* __restrict__ hw,
                                                     bf16* __restrict__ dd1,
                                                     float* __restrict__ part, int64_t npix,
                                                     Pyramid P) {
  __shared__ float sh[kThreads / 64];
  float acc[17];
#pragma unroll
  for (int c = 0; c < 17; c++) acc[c] = 0.f;
  float l = 0.f;
  const float inv = 1.0f / (float)npix;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npix;
       p += (int64_t)gridDim.x * blockDim.x) {
    const float e = out[p] - y[p];
    l += e * e;
    float es = e;
    if (P.S > 0) {
      const int x = (int)(p % P.W);
      const int64_t r = p / P.W;
      const int yy = (int)(r % P.H);
      const int64_t b = r / P.H;
      for (int sc = 1; sc <= P.S; sc++)
        es += P.lam[sc] * P.m[sc][(b * (P.H >> sc) + (yy >> sc)) * (P.W >> sc) + (x >> sc)];
    }
    const float g = 2.0f * es * inv;
    float f[16], o[16];
    unpack8(ld16(d1 + p * 16), f);
    unpack8(ld16(d1 + p * 16 + 8), f + 8);
#pragma unroll
    for (int c = 0; c < 16; c++) {
      acc[c] += g * f[c];
      o[c] = g * hw[c];
    }
    acc[16] += g;
    *(uint4*)(dd1 + p * 16) = pack8(o);
    *(uint4*)(dd1 + p * 16 + 8) = pack8(o + 8);
  }
  // per-workgroup partials [block][18]: 17 head gradients, then the loss
  for (int c = 0; c < 17; c++) {
    const float s = block_sum(acc[c], sh);
    if (threadIdx.x == 0) part[blockIdx.x * 18 + c] = s;
  }
  const float s = block_sum(l, sh);
  if (threadIdx.x == 0) part[blockIdx.x * 18 + 17] = s * inv;
}

// dhw[c] += sum of the partials, loss += sum of the loss partials.  Workgroup c (18 of
// them, one wave each) sums column c: lane l adds partials l, l + 64, ... (eight loads in
// flight), then the wave's lanes are added by a butterfly.
__global__ __launch_bounds__(64) void head_reduce(const float* __restrict__ part, int nblk,
                                                  float* __restrict__ dhw,
                                                  float* __restrict__ loss) {
  const int c = blockIdx.x;
  float s = 0.f;
  constexpr int U = 8;
  int b = threadIdx.x;
  for (; b + (U - 1) * 64 < nblk; b += U * 64) {
    float v[U];
#pragma unroll
    for (int u = 0; u < U; u++) v[u] = part[(b + u * 64) * 18 + c];
#pragma unroll
    for (int u = 0; u < U; u++) s += v[u];
  }
  for (; b < nblk; b += 64) s += part[b * 18 + c];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (threadIdx.x != 0) return;
  if (c < 17)
    dhw[c] += s;
  else
    *loss += s;
}

// g[b, y, x, c] = 0.25 gp[b, y/2, x/2, c] + gs[b, y, x, soff + c]   (C channels, H x W)
__global__ __launch_bounds__(kThreads) void grad_pool_skip(const bf16* __restrict__ gp,
                                                           const bf16* __restrict__ gs, int gss,
                                                           int soff, bf16* __restrict__ g, int C,
                                                           int H, int W, int64_t total8) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total8;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int c8 = (int)(e % (C / 8)) * 8;
    const int64_t pix = e / (C / 8);
    const int x = (int)(pix % W);
    const int64_t r = pix / W;
    const int y = (int)(r % H);
    const int64_t b = r / H;
    const int64_t pp = (b * (H / 2) + y / 2) * (W / 2) + x / 2;
    float a[8], s[8];
    unpack8(ld16(gp + pp * C + c8), a);
    unpack8(ld16(gs + pix * gss + soff + c8), s);
#pragma unroll
    for (int j = 0; j < 8; j++) a[j] = 0.25f * a[j] + s[j];
    *(uint4*)(g + pix * C + c8) = pack8(a);
  }
}

// g[b, y, x, c] = sum over the 2x2 block of gu[b, 2y + i, 2x + j, c]  (gu channel stride gus)
__global__ __launch_bounds__(kThreads) void grad_up(const bf16* __restrict__ gu, int gus,
                                                    bf16* __restrict__ g, int C, int H, int W,
                                                    int64_t total8) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total8;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int c8 = (int)(e % (C / 8)) * 8;
    const int64_t pix = e / (C / 8);
    const int x = (int)(pix % W);
    const int64_t r = pix / W;
    const int y = (int)(r % H);
    const int64_t b = r / H;
    const int64_t base = (b * 2 * H + 2 * y) * (2 * W) + 2 * x;
    const int64_t src[4] = {base, base + 1, base + 2 * W, base + 2 * W + 1};
    float a[8] = {0, 0, 0, 0, 0, 0, 0, 0}, t[8];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      unpack8(ld16(gu + src[k] * gus + c8), t);
#pragma unroll
      for (int j = 0; j < 8; j++) a[j] += t[j];
    }
    *(uint4*)(g + pix * C + c8) = pack8(a);
  }
}

// ---- parameters -------------------------------------------------------------------------
// fp32 W[9][Cin][Cout] -> packed forward weights and packed data-gradient weights (the
// transposed, tap-flipped kernel: a conv with Cin' = Cout, Cout' = Cin).
__device__ __forceinline__ void pack_layer_body(const float* __restrict__ W, int Cin,
                                                int Cout, bf16* __restrict__ pf,
                                                bf16* __restrict__ pt) {
  const int Cp = (Cin + 15) / 16 * 16;
  const int nf = 9 * Cp * Cout;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nf; e += gridDim.x * blockDim.x) {
    // e = ((ci / 16 * 9 + tap) * Cout + co) * 16 + ci % 16
    const int k = e % 16, r = e / 16;
    const int co = r % Cout, r2 = r / Cout;
    const int tap = r2 % 9, ch = r2 / 9;
    const int ci = ch * 16 + k;
    pf[e] = ci < Cin ? f2bf(W[((size_t)tap * Cin + ci) * Cout + co]) : (bf16)0;
  }
  if (!pt) return;
  const int nt = 9 * Cout * Cin;  // Cout is a multiple of 16
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nt; e += gridDim.x * blockDim.x) {
    // e = ((co / 16 * 9 + tap') * Cin + ci) * 16 + co % 16
    const int k = e % 16, r = e / 16;
    const int ci = r % Cin, r2 = r / Cin;
    const int tp = r2 % 9, ch = r2 / 9;
    const int co = ch * 16 + k;
    pt[e] = f2bf(W[((size_t)(8 - tp) * Cin + ci) * Cout + co]);
  }
}

// Every layer in one launch: blockIdx.y = layer.
struct PackTable {
  int32_t w_off[12], cin[12], cout[12];
  int64_t pf_off[12], pt_off[12];
};
__global__ __launch_bounds__(kThreads) void pack_layers(const float* __restrict__ w,
                                                        bf16* __restrict__ pf,
                                                        bf16* __restrict__ pt, PackTable t) {
  const int l = blockIdx.y;
  pack_layer_body(w + t.w_off[l], t.cin[l], t.cout[l], pf + t.pf_off[l],
                  pt ? pt + t.pt_off[l] : nullptr);
}

__global__ __launch_bounds__(kThreads) void adam_step(float* __restrict__ w,
                                                      const float* __restrict__ g,
                                                      float* __restrict__ m,
                                                      float* __restrict__ v, int64_t n,
                                                      float lr, float b1, float b2, float eps,
                                                      float c1, float c2) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
  }
}

__global__ __launch_bounds__(kThreads) void fill_f32(float* __restrict__ p, int64_t n, float v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// ---- use inside a space-charge step -----------------------------------------------------
// Charge-normalised images for the network and back.  Each grid is split into kRhoParts
// equal slices; one workgroup of kRhoThreads sums a slice (pass 1), and in pass 2 every
// workgroup of the grid adds the kRhoParts partial sums in slice order (so all of them see
// the same total), scales its slice and, for slice 0, stores the total.
constexpr int kRhoParts = 16;
constexpr int kRhoThreads = 256;

// Up to 64 grids per launch: grid j is rho[j] with skip flag skip[j] and image j.
struct RhoJobs {
  double* rho[64];
  const int32_t* skip[64];
};

// Sum of v[a..b) by one workgroup, eight loads in flight per thread, in a fixed order.
template <typename T>
__device__ __forceinline__ double slice_sum(const T* __restrict__ v, int a, int b) {
  __shared__ double sh[kRhoThreads / 64];
  constexpr int U = 8;
  double s = 0.0;
  int i = a + threadIdx.x;
  for (; i + (U - 1) * kRhoThreads < b; i += U * kRhoThreads) {
    double t[U];
#pragma unroll
    for (int u = 0; u < U; u++) t[u] = (double)v[i + u * kRhoThreads];
#pragma unroll
    for (int u = 0; u < U; u++) s += t[u];
  }
  for (; i < b; i += kRhoThreads) s += (double)v[i];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int w = 0; w < kRhoThreads / 64; w++) t += sh[w];
  return t;
}

// dst[a..b) = src[a..b) * scale, eight elements in flight per thread.
template <typename S, typename D>
__device__ __forceinline__ void slice_scale(const S* __restrict__ src, int a, int b,
                                            double scale, D* __restrict__ dst) {
  constexpr int U = 8;
  int i = a + threadIdx.x;
  for (; i + (U - 1) * kRhoThreads < b; i += U * kRhoThreads) {
    double t[U];
#pragma unroll
    for (int u = 0; u < U; u++) t[u] = (double)src[i + u * kRhoThreads];
#pragma unroll
    for (int u = 0; u < U; u++) dst[i + u * kRhoThreads] = (D)(t[u] * scale);
  }
  for (; i < b; i += kRhoThreads) dst[i] = (D)((double)src[i] * scale);
}

__device__ __forceinline__ double parts_total(const double* __restrict__ part) {
  double t = 0.0;
#pragma unroll
  for (int k = 0; k < kRhoParts; k++) t += part[k];
  return t;
}

__device__ __forceinline__ int slice_begin(int n, int k) {
  return (int)(((int64_t)n * k) / kRhoParts);
}

// Pass 1 over rho: part[j * kRhoParts + k] = sum of slice k of grid j.  Grid (kRhoParts, count).
__global__ __launch_bounds__(kRhoThreads) void rho_parts(RhoJobs jobs, int n,
                                                         double* __restrict__ part) {
  const int j = blockIdx.y, k = blockIdx.x;
  if (*jobs.skip[j]) return;
  const double t = slice_sum(jobs.rho[j], slice_begin(n, k), slice_begin(n, k + 1));
  if (threadIdx.x == 0) part[j * kRhoParts + k] = t;
}

// Pass 2: image j = rho_j * n / total_j (mean 1); total[j] = total_j.
__global__ __launch_bounds__(kRhoThreads) void rho_images(RhoJobs jobs, int n,
                                                          const double* __restrict__ part,
                                                          float* __restrict__ x,
                                                          double* __restrict__ total) {
  const int j = blockIdx.y, k = blockIdx.x;
  if (*jobs.skip[j]) return;
  const double t = parts_total(part + j * kRhoParts);
  slice_scale(jobs.rho[j], slice_begin(n, k), slice_begin(n, k + 1),
              t > 0.0 ? (double)n / t : 0.0, x + (size_t)j * n);
  if (k == 0 && threadIdx.x == 0) total[j] = t;
}

// Pass 1 over the network's output images.
__global__ __launch_bounds__(kRhoThreads) void image_parts(const float* __restrict__ out, int n,
                                                           RhoJobs jobs,
                                                           double* __restrict__ part) {
  const int j = blockIdx.y, k = blockIdx.x;
  if (*jobs.skip[j]) return;
  const double t = slice_sum(out + (size_t)j * n, slice_begin(n, k), slice_begin(n, k + 1));
  if (threadIdx.x == 0) part[j * kRhoParts + k] = t;
}

// Pass 2: rho_j = out_j * total_j / sum(out_j), the denoised grid with the deposited charge;
// a zero-sum output keeps the deposited grid.
__global__ __launch_bounds__(kRhoThreads) void images_rho(const float* __restrict__ out, int n,
                                                          RhoJobs jobs,
                                                          const double* __restrict__ part,
                                                          const double* __restrict__ total) {
  const int j = blockIdx.y, k = blockIdx.x;
  if (*jobs.skip[j]) return;
  const double t = parts_total(part + j * kRhoParts);
  if (t == 0.0) return;
  slice_scale(out + (size_t)j * n, slice_begin(n, k), slice_begin(n, k + 1), total[j] / t,
              jobs.rho[j]);
}

// dst image j <- src image idx[j] (j < count), images of n floats.
struct GatherIdx {
  int32_t idx[64];
};
__global__ __launch_bounds__(kThreads) void gather_images(const float* __restrict__ src,
                                                          float* __restrict__ dst, int64_t n,
                                                          GatherIdx g) {
  const int j = blockIdx.y;
  const float* a = src + (int64_t)g.idx[j] * n;
  float* b = dst + (int64_t)j * n;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

// ---- training pairs ------------------------------------------------------------------------
// Counter-based uniform variates: splitmix64 of (seed, pair, particle, draw).
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ float unif(uint64_t h) {  // (0, 1)
  return ((float)(h >> 40) + 0.5f) * (1.0f / 16777216.0f);
}

struct PairSpec {        // one training pair's distribution
  int32_t family;        // 0 Gaussian, 1 KV (uniform disk), 2 waterbag (4D ball), 3 hollow
  float sa, sb;          // per-axis scales (aspect ratio)
  float cs, sn;          // rotation
};

// Particle i of pair k, sample `which` (0: low-N, 1: high-N).
__device__ __forceinline__ void particle(const PairSpec& s, uint64_t seed, int k, int which,
                                         int64_t i, float& x, float& y) {
  const uint64_t h0 = mix64(seed ^ mix64(((uint64_t)k << 33) ^ ((uint64_t)which << 32) ^ 0x5bd1e995ull) ^
                            (uint64_t)i * 0x2545F4914F6CDD1Dull);
  const uint64_t h1 = mix64(h0);
  const float u1 = unif(h0), u2 = unif(h1);
  float r, c, sn;
  __sincosf(6.283185307179586f * u2, &sn, &c);
  switch (s.family) {
    case 0: r = sqrtf(-2.0f * __logf(u1)); break;                    // Box-Muller, sigma 1
    case 1: r = 2.0f * sqrtf(u1); break;                             // uniform disk, rms 1
    case 2: r = 2.449489743f * sqrtf(1.0f - sqrtf(1.0f - u1)); break;  // 4D ball, rms 1
    default: r = 1.714985851f * sqrtf(0.36f + 0.64f * u1); break;    // annulus 0.6-1, rms 1
  }
  const float px = r * c * s.sa, py = r * sn * s.sb;
  x = s.cs * px - s.sn * py;
  y = s.sn * px + s.cs * py;
}

// Particles of a pair come from the generator, or (bx != nullptr) are particles 0, stride,
// 2 stride, ... of a bunch's x and y arrays.
struct PairSource {
  const PairSpec* spec;
  uint64_t seed;
  const double* bx;
  const double* by;
};

__device__ __forceinline__ void source_particle(const PairSource& src, const PairSpec& s, int k,
                                                int which, int64_t i, int64_t stride, float& x,
                                                float& y) {
  if (src.bx) {
    x = (float)src.bx[i * stride];
    y = (float)src.by[i * stride];
  } else {
    particle(s, src.seed, k, which, i, x, y);
  }
}

// Extrema of the low-N sample of each pair: grid (pairs, blocks per pair).
__global__ __launch_bounds__(kThreads) void pair_extrema(PairSource src, int64_t n,
                                                         int64_t stride,
                                                         float* __restrict__ partial) {
  const int k = blockIdx.y;
  const PairSpec s = src.bx ? PairSpec{} : src.spec[k];
  float lo[2] = {FLT_MAX, FLT_MAX}, hi[2] = {-FLT_MAX, -FLT_MAX};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    float x, y;
    source_particle(src, s, k, 0, i, stride, x, y);
    lo[0] = fminf(lo[0], x);
    hi[0] = fmaxf(hi[0], x);
    lo[1] = fminf(lo[1], y);
    hi[1] = fmaxf(hi[1], y);
  }
  __shared__ float sh[4][kThreads / 64];
  float v[4] = {lo[0], -hi[0], lo[1], -hi[1]};
  for (int j = 0; j < 4; j++)
    for (int o = 32; o > 0; o >>= 1) v[j] = fminf(v[j], __shfl_xor(v[j], o, 64));
  if ((threadIdx.x & 63) == 0)
    for (int j = 0; j < 4; j++) sh[j][threadIdx.x >> 6] = v[j];
  __syncthreads();
  if (threadIdx.x < 4) {
    float m = sh[threadIdx.x][0];
    for (int w = 1; w < kThreads / 64; w++) m = fminf(m, sh[threadIdx.x][w]);
    partial[((size_t)k * gridDim.x + blockIdx.x) * 4 + threadIdx.x] = m;
  }
}

// Grid of each pair: square box around the low-N extrema (xy ratio 1, as
// SpaceChargeCalc2p5Drb::bunchAnalysis sizes its grid), nodes xMin + i dx, i < H.
__global__ void pair_grid(const float* __restrict__ partial, int nblk, int npairs, int H, int W,
                          float* __restrict__ geo) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= npairs) return;
  float v[4];
  for (int j = 0; j < 4; j++) {
    float m = FLT_MAX;
    for (int b = 0; b < nblk; b++) m = fminf(m, partial[((size_t)k * nblk + b) * 4 + j]);
    v[j] = m;
  }
  const float xmin = v[0], xmax = -v[1], ymin = v[2], ymax = -v[3];
  const float xc = 0.5f * (xmin + xmax), yc = 0.5f * (ymin + ymax);
  const float half = 0.5f * fmaxf(xmax - xmin, ymax - ymin);
  geo[4 * k + 0] = xc - half;
  geo[4 * k + 1] = 2 * half / (H - 1);
  geo[4 * k + 2] = yc - half;
  geo[4 * k + 3] = 2 * half / (W - 1);
}

// TSC deposit (Grid2D.cc:337-350 indices and weights) of n particles of each pair into
// out[k] (H x W, normalized so that the mean over the grid of n/H/W particles per node is
// 1).  The grid is split into row slabs held in LDS; every slab block regenerates the
// particles of its chunk and keeps the contributions that land in its slab.
// grid: (chunks, slabs, pairs).
__global__ __launch_bounds__(kThreads) void pair_deposit(PairSource src, int which, int64_t n,
                                                         int64_t stride,
                                                         const float* __restrict__ geo, int H,
                                                         int W, int slab_rows,
                                                         float* __restrict__ out) {
  extern __shared__ float acc[];
  const int k = blockIdx.z;
  const PairSpec s = src.bx ? PairSpec{} : src.spec[k];
  const float xmin = geo[4 * k], dx = geo[4 * k + 1], ymin = geo[4 * k + 2], dy = geo[4 * k + 3];
  const int r0 = blockIdx.y * slab_rows, r1 = min(H, r0 + slab_rows);
  const int cells = (r1 - r0) * W;
  for (int c = threadIdx.x; c < cells; c += blockDim.x) acc[c] = 0.f;
  __syncthreads();
  const float norm = (float)H * (float)W / (float)n;
  const int64_t per = (n + gridDim.x - 1) / gridDim.x;
  const int64_t i0 = (int64_t)blockIdx.x * per, i1 = min(n, i0 + per);
  for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    float x, y;
    source_particle(src, s, k, which, i, stride, x, y);
    const float fx = (x - xmin) / dx, fy = (y - ymin) / dy;
    if (!(fx >= 0.f && fx <= (float)(H - 1) && fy >= 0.f && fy <= (float)(W - 1))) continue;
    int ix = (int)(fx + 0.5f), iy = (int)(fy + 0.5f);
    ix = min(max(ix, 1), H - 2);
    iy = min(max(iy, 1), W - 2);
    if (ix + 1 < r0 || ix - 1 >= r1) continue;
    const float xf = fx - ix, yf = fy - iy;
    const float wx[3] = {0.5f * (0.5f - xf) * (0.5f - xf), 0.75f - xf * xf,
                         0.5f * (0.5f + xf) * (0.5f + xf)};
    const float wy[3] = {0.5f * (0.5f - yf) * (0.5f - yf), 0.75f - yf * yf,
                         0.5f * (0.5f + yf) * (0.5f + yf)};
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const int row = ix - 1 + a;
      if (row < r0 || row >= r1) continue;
#pragma unroll
      for (int c = 0; c < 3; c++) atomicAdd(&acc[(row - r0) * W + iy - 1 + c], wx[a] * wy[c] * norm);
    }
  }
  __syncthreads();
  float* dst = out + (size_t)k * H * W + (size_t)r0 * W;
  for (int c = threadIdx.x; c < cells; c += blockDim.x)
    if (acc[c] != 0.f) atomicAdd(&dst[c], acc[c]);
}

}  // namespace dn
}  // namespace ob

// ============================================================================================
// Host side
// ============================================================================================
namespace {
using ob::dn::bf16;

struct Layer {
  int Cin, Cout, H;       // H: level (0: full, 1: half, 2: quarter)
  size_t w_off, b_off;    // in the FP32 parameter vector
  size_t pf_off, pt_off;  // in the packed BF16 buffers
};

// The fixed architecture: 12 conv3x3 layers, then the head (16 weights + bias).
const int kLayerShape[12][3] = {{1, 16, 0},  {16, 16, 0}, {16, 32, 1}, {32, 32, 1},
                                {32, 64, 2}, {64, 64, 2}, {64, 64, 2}, {64, 64, 2},
                                {96, 32, 1}, {32, 32, 1}, {48, 16, 0}, {16, 16, 0}};
}  // namespace

struct ob_denoiser {
  int H = 0, W = 0, max_batch = 0;
  Layer L[12];
  size_t nparam = 0, head_off = 0, npf = 0, npt = 0;
  float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  // FP32 params, grads, Adam
  bf16 *pf = nullptr, *pt = nullptr;                              // packed BF16 weights
  int step = 0;
  // activations (BF16 NHWC) and their gradients, sized for max_batch images
  bf16* act[12] = {};
  bf16* gact[12] = {};
  bf16* gcat4 = nullptr;  // gradient of [up(e3), e2]   (96 ch, H/2)
  bf16* gcat5 = nullptr;  // gradient of [up(d2), e1]   (48 ch, H)
  bf16* gtmp0 = nullptr;  // gradient of pooled inputs  (p1: 16 ch H/2, p2: 32 ch H/4)
  bf16* gtmp1 = nullptr;
  float* out = nullptr;   // head output, FP32
  float* loss = nullptr;  // device scalar
  float* wpart = nullptr;  // weight-gradient partials (kWpartFloats)
  float* bpart = nullptr;  // bias-gradient partials (kWgradBlocks * 64)
  float* hpart = nullptr;  // head / loss partials (kHeadBlocks * 18)
  float* pyr = nullptr;   // error pyramid of the multi-scale loss
  int nscales = 0;
  float lam[9] = {};
  float* xin = nullptr;   // normalized rho images (space-charge use), max_batch of them
  double* total = nullptr;  // its total charge
  double* rpart = nullptr;  // per-slice charge sums, kRhoParts per image
};

namespace ob {
namespace dn {
namespace {

int level_dim(const ob_denoiser* d, int lev, bool w) { return (w ? d->W : d->H) >> lev; }

void launch_conv(const ConvArgs& a, int B, hipStream_t s) {
  OB_REQUIRE(a.H % kTile == 0 && a.W % kTile == 0, "denoiser: image sides must be multiples of 64");
  const int tiles = (a.H / kTile) * (a.W / kTile);
  OB_REQUIRE(a.Cout % 16 == 0, "denoiser: output channels must be a multiple of 16");
  // 48 and 96 output channels (the data gradients of the concatenating layers) take three
  // 16-channel tiles per workgroup, so the input tile is staged once or twice, not 3 times
  const int nt = (a.Cout % 64 == 0) ? 4 : (a.Cout % 48 == 0) ? 3 : (a.Cout % 32 == 0) ? 2 : 1;
  const dim3 grid(tiles, a.Cout / (16 * nt), B);
  ProfScope ps("dn_conv");
  if (nt == 4)
    hipLaunchKernelGGL(conv3x3_fwd<4>, grid, dim3(kThreads), 0, s, a);
  else if (nt == 3)
    hipLaunchKernelGGL(conv3x3_fwd<3>, grid, dim3(kThreads), 0, s, a);
  else if (nt == 2)
    hipLaunchKernelGGL(conv3x3_fwd<2>, grid, dim3(kThreads), 0, s, a);
  else
    hipLaunchKernelGGL(conv3x3_fwd<1>, grid, dim3(kThreads), 0, s, a);
}

// Workgroups of the weight-gradient pass: about four per CU in all, split over the chunks
// (measured 1.3% faster training than two).
int wgrad_blocks(int ntiles, int chunks) {
  return std::max(1, std::min(ntiles, (kWgradBlocks + chunks - 1) / chunks));
}

// Chunks per workgroup of the weight-gradient pass: as many as the accumulators allow
// (CH x NP f4 per lane), so dY and Y are read once per CH chunks.  At MT = 2, CH = 3 needs
// 247 registers and measured slower than CH = 2 (241 against 148 us for 6 chunks).
int wgrad_chunks_per_block(int mt, int chunks) {
  if (mt == 1) return chunks % 3 == 0 ? 3 : 1;
  if (mt == 2) return chunks % 2 == 0 ? 2 : 1;
  return 1;
}

void launch_wgrad(const ConvArgs& a, int B, hipStream_t s) {
  const int ntiles = (a.H / kTile) * (a.W / kTile) * B;
  const int chunks = (a.Cin + 15) / 16;
  const int mt = a.Cout / 16;
  const int ch = wgrad_chunks_per_block(mt, chunks);
  const int blocks = wgrad_blocks(ntiles, chunks / ch);
  OB_REQUIRE((size_t)chunks * blocks * 9 * 16 * a.Cout <= kWpartFloats,
             "denoiser: weight-gradient partials exceed their buffer");
  const dim3 grid(blocks, chunks / ch);
  {
    ProfScope ps("dn_wgrad");
    const int sel = mt * 4 + ch;
    switch (sel) {
      case 1 * 4 + 1: hipLaunchKernelGGL((conv3x3_wgrad<1, 1>), grid, dim3(kThreads), 0, s, a, ntiles); break;
      case 1 * 4 + 3: hipLaunchKernelGGL((conv3x3_wgrad<1, 3>), grid, dim3(kThreads), 0, s, a, ntiles); break;
      case 2 * 4 + 1: hipLaunchKernelGGL((conv3x3_wgrad<2, 1>), grid, dim3(kThreads), 0, s, a, ntiles); break;
      case 2 * 4 + 2: hipLaunchKernelGGL((conv3x3_wgrad<2, 2>), grid, dim3(kThreads), 0, s, a, ntiles); break;
      case 4 * 4 + 1: hipLaunchKernelGGL((conv3x3_wgrad<4, 1>), grid, dim3(kThreads), 0, s, a, ntiles); break;
      default: OB_REQUIRE(false, "denoiser: unsupported output channel count");
    }
  }
  ProfScope ps("dn_wgrad_reduce");
  const int ne = 9 * a.Cin * a.Cout + a.Cout;
  hipLaunchKernelGGL(conv3x3_wgrad_reduce, dim3((ne + 63) / 64), dim3(1024), 0, s, a.wpart,
                     a.bpart, blocks, a.Cin, a.Cout, a.dw, a.db);
}

int grid_for(int64_t n) {
  return (int)std::max<int64_t>(1, std::min<int64_t>((n + kThreads - 1) / kThreads, 8 * device_cu_count()));
}

Src src(const bf16* p, int C, int cs = -1, int coff = 0, int mode = SRC_SAME,
        const bf16* mask = nullptr) {
  return Src{p, mask, C, cs < 0 ? C : cs, coff, mode};
}
Src none() { return Src{nullptr, nullptr, 0, 0, 0, 0}; }

// Logical input of layer l in the forward graph.
void layer_inputs(const ob_denoiser* d, int l, const float* x, ConvArgs& a) {
  bf16* const* A = d->act;
  a.xin = nullptr;
  a.s0 = none();
  a.s1 = none();
  switch (l) {
    case 0: a.xin = x; break;
    case 1: a.s0 = src(A[0], 16); break;
    case 2: a.s0 = src(A[1], 16, 16, 0, SRC_POOL); break;
    case 3: a.s0 = src(A[2], 32); break;
    case 4: a.s0 = src(A[3], 32, 32, 0, SRC_POOL); break;
    case 5: a.s0 = src(A[4], 64); break;
    case 6: a.s0 = src(A[5], 64); break;
    case 7: a.s0 = src(A[6], 64); break;
    case 8: a.s0 = src(A[7], 64, 64, 0, SRC_UP); a.s1 = src(A[3], 32); break;
    case 9: a.s0 = src(A[8], 32); break;
    case 10: a.s0 = src(A[9], 32, 32, 0, SRC_UP); a.s1 = src(A[1], 16); break;
    case 11: a.s0 = src(A[10], 16); break;
  }
}

ConvArgs fwd_args(ob_denoiser* d, int l, const float* x) {
  ConvArgs a;
  std::memset(&a, 0, sizeof(a));
  const Layer& L = d->L[l];
  layer_inputs(d, l, x, a);
  a.w = d->pf + L.pf_off;
  a.bias = d->w + L.b_off;
  a.Cin = L.Cin;
  a.Cout = L.Cout;
  a.H = level_dim(d, L.H, false);
  a.W = level_dim(d, L.H, true);
  a.relu = 1;
  a.out = d->act[l];
  a.ocs = L.Cout;
  a.ocoff = 0;
  return a;
}

void forward(ob_denoiser* d, const float* x, int B, hipStream_t s) {
  for (int l = 0; l < 12; l++) launch_conv(fwd_args(d, l, x), B, s);
  const int64_t npix = (int64_t)B * d->H * d->W;
  ProfScope ps("dn_head");
  hipLaunchKernelGGL(head_fwd, dim3(grid_for(npix)), dim3(kThreads), 0, s, d->act[11], x,
                     d->w + d->head_off, d->out, npix);
}

// Backward of the whole net for the loss against `target`; gradients accumulate into d->g
// (cleared here).
void backward(ob_denoiser* d, const float* x, const float* target, int B, hipStream_t s) {
  const int64_t npix = (int64_t)B * d->H * d->W;
  hipLaunchKernelGGL(fill_f32, dim3(grid_for(d->nparam)), dim3(kThreads), 0, s, d->g,
                     (int64_t)d->nparam, 0.f);
  hipLaunchKernelGGL(fill_f32, dim3(1), dim3(kThreads), 0, s, d->loss, (int64_t)1, 0.f);
  Pyramid P;
  std::memset(&P, 0, sizeof(P));
  P.S = d->nscales;
  P.H = d->H;
  P.W = d->W;
  size_t off = 0;
  for (int sc = 1; sc <= P.S; sc++) {
    P.m[sc] = d->pyr + off;
    P.lam[sc] = d->lam[sc];
    off += (size_t)B * (d->H >> sc) * (d->W >> sc);
  }
  for (int sc = 1; sc <= P.S; sc++) {
    ProfScope ps("dn_loss_pyramid");
    const int64_t n = (int64_t)B * (d->H >> sc) * (d->W >> sc);
    hipLaunchKernelGGL(pyramid_level, dim3(grid_for(n)), dim3(kThreads), 0, s, d->out, target, P,
                       sc, B, d->loss);
  }
  {
    ProfScope ps("dn_head_bwd");
    const int hb = (int)std::min<int64_t>(kHeadBlocks, (npix + kThreads - 1) / kThreads);
    hipLaunchKernelGGL(head_bwd, dim3(hb), dim3(kThreads), 0, s, d->act[11], d->out, target,
                       d->w + d->head_off, d->gact[11], d->hpart, npix, P);
    hipLaunchKernelGGL(head_reduce, dim3(18), dim3(64), 0, s, d->hpart, hb, d->g + d->head_off,
                       d->loss);
  }
  // gradient w.r.t. each layer's output arrives in gact[l] (before its ReLU mask)
  for (int l = 11; l >= 0; l--) {
    const Layer& L = d->L[l];
    const int H = level_dim(d, L.H, false), W = level_dim(d, L.H, true);
    // weight gradient
    ConvArgs a;
    std::memset(&a, 0, sizeof(a));
    layer_inputs(d, l, x, a);
    a.Cin = L.Cin;
    a.Cout = L.Cout;
    a.H = H;
    a.W = W;
    a.dy = d->gact[l];
    a.y = d->act[l];
    a.dw = d->g + L.w_off;
    a.db = d->g + L.b_off;
    a.wpart = d->wpart;
    a.bpart = d->bpart;
    launch_wgrad(a, B, s);
    if (l == 0) break;
    // data gradient: conv of dY (masked by Y > 0) with the flipped transposed kernel
    ConvArgs b;
    std::memset(&b, 0, sizeof(b));
    b.s0 = src(d->gact[l], L.Cout, L.Cout, 0, SRC_SAME, d->act[l]);
    b.w = d->pt + L.pt_off;
    b.Cin = L.Cout;
    b.Cout = L.Cin;
    b.H = H;
    b.W = W;
    b.relu = 0;
    b.ocoff = 0;
    b.ocs = L.Cin;
    switch (l) {
      case 1: case 3: case 5: case 6: case 7: case 9: case 11:
        b.out = d->gact[l - 1];  // plain chain
        break;
      case 2: b.out = d->gtmp0; break;   // grad of pool(e1): 16 ch, H/2
      case 4: b.out = d->gtmp1; break;   // grad of pool(e2): 32 ch, H/4
      case 8: b.out = d->gcat4; break;   // grad of [up(e3), e2]: 96 ch, H/2
      case 10: b.out = d->gcat5; break;  // grad of [up(d2), e1]: 48 ch, H
    }
    launch_conv(b, B, s);
    ProfScope ps("dn_grad_route");
    switch (l) {
      case 10: {  // d2 <- up, e1 skip kept in gcat5[32:48]
        const int h = d->H / 2, w = d->W / 2;
        const int64_t t8 = (int64_t)B * h * w * 32 / 8;
        hipLaunchKernelGGL(grad_up, dim3(grid_for(t8)), dim3(kThreads), 0, s, d->gcat5, 48,
                           d->gact[9], 32, h, w, t8);
        break;
      }
      case 8: {  // e3 <- up, e2 skip kept in gcat4[64:96]
        const int h = d->H / 4, w = d->W / 4;
        const int64_t t8 = (int64_t)B * h * w * 64 / 8;
        hipLaunchKernelGGL(grad_up, dim3(grid_for(t8)), dim3(kThreads), 0, s, d->gcat4, 96,
                           d->gact[7], 64, h, w, t8);
        break;
      }
      case 4: {  // e2 = pool grad + skip grad
        const int h = d->H / 2, w = d->W / 2;
        const int64_t t8 = (int64_t)B * h * w * 32 / 8;
        hipLaunchKernelGGL(grad_pool_skip, dim3(grid_for(t8)), dim3(kThreads), 0, s, d->gtmp1,
                           d->gcat4, 96, 64, d->gact[3], 32, h, w, t8);
        break;
      }
      case 2: {  // e1 = pool grad + skip grad
        const int h = d->H, w = d->W;
        const int64_t t8 = (int64_t)B * h * w * 16 / 8;
        hipLaunchKernelGGL(grad_pool_skip, dim3(grid_for(t8)), dim3(kThreads), 0, s, d->gtmp0,
                           d->gcat5, 48, 32, d->gact[1], 16, h, w, t8);
        break;
      }
      default: break;
    }
  }
}

void pack_all(ob_denoiser* d, hipStream_t s) {
  ProfScope ps("dn_pack");
  PackTable t;
  for (int l = 0; l < 12; l++) {
    const Layer& L = d->L[l];
    t.w_off[l] = (int32_t)L.w_off;
    t.cin[l] = L.Cin;
    t.cout[l] = L.Cout;
    t.pf_off[l] = (int64_t)L.pf_off;
    t.pt_off[l] = (int64_t)L.pt_off;
  }
  hipLaunchKernelGGL(pack_layers, dim3(64, 12), dim3(kThreads), 0, s, d->w, d->pf, d->pt, t);
}

// rho grids -> normalised images -> network -> rho grids with their deposited charge.
void denoise_grids(ob_denoiser* d, const RhoJobs& rj, int count, hipStream_t s) {
  const int n = d->H * d->W;
  const dim3 grid(kRhoParts, count);
  hipLaunchKernelGGL(rho_parts, grid, dim3(kRhoThreads), 0, s, rj, n, d->rpart);
  hipLaunchKernelGGL(rho_images, grid, dim3(kRhoThreads), 0, s, rj, n, d->rpart, d->xin,
                     d->total);
  forward(d, d->xin, count, s);
  hipLaunchKernelGGL(image_parts, grid, dim3(kRhoThreads), 0, s, d->out, n, rj, d->rpart);
  hipLaunchKernelGGL(images_rho, grid, dim3(kRhoThreads), 0, s, d->out, n, rj, d->rpart,
                     d->total);
  OB_HIP(hipGetLastError());
}

void check_batch(ob_denoiser* d, int B) {
  OB_REQUIRE(B >= 1 && B <= d->max_batch, "denoiser: batch size out of range");
}

}  // namespace
}  // namespace dn

void dn_check_shape(ob_denoiser* d, int nx, int ny) {
  OB_REQUIRE(d->H == nx && d->W == ny, "denoiser: image size differs from the grid");
}

void dn_denoise_rhos(const std::vector<DenoiseJob>& jobs) {
  hipStream_t s = stream();
  size_t a = 0;
  while (a < jobs.size()) {
    ob_denoiser* d = jobs[a].d;
    dn_check_shape(d, jobs[a].nx, jobs[a].ny);
    const int cap = std::min(64, d->max_batch);
    dn::RhoJobs rj;
    int count = 0;
    while (a < jobs.size() && jobs[a].d == d && count < cap) {
      rj.rho[count] = jobs[a].rho;
      rj.skip[count] = jobs[a].skip;
      count++;
      a++;
    }
    ProfScope ps("dn_sc_denoise_batch");
    dn::denoise_grids(d, rj, count, s);
  }
}

void dn_denoise_rho(ob_denoiser* d, double* rho, int nx, int ny, const int32_t* skip) {
  dn_check_shape(d, nx, ny);
  hipStream_t s = stream();
  ProfScope ps("dn_sc_denoise");
  dn::RhoJobs rj;
  rj.rho[0] = rho;
  rj.skip[0] = skip;
  dn::denoise_grids(d, rj, 1, s);
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

OB_API int ob_dn_create(int H, int W, int max_batch, ob_denoiser** out) {
  OB_TRY OB_REQUIRE(out != nullptr, "null output pointer");
  OB_REQUIRE(H >= 64 && W >= 64 && H % 64 == 0 && W % 64 == 0 && H <= 4096 && W <= 4096,
             "denoiser: image sides must be multiples of 64 in [64, 4096]");
  OB_REQUIRE(max_batch >= 1, "denoiser: max_batch must be >= 1");
  ob::ensure_init();
  ob_denoiser* d = new ob_denoiser();
  try {
    d->H = H;
    d->W = W;
    d->max_batch = max_batch;
    size_t off = 0, pf = 0, pt = 0;
    for (int l = 0; l < 12; l++) {
      Layer& L = d->L[l];
      L.Cin = kLayerShape[l][0];
      L.Cout = kLayerShape[l][1];
      L.H = kLayerShape[l][2];
      L.w_off = off;
      off += (size_t)9 * L.Cin * L.Cout;
      L.b_off = off;
      off += L.Cout;
      L.pf_off = pf;
      pf += (size_t)9 * ((L.Cin + 15) / 16 * 16) * L.Cout;
      L.pt_off = pt;
      pt += (size_t)9 * L.Cin * L.Cout;
    }
    d->head_off = off;
    off += 17;
    d->nparam = off;
    d->npf = pf;
    d->npt = pt;
    OB_HIP(hipMalloc((void**)&d->w, sizeof(float) * off));
    OB_HIP(hipMalloc((void**)&d->g, sizeof(float) * off));
    OB_HIP(hipMalloc((void**)&d->m, sizeof(float) * off));
    OB_HIP(hipMalloc((void**)&d->v, sizeof(float) * off));
    OB_HIP(hipMemset(d->w, 0, sizeof(float) * off));
    OB_HIP(hipMemset(d->g, 0, sizeof(float) * off));
    OB_HIP(hipMemset(d->m, 0, sizeof(float) * off));
    OB_HIP(hipMemset(d->v, 0, sizeof(float) * off));
    OB_HIP(hipMalloc((void**)&d->pf, sizeof(bf16) * pf));
    OB_HIP(hipMalloc((void**)&d->pt, sizeof(bf16) * pt));
    const size_t img = (size_t)max_batch * H * W;
    for (int l = 0; l < 12; l++) {
      const size_t n = img * d->L[l].Cout >> (2 * d->L[l].H);
      OB_HIP(hipMalloc((void**)&d->act[l], sizeof(bf16) * n));
      OB_HIP(hipMalloc((void**)&d->gact[l], sizeof(bf16) * n));
    }
    OB_HIP(hipMalloc((void**)&d->gcat4, sizeof(bf16) * (img / 4) * 96));
    OB_HIP(hipMalloc((void**)&d->gcat5, sizeof(bf16) * img * 48));
    OB_HIP(hipMalloc((void**)&d->gtmp0, sizeof(bf16) * (img / 4) * 16));
    OB_HIP(hipMalloc((void**)&d->gtmp1, sizeof(bf16) * (img / 16) * 32));
    OB_HIP(hipMalloc((void**)&d->out, sizeof(float) * img));
    OB_HIP(hipMalloc((void**)&d->loss, sizeof(float)));
    OB_HIP(hipMalloc((void**)&d->xin, sizeof(float) * img));
    OB_HIP(hipMalloc((void**)&d->pyr, sizeof(float) * (img / 3 + 1)));
    OB_HIP(hipMalloc((void**)&d->wpart, sizeof(float) * ob::dn::kWpartFloats));
    OB_HIP(hipMalloc((void**)&d->bpart, sizeof(float) * ob::dn::kWgradBlocks * 64));
    OB_HIP(hipMalloc((void**)&d->hpart, sizeof(float) * ob::dn::kHeadBlocks * 18));
    OB_HIP(hipMalloc((void**)&d->total, sizeof(double) * max_batch));
    OB_HIP(hipMalloc((void**)&d->rpart, sizeof(double) * ob::dn::kRhoParts * max_batch));
  } catch (...) {
    delete d;  // device buffers of a failed create are reclaimed at process exit
    throw;
  }
  *out = d;
  OB_CATCH
}

OB_API int ob_dn_destroy(ob_denoiser* d) {
  OB_TRY if (!d) return 0;
  (void)hipStreamSynchronize(ob::stream());
  for (void* p : {(void*)d->w, (void*)d->g, (void*)d->m, (void*)d->v, (void*)d->pf, (void*)d->pt,
                  (void*)d->gcat4, (void*)d->gcat5, (void*)d->gtmp0, (void*)d->gtmp1,
                  (void*)d->out, (void*)d->loss, (void*)d->xin, (void*)d->total, (void*)d->rpart,
                  (void*)d->pyr, (void*)d->wpart, (void*)d->bpart, (void*)d->hpart})
    if (p) (void)hipFree(p);
  for (int l = 0; l < 12; l++) {
    if (d->act[l]) (void)hipFree(d->act[l]);
    if (d->gact[l]) (void)hipFree(d->gact[l]);
  }
  delete d;
  OB_CATCH
}

OB_API int ob_dn_num_params(ob_denoiser* d, int64_t* n) {
  OB_TRY OB_REQUIRE(d != nullptr && n != nullptr, "null handle");
  *n = (int64_t)d->nparam;
  OB_CATCH
}

// Parameters as one FP32 vector: per layer W[9][Cin][Cout] then b[Cout]; then the head's
// 16 weights and its bias.  Setting them resets the optimizer state.
OB_API int ob_dn_set_params(ob_denoiser* d, const float* host, int64_t n) {
  OB_TRY OB_REQUIRE(d != nullptr && host != nullptr, "null handle");
  OB_REQUIRE(n == (int64_t)d->nparam, "denoiser: wrong parameter count");
  hipStream_t s = ob::stream();
  OB_HIP(hipMemcpyAsync(d->w, host, sizeof(float) * n, hipMemcpyHostToDevice, s));
  OB_HIP(hipMemsetAsync(d->m, 0, sizeof(float) * n, s));
  OB_HIP(hipMemsetAsync(d->v, 0, sizeof(float) * n, s));
  d->step = 0;
  ob::dn::pack_all(d, s);
  OB_HIP(hipStreamSynchronize(s));
  OB_CATCH
}

OB_API int ob_dn_get_params(ob_denoiser* d, float* host, int64_t n) {
  OB_TRY OB_REQUIRE(d != nullptr && host != nullptr, "null handle");
  OB_REQUIRE(n == (int64_t)d->nparam, "denoiser: wrong parameter count");
  OB_HIP(hipMemcpyAsync(host, d->w, sizeof(float) * n, hipMemcpyDeviceToHost, ob::stream()));
  OB_HIP(hipStreamSynchronize(ob::stream()));
  OB_CATCH
}

// Inference on B images resident on the device (FP32, [B][H][W]); y may equal x.
OB_API int ob_dn_forward_dev(ob_denoiser* d, const float* x_dev, float* y_dev, int B) {
  OB_TRY OB_REQUIRE(d != nullptr && x_dev != nullptr && y_dev != nullptr, "null handle");
  ob::dn::check_batch(d, B);
  hipStream_t s = ob::stream();
  ob::dn::forward(d, x_dev, B, s);
  OB_HIP(hipMemcpyAsync(y_dev, d->out, sizeof(float) * B * d->H * d->W, hipMemcpyDeviceToDevice, s));
  OB_HIP(hipGetLastError());
  OB_CATCH
}

// Host-buffer variants (tests, small problems).
OB_API int ob_dn_forward(ob_denoiser* d, const float* x, float* y, int B) {
  OB_TRY OB_REQUIRE(d != nullptr && x != nullptr && y != nullptr, "null handle");
  ob::dn::check_batch(d, B);
  hipStream_t s = ob::stream();
  const size_t n = (size_t)B * d->H * d->W;
  float* xd = nullptr;
  OB_HIP(hipMalloc((void**)&xd, sizeof(float) * n));
  try {
    OB_HIP(hipMemcpyAsync(xd, x, sizeof(float) * n, hipMemcpyHostToDevice, s));
    ob::dn::forward(d, xd, B, s);
    OB_HIP(hipMemcpyAsync(y, d->out, sizeof(float) * n, hipMemcpyDeviceToHost, s));
    OB_HIP(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipFree(xd);
    throw;
  }
  OB_HIP(hipFree(xd));
  OB_CATCH
}

// Loss (MSE) and the gradient of every parameter (FP32, layout of ob_dn_set_params) for
// B host images and targets.  Leaves the parameters unchanged.
OB_API int ob_dn_loss_grad(ob_denoiser* d, const float* x, const float* target, int B,
                           float* loss, float* grad) {
  OB_TRY OB_REQUIRE(d != nullptr && x != nullptr && target != nullptr, "null handle");
  ob::dn::check_batch(d, B);
  hipStream_t s = ob::stream();
  const size_t n = (size_t)B * d->H * d->W;
  float* buf = nullptr;
  OB_HIP(hipMalloc((void**)&buf, sizeof(float) * 2 * n));
  try {
    OB_HIP(hipMemcpyAsync(buf, x, sizeof(float) * n, hipMemcpyHostToDevice, s));
    OB_HIP(hipMemcpyAsync(buf + n, target, sizeof(float) * n, hipMemcpyHostToDevice, s));
    ob::dn::forward(d, buf, B, s);
    ob::dn::backward(d, buf, buf + n, B, s);
    if (loss) OB_HIP(hipMemcpyAsync(loss, d->loss, sizeof(float), hipMemcpyDeviceToHost, s));
    if (grad) OB_HIP(hipMemcpyAsync(grad, d->g, sizeof(float) * d->nparam, hipMemcpyDeviceToHost, s));
    OB_HIP(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipFree(buf);
    throw;
  }
  OB_HIP(hipFree(buf));
  OB_CATCH
}

// One Adam step on B device-resident pairs (x_dev, target_dev: FP32 [B][H][W]).  The loss
// of the step stays on the device until ob_dn_last_loss.
OB_API int ob_dn_train_step_dev(ob_denoiser* d, const float* x_dev, const float* target_dev, int B,
                                float lr) {
  OB_TRY OB_REQUIRE(d != nullptr && x_dev != nullptr && target_dev != nullptr, "null handle");
  ob::dn::check_batch(d, B);
  OB_REQUIRE(lr > 0.f, "denoiser: learning rate must be positive");
  hipStream_t s = ob::stream();
  ob::dn::forward(d, x_dev, B, s);
  ob::dn::backward(d, x_dev, target_dev, B, s);
  d->step++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f / (1.0f - powf(b1, (float)d->step));
  const float c2 = 1.0f / (1.0f - powf(b2, (float)d->step));
  {
    ob::ProfScope ps("dn_adam");
    hipLaunchKernelGGL(ob::dn::adam_step, dim3(ob::dn::grid_for(d->nparam)), dim3(ob::dn::kThreads),
                       0, s, d->w, d->g, d->m, d->v, (int64_t)d->nparam, lr, b1, b2, 1e-8f, c1, c2);
  }
  ob::dn::pack_all(d, s);
  OB_HIP(hipGetLastError());
  OB_CATCH
}

// Loss = mean (out - y)^2 + sum_{s=1..nscales} lam[s-1] * mean over 2^s x 2^s blocks of
// (block mean of out - y)^2.  nscales = 0: plain MSE.
OB_API int ob_dn_set_loss(ob_denoiser* d, int nscales, const float* lam) {
  OB_TRY OB_REQUIRE(d != nullptr, "null handle");
  OB_REQUIRE(nscales >= 0 && nscales <= ob::dn::kMaxScales && (nscales == 0 || lam != nullptr),
             "denoiser: 0..8 loss scales");
  OB_REQUIRE((d->H >> nscales) >= 1 && (d->W >> nscales) >= 1 && (d->H >> nscales) << nscales == d->H &&
                 (d->W >> nscales) << nscales == d->W,
             "denoiser: the image is not divisible into 2^nscales blocks");
  d->nscales = nscales;
  for (int sc = 1; sc <= nscales; sc++) d->lam[sc] = lam[sc - 1];
  OB_CATCH
}

OB_API int ob_dn_last_loss(ob_denoiser* d, float* loss) {
  OB_TRY OB_REQUIRE(d != nullptr && loss != nullptr, "null handle");
  OB_HIP(hipMemcpyAsync(loss, d->loss, sizeof(float), hipMemcpyDeviceToHost, ob::stream()));
  OB_HIP(hipStreamSynchronize(ob::stream()));
  OB_CATCH
}

// ---- training pairs --------------------------------------------------------------------
// specs: per pair {family, scale_a, scale_b, cos, sin} as 5 floats (family as a float).
// Writes x[k] (n_low particles) and y[k] (n_high particles) deposited on the grid of the
// low-N sample, both normalized to mean 1 over the H x W grid.  x, y: device buffers of
// npairs * H * W floats (may be allocated with ob_dn_alloc).
namespace {
// Extrema -> square grid -> the two deposits, for npairs pairs from `src`.
void make_pairs(int H, int W, int npairs, const ob::dn::PairSource& src, int64_t n_low,
                int64_t low_stride, int64_t n_high, float* x_dev, float* y_dev) {
  hipStream_t s = ob::stream();
  const int nblk = 64;
  float *part = nullptr, *geo = nullptr;
  OB_HIP(hipMalloc((void**)&part, sizeof(float) * 4 * nblk * npairs));
  try {
    OB_HIP(hipMalloc((void**)&geo, sizeof(float) * 4 * npairs));
    hipLaunchKernelGGL(ob::dn::pair_extrema, dim3(nblk, npairs), dim3(ob::dn::kThreads), 0, s,
                       src, n_low, low_stride, part);
    hipLaunchKernelGGL(ob::dn::pair_grid, dim3((npairs + 63) / 64), dim3(64), 0, s, part, nblk,
                       npairs, H, W, geo);
    const size_t n = (size_t)npairs * H * W;
    OB_HIP(hipMemsetAsync(x_dev, 0, sizeof(float) * n, s));
    OB_HIP(hipMemsetAsync(y_dev, 0, sizeof(float) * n, s));
    const int slab_rows = std::max(1, std::min(H, (int)(48 * 1024 / (sizeof(float) * W))));
    const int slabs = (H + slab_rows - 1) / slab_rows;
    const size_t lds = sizeof(float) * slab_rows * W;
    for (int which = 0; which < 2; which++) {
      const int64_t np = which ? n_high : n_low;
      const int64_t stride = which ? 1 : low_stride;
      const int chunks = (int)std::max<int64_t>(1, std::min<int64_t>(np / 16384, 256));
      hipLaunchKernelGGL(ob::dn::pair_deposit, dim3(chunks, slabs, npairs),
                         dim3(ob::dn::kThreads), lds, s, src, which, np, stride, geo, H, W,
                         slab_rows, which ? y_dev : x_dev);
    }
    OB_HIP(hipGetLastError());
    OB_HIP(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipFree(part);
    (void)hipFree(geo);
    throw;
  }
  (void)hipFree(part);
  (void)hipFree(geo);
}
}  // namespace

OB_API int ob_dn_make_pairs(int H, int W, int npairs, const float* specs, uint64_t seed,
                            int64_t n_low, int64_t n_high, float* x_dev, float* y_dev) {
  OB_TRY OB_REQUIRE(specs != nullptr && x_dev != nullptr && y_dev != nullptr, "null pointer");
  OB_REQUIRE(H >= 4 && W >= 4 && npairs >= 1 && n_low >= 2 && n_high >= 1, "invalid pair sizes");
  ob::ensure_init();
  std::vector<ob::dn::PairSpec> hs(npairs);
  for (int k = 0; k < npairs; k++) {
    const float* p = specs + 5 * k;
    hs[k] = ob::dn::PairSpec{(int32_t)p[0], p[1], p[2], p[3], p[4]};
    OB_REQUIRE(hs[k].family >= 0 && hs[k].family <= 3, "invalid distribution family");
  }
  ob::dn::PairSpec* dspec = nullptr;
  OB_HIP(hipMalloc((void**)&dspec, sizeof(ob::dn::PairSpec) * npairs));
  try {
    OB_HIP(hipMemcpy(dspec, hs.data(), sizeof(ob::dn::PairSpec) * npairs, hipMemcpyHostToDevice));
    make_pairs(H, W, npairs, ob::dn::PairSource{dspec, seed, nullptr, nullptr}, n_low, 1, n_high,
               x_dev, y_dev);
  } catch (...) {
    (void)hipFree(dspec);
    throw;
  }
  (void)hipFree(dspec);
  OB_CATCH
}

// One training pair from a bunch: x_dev gets particles 0, stride, 2 stride, ... (the
// low-N sample), y_dev all particles, both on the square grid around the low-N sample
// (as the generator does).  Pending maps of the bunch are applied first.
OB_API int ob_dn_pair_from_bunch(int H, int W, ob_bunch* b, int64_t stride, float* x_dev,
                                 float* y_dev) {
  OB_TRY OB_REQUIRE(b != nullptr && x_dev != nullptr && y_dev != nullptr, "null pointer");
  OB_REQUIRE(H >= 4 && W >= 4 && stride >= 1, "invalid pair sizes");
  ob::bunch_flush(b, nullptr);
  const ob::DeviceBunchView v = ob::bunch_device_view(b);
  const int64_t n_low = (v.n + stride - 1) / stride;
  OB_REQUIRE(n_low >= 2, "denoiser: the low-N sample needs at least 2 particles");
  make_pairs(H, W, 1, ob::dn::PairSource{nullptr, 0, v.c[0], v.c[2]}, n_low, stride, v.n,
             x_dev, y_dev);
  OB_CATCH
}

// dst image j <- src image idx[j], j < count <= 64, images of n floats (device buffers);
// src holds nsrc images.
OB_API int ob_dn_gather(const float* src, int64_t nsrc, float* dst, int64_t n, const int* idx,
                        int count) {
  OB_TRY OB_REQUIRE(src != nullptr && dst != nullptr && idx != nullptr, "null pointer");
  OB_REQUIRE(count >= 1 && count <= 64 && n >= 1, "gather: 1..64 images");
  ob::dn::GatherIdx g;
  for (int j = 0; j < count; j++) {
    OB_REQUIRE(idx[j] >= 0 && idx[j] < nsrc, "gather: image index out of range");
    g.idx[j] = idx[j];
  }
  hipLaunchKernelGGL(ob::dn::gather_images, dim3(std::max(1, (int)std::min<int64_t>(64, n / 4096 + 1)), count),
                     dim3(ob::dn::kThreads), 0, ob::stream(), src, dst, n, g);
  OB_HIP(hipGetLastError());
  OB_CATCH
}

// Device float buffers for datasets (freed with ob_dn_free).
OB_API int ob_dn_alloc(int64_t nfloats, float** out) {
  OB_TRY OB_REQUIRE(out != nullptr && nfloats > 0, "invalid allocation");
  ob::ensure_init();
  OB_HIP(hipMalloc((void**)out, sizeof(float) * nfloats));
  OB_CATCH
}
OB_API int ob_dn_free(float* p) {
  OB_TRY if (p) {
    OB_HIP(hipStreamSynchronize(ob::stream()));
    OB_HIP(hipFree(p));
  }
  OB_CATCH
}
OB_API int ob_dn_copy(float* dst, const float* src, int64_t nfloats, int kind) {
  OB_TRY OB_REQUIRE(dst != nullptr && src != nullptr && nfloats >= 0, "invalid copy");
  OB_REQUIRE(kind >= 0 && kind <= 2, "copy kind: 0 host->device, 1 device->host, 2 device->device");
  const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice
                          : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  OB_HIP(hipMemcpyAsync(dst, src, sizeof(float) * nfloats, k, ob::stream()));
  OB_HIP(hipStreamSynchronize(ob::stream()));
  OB_CATCH
}
