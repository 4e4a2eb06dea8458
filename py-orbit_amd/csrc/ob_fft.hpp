// ob_fft.hpp -- FP64 complex FFTs staged in LDS, used by the Hockney Poisson solvers.
//
// A workgroup transforms `lines` lines of power-of-two length N held in LDS as split
// re/im arrays (line l starts at l*ls; fft_line_stride(N), 1 modulo 32 doubles, spreads
// lines over the banks).  The transform is a Stockham autosort FFT: stages of radix 8 (4
// or 2 for the remaining factor of N), natural order in and out, so forward -> multiply
// -> inverse needs no permutation.  A stage reads each radix-R butterfly into registers,
// applies the twiddles and an R-point DFT, and after a barrier writes it back in place;
// for N >= 128 the exchanges between stages are padded against bank conflicts (a line
// then spans fft_line(N) > N doubles).  Conventions follow FFTW: forward exponent -1,
// unnormalised.
//
// Twiddles come from a table tw[q] = exp(-2 pi i q / nmax), q < nmax, computed on the
// host in long double from the exact rational angle and rounded once; a butterfly loads
// its first twiddle w and forms w^2 .. w^(R-1) by complex products.
#pragma once

#include <type_traits>

#include "ob_internal.hpp"

namespace ob {

__host__ __device__ constexpr int clog2(int n) { return n <= 1 ? 0 : 1 + clog2(n / 2); }

// Stage radices for N = 2^k: 8s, with the remaining factor 2^(k mod 3) taken first as a 4
// (k mod 3 = 2), two 4s in place of an 8 and a 2 (k mod 3 = 1, k >= 4), or a 2 (N = 2).
__host__ __device__ constexpr int fft_stages(int n) {
  return clog2(n) % 3 == 0 ? clog2(n) / 3
                           : (clog2(n) % 3 == 2 ? 1 + clog2(n) / 3
                                                : (clog2(n) == 1 ? 1 : 2 + (clog2(n) - 4) / 3));
}
__host__ __device__ constexpr int fft_radix(int n, int s) {
  return clog2(n) % 3 == 0 ? 8
                           : (clog2(n) % 3 == 2 ? (s == 0 ? 4 : 8)
                                                : (clog2(n) == 1 ? 2 : (s < 2 ? 4 : 8)));
}
__host__ __device__ constexpr int fft_span(int n, int s) {  // product of the radices before s
  return s == 0 ? 1 : fft_span(n, s - 1) * fft_radix(n, s - 1);
}

// In-register R-point DFTs, natural order; SIGN -1 forward, +1 inverse.  The FFT arithmetic
// may fuse multiply-adds (the solver is checked to a relative tolerance, not bit for bit;
// the rest of the library is built without contraction).
template <int SIGN>
__device__ __forceinline__ void dft2(double* r, double* i) {
#pragma clang fp contract(fast)
  const double ar = r[0], ai = i[0];
  r[0] = ar + r[1];
  i[0] = ai + i[1];
  r[1] = ar - r[1];
  i[1] = ai - i[1];
}

template <int SIGN>
__device__ __forceinline__ void dft4(double* r, double* i) {
#pragma clang fp contract(fast)
  const double t0r = r[0] + r[2], t0i = i[0] + i[2];
  const double t1r = r[0] - r[2], t1i = i[0] - i[2];
  const double t2r = r[1] + r[3], t2i = i[1] + i[3];
  const double dr = r[1] - r[3], di = i[1] - i[3];
  // (x1 - x3) * exp(SIGN i pi / 2) = (x1 - x3) * SIGN i
  const double t3r = SIGN > 0 ? -di : di, t3i = SIGN > 0 ? dr : -dr;
  r[0] = t0r + t2r;
  i[0] = t0i + t2i;
  r[2] = t0r - t2r;
  i[2] = t0i - t2i;
  r[1] = t1r + t3r;
  i[1] = t1i + t3i;
  r[3] = t1r - t3r;
  i[3] = t1i - t3i;
}

template <int SIGN>
__device__ __forceinline__ void dft8(double* r, double* i) {
#pragma clang fp contract(fast)
  double er[4] = {r[0], r[2], r[4], r[6]}, ei[4] = {i[0], i[2], i[4], i[6]};
  double orr[4] = {r[1], r[3], r[5], r[7]}, oi[4] = {i[1], i[3], i[5], i[7]};
  dft4<SIGN>(er, ei);
  dft4<SIGN>(orr, oi);
  constexpr double h = 0.70710678118654752440;  // sqrt(1/2)
  constexpr double s = SIGN > 0 ? 1.0 : -1.0;
  // X[k] = E[k] + W^k O[k], X[k+4] = E[k] - W^k O[k], W = exp(SIGN 2 pi i / 8)
  double xr, xi;
  r[0] = er[0] + orr[0];
  i[0] = ei[0] + oi[0];
  r[4] = er[0] - orr[0];
  i[4] = ei[0] - oi[0];
  xr = h * orr[1] - s * h * oi[1];  // W^1 = h + s h i
  xi = s * h * orr[1] + h * oi[1];
  r[1] = er[1] + xr;
  i[1] = ei[1] + xi;
  r[5] = er[1] - xr;
  i[5] = ei[1] - xi;
  xr = -s * oi[2];                  // W^2 = s i
  xi = s * orr[2];
  r[2] = er[2] + xr;
  i[2] = ei[2] + xi;
  r[6] = er[2] - xr;
  i[6] = ei[2] - xi;
  xr = -h * orr[3] - s * h * oi[3];  // W^3 = -h + s h i
  xi = s * h * orr[3] - h * oi[3];
  r[3] = er[3] + xr;
  i[3] = ei[3] + xi;
  r[7] = er[3] - xr;
  i[7] = ei[3] - xi;
}

template <int R, int SIGN>
__device__ __forceinline__ void dft(double* r, double* i) {
  if constexpr (R == 8)
    dft8<SIGN>(r, i);
  else if constexpr (R == 4)
    dft4<SIGN>(r, i);
  else
    dft2<SIGN>(r, i);
}

// v[m] *= w^m (m = 1..R-1) with w = tw[i1] (conjugated for SIGN > 0): one table load per
// butterfly, the higher powers by complex products (a few ulp from the table's values).
template <int R, int SIGN>
__device__ __forceinline__ void twiddle_powers(double* vr, double* vi,
                                               const double2* __restrict__ tw, int i1) {
#pragma clang fp contract(fast)
  const double2 w = tw[i1];
  const double wr = w.x, wi = SIGN < 0 ? w.y : -w.y;
  double pr = wr, pi = wi;
#pragma unroll
  for (int m = 1; m < R; m++) {
    const double a = vr[m], c = vi[m];
    vr[m] = a * pr - c * pi;
    vi[m] = a * pi + c * pr;
    if (m + 1 < R) {
      const double t = pr * wr - pi * wi;
      pi = pr * wi + pi * wr;
      pr = t;
    }
  }
}

// Exchange layouts of lds_fft.  Exchange e holds the data between stage e-1 and stage e
// (exchange 0 is the input, exchange S the output).  Stage e-1 writes groups of
// Ns = span(e-1) consecutive points every Ns*R points, j-fastest; for Ns < 32 the groups
// written by one half-wave (32 lanes of doubles cover the 64 four-byte banks once) would
// share banks.  For N >= 128 such an exchange inserts Ns doubles of padding after every
// Q = max(Ns*R, 32) points.  The 32/Ns groups of a half-wave then start, modulo 32
// doubles, at distinct multiples of Ns (group g = a*(Q/(Ns*R)) + b, b < Q/(Ns*R), starts
// at a*Q + Ns*(b*R + a)), so its 32 lanes use 32 distinct bank pairs.  Stage e reads 32
// consecutive points starting at a multiple of 32, which stay inside one padded block.
// The input and output keep element i of line l at l*ls + i; a line needs fft_line(N)
// doubles.
__host__ __device__ constexpr int fft_pad_p(int n, int e) {
  return (n < 128 || e == 0 || e >= fft_stages(n) || fft_span(n, e - 1) >= 32)
             ? 0
             : fft_span(n, e - 1);
}
__host__ __device__ constexpr int fft_pad_q(int n, int e) {
  return fft_span(n, e) > 32 ? fft_span(n, e) : 32;
}
__host__ __device__ constexpr int fft_line_from(int n, int e) {
  return e > fft_stages(n)
             ? n
             : (n + fft_pad_p(n, e) * (n / fft_pad_q(n, e)) > fft_line_from(n, e + 1)
                    ? n + fft_pad_p(n, e) * (n / fft_pad_q(n, e))
                    : fft_line_from(n, e + 1));
}
// Doubles per line (re or im) that lds_fft touches.
__host__ __device__ constexpr int fft_line(int n) { return fft_line_from(n, 0); }
// A line stride for several lines per block: at least fft_line(n), 1 modulo 32 for the
// padded lengths (so lines that share a half-wave start on different banks), n + 1 below.
__host__ __device__ constexpr int fft_line_stride(int n) {
  return fft_line(n) == n ? n + 1 : fft_line(n) + ((33 - fft_line(n) % 32) % 32);
}
template <int N, int E>
__device__ __forceinline__ int fft_xidx(int i) {
  constexpr int P = fft_pad_p(N, E), Q = fft_pad_q(N, E);
  if constexpr (P == 0)
    return i;
  else
    return i + P * (int)((unsigned)i / Q);
}

// One Stockham stage S of an N-point transform over `lines` (<= LMAX) lines, NT threads.
template <int N, int S, int LMAX, int NT, int SIGN>
__device__ __forceinline__ void fft_stage(double* __restrict__ re, double* __restrict__ im,
                                          int lines, int ls, const double2* __restrict__ tw,
                                          int nmax) {
#pragma clang fp contract(fast)
  constexpr int R = fft_radix(N, S);
  constexpr int Ns = fft_span(N, S);
  constexpr int NB = N / R;  // butterflies per line
  constexpr int BPT = (LMAX * NB + NT - 1) / NT;
  const int total = lines * NB;
  const int tstep = nmax / (Ns * R);
  double vr[BPT][R], vi[BPT][R];
#pragma unroll
  for (int q = 0; q < BPT; q++) {
    const int b = (int)threadIdx.x + q * NT;
    if (b < total) {
      const int line = b / NB, j = b - line * NB;
      const double* lr = re + line * ls;
      const double* li = im + line * ls;
#pragma unroll
      for (int m = 0; m < R; m++) {
        const int a = fft_xidx<N, S>(j + m * NB);
        vr[q][m] = lr[a];
        vi[q][m] = li[a];
      }
      if constexpr (Ns > 1) {  // v[m] *= exp(SIGN 2 pi i m k / (Ns R))
        const int k = j & (Ns - 1);
        twiddle_powers<R, SIGN>(vr[q], vi[q], tw, k * tstep);
      }
      dft<R, SIGN>(vr[q], vi[q]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < BPT; q++) {
    const int b = (int)threadIdx.x + q * NT;
    if (b < total) {
      const int line = b / NB, j = b - line * NB;
      const int k = j & (Ns - 1);
      double* dr = re + line * ls;
      double* di = im + line * ls;
#pragma unroll
      for (int m = 0; m < R; m++) {
        const int a = fft_xidx<N, S + 1>((j - k) * R + k + m * Ns);
        dr[a] = vr[q][m];
        di[a] = vi[q][m];
      }
    }
  }
  __syncthreads();
}

template <int N, int S, int LMAX, int NT, int SIGN>
__device__ __forceinline__ void fft_stages_from(double* re, double* im, int lines, int ls,
                                                const double2* tw, int nmax) {
  if constexpr (S < fft_stages(N)) {
    fft_stage<N, S, LMAX, NT, SIGN>(re, im, lines, ls, tw, nmax);
    fft_stages_from<N, S + 1, LMAX, NT, SIGN>(re, im, lines, ls, tw, nmax);
  }
}

// N-point FFT of `lines` (<= LMAX) lines in LDS with NT threads; every thread of the block
// must call it, and the data must be visible to all threads on entry (barrier before).
// Line l occupies re/im[l*ls .. l*ls + fft_line(N)), so ls >= fft_line(N) when lines > 1.
// Returns after a barrier.
template <int N, int LMAX, int NT, int SIGN>
__device__ __forceinline__ void lds_fft(double* __restrict__ re, double* __restrict__ im,
                                        int lines, int ls, const double2* __restrict__ tw,
                                        int nmax) {
  fft_stages_from<N, 0, LMAX, NT, SIGN>(re, im, lines, ls, tw, nmax);
}

// ---- transforms with register-resident ends ------------------------------------------
// A plan runs the radices of fft_radix in order (REV = false) or reversed (REV = true); any
// order is a valid Stockham factorisation.  The first stage of a plan takes its butterflies
// from registers filled by the caller (e.g. straight from global memory) and the last stage
// leaves its results in registers, so only the exchanges between stages go through LDS.
// At the ends thread t, butterfly q handles b = t + q*NT as line = b % LMAX, j = b / LMAX
// (lines fastest: a wave touches LMAX lines x a few consecutive points, contiguous when the
// lines are adjacent in memory); it holds points p = j + m*N/R0 before the first stage and
// outputs p = j + m*N/RL after the last.  A forward plan's output registers are the input
// registers of the reversed plan, so forward -> pointwise multiply -> inverse needs no LDS
// exchange in between.  Middle stages run j-fastest within a line.
//
// LDS layout per exchange e (data written by stage e-1, read by stage e): element i of line
// l is at l*LS + i + P*(i/Q), chosen for the 64 banks of 4 bytes (32 doubles per row):
//  - stage e-1 writes groups of Ns = span(e-1) consecutive points every Q = Ns*R points;
//    when Ns < 32 the groups of 32 consecutive butterflies would share banks, so Ns points
//    of padding follow every Q points;
//  - the first exchange is written by line-fastest threads at stride R0: LS = 1 (mod 32);
//  - the last exchange is read by line-fastest threads, 8 lines x 4 points per half wave:
//    LS = 4 (mod 32).
template <int N, bool REV>
__host__ __device__ constexpr int plan_radix(int s) {
  return REV ? fft_radix(N, fft_stages(N) - 1 - s) : fft_radix(N, s);
}
template <int N, bool REV>
__host__ __device__ constexpr int plan_span(int s) {
  return s == 0 ? 1 : plan_span<N, REV>(s - 1) * plan_radix<N, REV>(s - 1);
}
__host__ __device__ constexpr int round_to(int v, int mod, int rem) {
  return v + ((rem - v % mod) % mod + mod) % mod;
}
template <int N, bool REV>
__host__ __device__ constexpr int xch_q(int e) { return plan_span<N, REV>(e); }
template <int N, bool REV>
__host__ __device__ constexpr int xch_p(int e) {
  return e <= 1 ? 0
                : ((plan_span<N, REV>(e - 1) < 32 && plan_span<N, REV>(e) % 32 == 0)
                       ? plan_span<N, REV>(e - 1)
                       : 0);
}
template <int N, bool REV>
__host__ __device__ constexpr int xch_ls(int e) {
  return round_to(N + xch_p<N, REV>(e) * (N / xch_q<N, REV>(e)), 32,
                  (e == fft_stages(N) - 1 && e != 1) ? 4 : 1);
}
template <int N, bool REV, int E>
__device__ __forceinline__ int xch_idx(int line, int i) {
  constexpr int Q = xch_q<N, REV>(E), P = xch_p<N, REV>(E), LS = xch_ls<N, REV>(E);
  if constexpr (P == 0)
    return line * LS + i;
  else
    return line * LS + i + P * (i / Q);
}
// Doubles per re (or im) array of LDS for plan REV over LMAX lines.
template <int N, bool REV, int E = 1>
__host__ __device__ constexpr int plan_lds_line() {
  if constexpr (E >= fft_stages(N))
    return 0;
  else
    return xch_ls<N, REV>(E) > plan_lds_line<N, REV, E + 1>() ? xch_ls<N, REV>(E)
                                                              : plan_lds_line<N, REV, E + 1>();
}

template <int N, bool REV, int LMAX, int NT>
struct RegPlan {
  static constexpr int S = fft_stages(N);
  static constexpr int R0 = plan_radix<N, REV>(0), NB0 = N / R0;
  static constexpr int B0 = (LMAX * NB0 + NT - 1) / NT;
  static constexpr int RL = plan_radix<N, REV>(S - 1), NBL = N / RL;
  static constexpr int BL = (LMAX * NBL + NT - 1) / NT;
  static constexpr int LDS_LINE = plan_lds_line<N, REV>();  // doubles per line, re or im
};

// Middle stage S of plan REV: LDS (exchange S) -> LDS (exchange S+1).
template <int N, int S, bool REV, int LMAX, int NT, int SIGN>
__device__ __forceinline__ void plan_stage_ll(double* __restrict__ re, double* __restrict__ im,
                                              int lines, const double2* __restrict__ tw,
                                              int nmax) {
#pragma clang fp contract(fast)
  constexpr int R = plan_radix<N, REV>(S);
  constexpr int Ns = plan_span<N, REV>(S);
  constexpr int NB = N / R;
  constexpr int BPT = (LMAX * NB + NT - 1) / NT;
  const int total = lines * NB;
  const int tstep = nmax / (Ns * R);
  double vr[BPT][R], vi[BPT][R];
#pragma unroll
  for (int q = 0; q < BPT; q++) {
    const int b = (int)threadIdx.x + q * NT;
    if (b < total) {
      const int line = b / NB, j = b - line * NB;
      const int k = j & (Ns - 1);
#pragma unroll
      for (int m = 0; m < R; m++) {
        const int a = xch_idx<N, REV, S>(line, j + m * NB);
        vr[q][m] = re[a];
        vi[q][m] = im[a];
      }
      twiddle_powers<R, SIGN>(vr[q], vi[q], tw, k * tstep);
      dft<R, SIGN>(vr[q], vi[q]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < BPT; q++) {
    const int b = (int)threadIdx.x + q * NT;
    if (b < total) {
      const int line = b / NB, j = b - line * NB;
      const int k = j & (Ns - 1);
#pragma unroll
      for (int m = 0; m < R; m++) {
        const int a = xch_idx<N, REV, S + 1>(line, (j - k) * R + k + m * Ns);
        re[a] = vr[q][m];
        im[a] = vi[q][m];
      }
    }
  }
  __syncthreads();
}

template <int N, int S, int SEND, bool REV, int LMAX, int NT, int SIGN>
__device__ __forceinline__ void plan_stages_ll(double* re, double* im, int lines,
                                               const double2* tw, int nmax) {
  if constexpr (S < SEND) {
    plan_stage_ll<N, S, REV, LMAX, NT, SIGN>(re, im, lines, tw, nmax);
    plan_stages_ll<N, S + 1, SEND, REV, LMAX, NT, SIGN>(re, im, lines, tw, nmax);
  }
}

// First stage (no twiddles) on registers holding points p = j + m*N/R0; for a plan of more
// than one stage the results go to LDS (exchange 1), followed by a barrier.  No thread may
// still be reading the LDS lines when it is called.
template <int N, bool REV, int LMAX, int NT, int SIGN>
__device__ __forceinline__ void plan_first_regs(
    double* __restrict__ re, double* __restrict__ im, int lines,
    double (&vr)[RegPlan<N, REV, LMAX, NT>::B0][RegPlan<N, REV, LMAX, NT>::R0],
    double (&vi)[RegPlan<N, REV, LMAX, NT>::B0][RegPlan<N, REV, LMAX, NT>::R0]) {
  using P = RegPlan<N, REV, LMAX, NT>;
#pragma unroll
  for (int q = 0; q < P::B0; q++) dft<P::R0, SIGN>(vr[q], vi[q]);
  if constexpr (P::S > 1) {
#pragma unroll
    for (int q = 0; q < P::B0; q++) {
      const int b = (int)threadIdx.x + q * NT;
      const int line = b % LMAX, j = b / LMAX;
      if (b < LMAX * P::NB0 && line < lines) {
#pragma unroll
        for (int m = 0; m < P::R0; m++) {
          const int a = xch_idx<N, REV, 1>(line, j * P::R0 + m);
          re[a] = vr[q][m];
          im[a] = vi[q][m];
        }
      }
    }
    __syncthreads();
  }
}

// Last stage: from LDS (exchange S-1) into registers holding outputs p = j + m*N/RL.  No
// barrier at the end; LDS may be overwritten only after one.
template <int N, bool REV, int LMAX, int NT, int SIGN>
__device__ __forceinline__ void plan_last_regs(
    const double* __restrict__ re, const double* __restrict__ im, int lines,
    const double2* __restrict__ tw, int nmax,
    double (&vr)[RegPlan<N, REV, LMAX, NT>::BL][RegPlan<N, REV, LMAX, NT>::RL],
    double (&vi)[RegPlan<N, REV, LMAX, NT>::BL][RegPlan<N, REV, LMAX, NT>::RL]) {
#pragma clang fp contract(fast)
  using P = RegPlan<N, REV, LMAX, NT>;
  constexpr int R = P::RL, NB = P::NBL;
  const int tstep = nmax / N;
#pragma unroll
  for (int q = 0; q < P::BL; q++) {
    const int b = (int)threadIdx.x + q * NT;
    const int line = b % LMAX, j = b / LMAX;
    if (b < LMAX * NB && line < lines) {
#pragma unroll
      for (int m = 0; m < R; m++) {
        const int a = xch_idx<N, REV, P::S - 1>(line, j + m * NB);
        vr[q][m] = re[a];
        vi[q][m] = im[a];
      }
      twiddle_powers<R, SIGN>(vr[q], vi[q], tw, j * tstep);  // k = j: span N/R
      dft<R, SIGN>(vr[q], vi[q]);
    }
  }
}

// Transform of registers in the first-stage layout of plan REV into registers in its
// last-stage layout (for a one-stage plan the two coincide).  re/im: LDS arrays of
// LMAX * RegPlan::LDS_LINE doubles each.
template <int N, bool REV, int LMAX, int NT, int SIGN>
__device__ __forceinline__ void plan_regs(
    double* __restrict__ re, double* __restrict__ im, int lines,
    const double2* __restrict__ tw, int nmax,
    double (&vr)[RegPlan<N, REV, LMAX, NT>::B0][RegPlan<N, REV, LMAX, NT>::R0],
    double (&vi)[RegPlan<N, REV, LMAX, NT>::B0][RegPlan<N, REV, LMAX, NT>::R0],
    double (&ur)[RegPlan<N, REV, LMAX, NT>::BL][RegPlan<N, REV, LMAX, NT>::RL],
    double (&ui)[RegPlan<N, REV, LMAX, NT>::BL][RegPlan<N, REV, LMAX, NT>::RL]) {
  using P = RegPlan<N, REV, LMAX, NT>;
  plan_first_regs<N, REV, LMAX, NT, SIGN>(re, im, lines, vr, vi);
  if constexpr (P::S == 1) {
#pragma unroll
    for (int q = 0; q < P::BL; q++)
#pragma unroll
      for (int m = 0; m < P::RL; m++) {
        ur[q][m] = vr[q][m];
        ui[q][m] = vi[q][m];
      }
  } else {
    plan_stages_ll<N, 1, P::S - 1, REV, LMAX, NT, SIGN>(re, im, lines, tw, nmax);
    plan_last_regs<N, REV, LMAX, NT, SIGN>(re, im, lines, tw, nmax, ur, ui);
  }
}

// Host: twiddle table of nmax entries (nmax a power of two).
double2* make_twiddles(int nmax);
// Padded transform length for an n-point free-space convolution: the reference pads to
// 2n (PoissonSolverFFT2D.cc:27-28); any length >= 2n-1 gives the same linear convolution,
// and the next power of two of 2n-1 equals 2n whenever n is a power of two.
inline int pad_len(int n) {
  int m = 1;
  while (m < 2 * n - 1) m <<= 1;
  return m < 2 ? 2 : m;
}

// Calls f(std::integral_constant<int, N>{}) for the runtime power of two n in [2, 4096].
template <typename F>
inline void dispatch_pow2(int n, F&& f) {
  switch (n) {
    case 2: f(std::integral_constant<int, 2>{}); break;
    case 4: f(std::integral_constant<int, 4>{}); break;
    case 8: f(std::integral_constant<int, 8>{}); break;
    case 16: f(std::integral_constant<int, 16>{}); break;
    case 32: f(std::integral_constant<int, 32>{}); break;
    case 64: f(std::integral_constant<int, 64>{}); break;
    case 128: f(std::integral_constant<int, 128>{}); break;
    case 256: f(std::integral_constant<int, 256>{}); break;
    case 512: f(std::integral_constant<int, 512>{}); break;
    case 1024: f(std::integral_constant<int, 1024>{}); break;
    case 2048: f(std::integral_constant<int, 2048>{}); break;
    case 4096: f(std::integral_constant<int, 4096>{}); break;
    default: throw Error{"FFT length is not a power of two in [2, 4096]"};
  }
}

}  // namespace ob
