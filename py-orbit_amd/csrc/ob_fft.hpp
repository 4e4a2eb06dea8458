// ob_fft.hpp -- FP64 complex FFTs staged in LDS, used by the Hockney Poisson solvers.
//
// A workgroup transforms `lines` lines of power-of-two length n held in LDS as split
// re/im arrays (line stride n).  Forward transforms use decimation in time (input in
// bit-reversed order, output natural); inverse transforms use decimation in frequency
// (input natural, output bit-reversed), so a forward -> multiply -> inverse sequence
// needs no permutation pass in between.  Conventions follow FFTW: forward exponent -1,
// unnormalised.
//
// Twiddles come from a table tw[k] = exp(-2 pi i k / nmax), k < nmax/2, computed on the
// host in long double from the exact rational angle and rounded once.
#pragma once

#include "ob_internal.hpp"

namespace ob {

__device__ __forceinline__ int bitrev(int j, int log2n) {
  return (int)(__brev((unsigned)j) >> (32 - log2n));
}

// Radix-2 DIT, in place; data enters in bit-reversed order.  sign: -1 forward, +1 inverse.
// Line l starts at l*ls (ls >= n; ls = n+1 spreads lines over the LDS banks).
__device__ __forceinline__ void lds_fft_dit(double* __restrict__ re, double* __restrict__ im,
                                            int n, int log2n, int lines,
                                            const double2* __restrict__ tw, int nmax,
                                            double sign, int ls = 0) {
  if (ls == 0) ls = n;
  const int nb = lines * (n >> 1);
  for (int s = 1; s <= log2n; s++) {
    const int half = 1 << (s - 1);
    const int tstep = nmax >> s;
    for (int t = threadIdx.x; t < nb; t += blockDim.x) {
      const int line = t >> (log2n - 1);
      const int j = t & ((n >> 1) - 1);
      const int k = j & (half - 1);
      const int a = line * ls + ((j >> (s - 1)) << s) + k;
      const int c = a + half;
      const double2 w = tw[k * tstep];
      const double wr = w.x, wi = sign < 0 ? w.y : -w.y;
      const double cr = re[c], ci = im[c];
      const double tr = cr * wr - ci * wi;
      const double ti = cr * wi + ci * wr;
      const double ar = re[a], ai = im[a];
      re[c] = ar - tr;
      im[c] = ai - ti;
      re[a] = ar + tr;
      im[a] = ai + ti;
    }
    __syncthreads();
  }
}

// Radix-2 DIF, in place; output leaves in bit-reversed order.
__device__ __forceinline__ void lds_fft_dif(double* __restrict__ re, double* __restrict__ im,
                                            int n, int log2n, int lines,
                                            const double2* __restrict__ tw, int nmax,
                                            double sign, int ls = 0) {
  if (ls == 0) ls = n;
  const int nb = lines * (n >> 1);
  for (int s = log2n; s >= 1; s--) {
    const int half = 1 << (s - 1);
    const int tstep = nmax >> s;
    for (int t = threadIdx.x; t < nb; t += blockDim.x) {
      const int line = t >> (log2n - 1);
      const int j = t & ((n >> 1) - 1);
      const int k = j & (half - 1);
      const int a = line * ls + ((j >> (s - 1)) << s) + k;
      const int c = a + half;
      const double2 w = tw[k * tstep];
      const double wr = w.x, wi = sign < 0 ? w.y : -w.y;
      const double ar = re[a], ai = im[a], cr = re[c], ci = im[c];
      re[a] = ar + cr;
      im[a] = ai + ci;
      const double dr = ar - cr, di = ai - ci;
      re[c] = dr * wr - di * wi;
      im[c] = dr * wi + di * wr;
    }
    __syncthreads();
  }
}

// Host: twiddle table for transforms up to length nmax (power of two).
double2* make_twiddles(int nmax);
__host__ __device__ inline int ilog2(int n) {
  int l = 0;
  while ((1 << l) < n) l++;
  return l;
}
// Padded transform length for an n-point free-space convolution: the reference pads to
// 2n (PoissonSolverFFT2D.cc:27-28); any length >= 2n-1 gives the same linear convolution,
// and the next power of two of 2n-1 equals 2n whenever n is a power of two.
inline int pad_len(int n) {
  int m = 1;
  while (m < 2 * n - 1) m <<= 1;
  return m < 2 ? 2 : m;
}

}  // namespace ob
