// In-LDS Stockham FFT core shared by the surface synthesis and the SAR
// focusing kernels (gfx950): in-register radix-2/4/8/16 inverse DFTs, radix
// passes over sequences held in LDS with padded addressing, the two-level
// twiddle tables, and streaming loads and stores of complex values.
// Forward transforms are the conjugate of the inverse of the conjugate.
#pragma once

#include "common.h"

namespace {

// Entries of a two-level twiddle table: exp(+2 pi i k / N) for k = 0..63,
// then for k = 64 m, m = 0..127 (N up to 8192).
constexpr int kTwEntries = 192;

// Streaming (nontemporal) store and load of one complex FP64 value: the
// workspace and the output fields are written once and read once by the next
// kernel, so they should not displace cached data (register spill slots, the
// coefficient rows, the twiddle tables).
typedef double nt_double2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void store_stream(double2* p, double2 v) {
  nt_double2 x;
  x.x = v.x;
  x.y = v.y;
  __builtin_nontemporal_store(x, (nt_double2*)p);
}

__device__ __forceinline__ double2 load_stream(const double2* p) {
  const nt_double2 x = __builtin_nontemporal_load((const nt_double2*)p);
  return make_double2(x.x, x.y);
}

// complex64 workspace elements: the FP64 value rounded to float32 parts
typedef float nt_float2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void store_stream32(float2* p, double2 v) {
  nt_float2 x;
  x.x = (float)v.x;
  x.y = (float)v.y;
  __builtin_nontemporal_store(x, (nt_float2*)p);
}

__device__ __forceinline__ double2 load_stream32(const float2* p) {
  const nt_float2 x = __builtin_nontemporal_load((const nt_float2*)p);
  return make_double2((double)x.x, (double)x.y);
}

__device__ __forceinline__ double2 cmul2(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// ---- in-register inverse DFTs, X[k] = sum_r x[r] exp(+2 pi i r k / R) ----
__device__ __forceinline__ double2 cadd2(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub2(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul_i(double2 a) { return make_double2(-a.y, a.x); }  // i a

__device__ __forceinline__ void dft4(double2& x0, double2& x1, double2& x2, double2& x3) {
  const double2 s02 = cadd2(x0, x2), d02 = csub2(x0, x2);
  const double2 s13 = cadd2(x1, x3), d13 = csub2(x1, x3);
  x0 = cadd2(s02, s13);
  x1 = cadd2(d02, cmul_i(d13));
  x2 = csub2(s02, s13);
  x3 = csub2(d02, cmul_i(d13));
}

// exp(+2 pi i m / 16) for m = 0..9 (the products r1 k2 of the radix-16 split)
__device__ __forceinline__ double2 w16(int m) {
  constexpr double c1 = 0.92387953251128675613, s1 = 0.38268343236508977173;
  constexpr double h = 0.70710678118654752440;
  switch (m) {
    case 0: return make_double2(1., 0.);
    case 1: return make_double2(c1, s1);
    case 2: return make_double2(h, h);
    case 3: return make_double2(s1, c1);
    case 4: return make_double2(0., 1.);
    case 6: return make_double2(-h, h);
    case 9: return make_double2(-c1, -s1);
    default: return make_double2(0., 0.);  // not used
  }
}

template <int R>
__device__ __forceinline__ void dft(double2 (&x)[R]) {
  if constexpr (R == 2) {
    const double2 a = x[0];
    x[0] = cadd2(a, x[1]);
    x[1] = csub2(a, x[1]);
  } else if constexpr (R == 4) {
    dft4(x[0], x[1], x[2], x[3]);
  } else if constexpr (R == 8) {
    // r = r1 + 4 r2, k = 2 k1 + k2: radix-2 over r2, twiddle w8^r1 on the odd
    // half (1, (1 + i) h, i, (-1 + i) h with h = 1/sqrt 2), radix-4 over r1
    constexpr double h = 0.70710678118654752440;
    double2 a0[4], a1[4];
#pragma unroll
    for (int r1 = 0; r1 < 4; ++r1) a0[r1] = cadd2(x[r1], x[r1 + 4]);
    a1[0] = csub2(x[0], x[4]);
    const double2 d1 = csub2(x[1], x[5]);
    a1[1] = make_double2(h * (d1.x - d1.y), h * (d1.x + d1.y));
    a1[2] = cmul_i(csub2(x[2], x[6]));
    const double2 d3 = csub2(x[3], x[7]);
    a1[3] = make_double2(-h * (d3.x + d3.y), h * (d3.x - d3.y));
    dft4(a0[0], a0[1], a0[2], a0[3]);
    dft4(a1[0], a1[1], a1[2], a1[3]);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      x[2 * k1] = a0[k1];
      x[2 * k1 + 1] = a1[k1];
    }
  } else {
    static_assert(R == 16, "radix 2, 4, 8 or 16");
    // r = r1 + 4 r2, k = 4 k1 + k2: radix-4 over r2, twiddle w16^(r1 k2), radix-4 over r1
    double2 a[4][4];
#pragma unroll
    for (int r1 = 0; r1 < 4; ++r1) {
      a[r1][0] = x[r1];
      a[r1][1] = x[r1 + 4];
      a[r1][2] = x[r1 + 8];
      a[r1][3] = x[r1 + 12];
      dft4(a[r1][0], a[r1][1], a[r1][2], a[r1][3]);
#pragma unroll
      for (int k2 = 1; k2 < 4; ++k2)
        if (r1 > 0) a[r1][k2] = cmul2(a[r1][k2], w16(r1 * k2));
    }
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) {
      dft4(a[0][k2], a[1][k2], a[2][k2], a[3][k2]);
#pragma unroll
      for (int k1 = 0; k1 < 4; ++k1) x[4 * k1 + k2] = a[k1][k2];
    }
  }
}

// FFT buffers in LDS hold element e at slot pad(e) = e + e / 8: the strided
// accesses of a Stockham pass (stride R elements of 16 bytes) then fall into
// distinct banks instead of a few.
__device__ __forceinline__ int pad(int e) { return e + (e >> 3); }
__host__ __device__ constexpr size_t padded(size_t n) { return n + (n >> 3); }

// Twiddle exp(+2 pi i k / N) from the two-level LDS table: tw[k & 63] for the
// low six bits times tw[64 + (k >> 6)] for the rest (one FP64 complex product,
// within 2 ulp of the directly rounded value).
__device__ __forceinline__ double2 twiddle(const double2* tw, int k) {
  const double2 lo = tw[k & 63];
  if (k < 64) return lo;
  return cmul2(lo, tw[64 + (k >> 6)]);
}

// Copy the first n entries of the two-level twiddle table of a transform
// length to LDS (64 + N / 64 entries are used by an N-point transform).
__device__ __forceinline__ void load_twiddles(double2* dst, const double2* src, int n = 128) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

// One radix-R pass of the inverse Stockham FFT over nseq sequences of length N
// (Ns = product of the radices of the passes before); element k of sequence s
// is at buf[pad(s * ss + k)], ss a multiple of 8. Butterfly j of a sequence
// reads x_r = e[j + r N/R], multiplies by w^(r jm N/(Ns R)) (jm = j mod Ns,
// w = exp(2 pi i / N); all 1 in the first pass), takes the radix-R DFT and
// writes y_r to e[(j - jm) R + jm + r Ns]. Every thread loads all of its
// butterflies before the barrier, so the pass runs in place; at most BPT
// butterflies per thread (checked by the caller's sizes). Padded addresses:
// pad(b + r m) = pad(b) + r (m + m / 8) when m is a multiple of 8, and
// pad(b + r) = pad(b) + r when b mod 8 + r < 8 (the first pass's writes).
// SHORT: every pass of the transform has N/R and Ns (after the first) multiples
// of 8, which holds for N >= 64 with radix 8/4 passes; otherwise (N = 16, 32)
// every address is padded on its own.
// The thread's index, opaque to the optimiser: every pass recomputes the
// index arithmetic derived from it instead of values hoisted out of the
// callers' loops staying live in registers (across a persistent kernel's
// work items they pushed the column pass of the synthesis into spills).
__device__ __forceinline__ int pass_tid() {
  int t = threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
}

template <int NT, int R, int BPT, bool SHORT>
__device__ __forceinline__ void stockham_pass(double2* buf, int logN, int Ns, int logNs, int nseq,
                                              int ss, const double2* tw) {
  constexpr int LR = R == 2 ? 1 : R == 4 ? 2 : R == 8 ? 3 : 4;
  const int tid = pass_tid();
  const int lnb = logN - LR;
  const int nb = 1 << lnb;
  const int total = nseq << lnb;
  const bool rd_fast = SHORT || (nb & 7) == 0;
  const int rd_step = nb + (nb >> 3);
  double2 v[BPT][R];
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    const int id = tid + b * NT;
    if (id < total) {
      const int s = id >> lnb, j = id & (nb - 1);
      const int base = s * ss + j;
      if (rd_fast) {
        const double2* p = buf + pad(base);
#pragma unroll
        for (int r = 0; r < R; ++r) v[b][r] = p[r * rd_step];
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) v[b][r] = buf[pad(base + r * nb)];
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    const int id = tid + b * NT;
    if (id < total) {
      const int s = id >> lnb, j = id & (nb - 1);
      const int jm = j & (Ns - 1);
      if (Ns > 1) {
        // w^(r step) by repeated multiplication with w^step (R - 2 products; the
        // rounding error grows by about one ulp per product, far below float32)
        const int step = jm << (logN - logNs - LR);   // jm * N / (Ns * R)
        const double2 w1 = twiddle(tw, step);
        double2 wr = w1;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          v[b][r] = cmul2(v[b][r], wr);
          if (r + 1 < R) wr = cmul2(wr, w1);
        }
      }
      dft<R>(v[b]);
      const int o = s * ss + (j - jm) * R + jm;
      double2* p = buf + pad(o);
      if (Ns == 1) {
#pragma unroll
        for (int r = 0; r < R; ++r) p[r] = v[b][r];
      } else if (SHORT || (Ns & 7) == 0) {
        const int st = Ns + (Ns >> 3);
#pragma unroll
        for (int r = 0; r < R; ++r) p[r * st] = v[b][r];
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) buf[pad(o + r * Ns)] = v[b][r];
      }
    }
  }
  __syncthreads();
}

// In-LDS inverse FFT of nseq power-of-two sequences of length N, element k of
// sequence s at buf[pad(s * ss + k)]; tw is the two-level twiddle table in LDS;
// results stay in buf. Passes take radix 2^min(left, LRMAX) (one step smaller
// when that would leave a radix-2 pass). EMAX bounds nseq * N / NT, the
// elements one thread holds in a pass. Consecutive threads take consecutive
// butterflies of one sequence. SHORT (N >= 64): see stockham_pass.
template <int NT, int EMAX, int LRMAX, bool SHORT>
__device__ __forceinline__ void lds_ifft(double2* buf, int N, int logN, int nseq, int ss,
                                         const double2* tw) {
  (void)N;
  int Ns = 1, logNs = 0, left = logN;
  while (left > 0) {
    int lr = left < LRMAX ? left : LRMAX;
    if (left - lr == 1 && lr > 2) --lr;
    switch (lr) {
      case 1:
        stockham_pass<NT, 2, (EMAX / 2 > 0 ? EMAX / 2 : 1), SHORT>(buf, logN, Ns, logNs, nseq, ss,
                                                                  tw);
        break;
      case 2:
        stockham_pass<NT, 4, (EMAX / 4 > 0 ? EMAX / 4 : 1), SHORT>(buf, logN, Ns, logNs, nseq, ss,
                                                                  tw);
        break;
      case 3:
        if constexpr (LRMAX >= 3)
          stockham_pass<NT, 8, (EMAX / 8 > 0 ? EMAX / 8 : 1), SHORT>(buf, logN, Ns, logNs, nseq,
                                                                    ss, tw);
        break;
      default:
        if constexpr (LRMAX >= 4)
          stockham_pass<NT, 16, (EMAX / 16 > 0 ? EMAX / 16 : 1), SHORT>(buf, logN, Ns, logNs, nseq,
                                                                        ss, tw);
        break;
    }
    Ns <<= lr;
    logNs += lr;
    left -= lr;
  }
}

// lds_ifft with the length fixed at compile time (N = 2^LOGN >= 64, sequences
// N apart): every pass's length, stride, twiddle step and padded offsets are
// constants, so the loads and stores of a butterfly share one address and
// take immediate offsets, and the pass loop is unrolled. Same arithmetic in
// the same order as lds_ifft, so the same bits.
template <int NT, int R, int BPT, int LOGN, int LOGNS>
__device__ __forceinline__ void stockham_pass_n(double2* buf, int nseq, const double2* tw) {
  constexpr int LR = R == 2 ? 1 : R == 4 ? 2 : R == 8 ? 3 : 4;
  constexpr int N = 1 << LOGN, Ns = 1 << LOGNS;
  constexpr int lnb = LOGN - LR, nb = 1 << lnb;
  static_assert(nb % 8 == 0 && (Ns == 1 || Ns % 8 == 0), "padded strides need multiples of 8");
  constexpr int rd_step = nb + (nb >> 3);
  constexpr int wr_step = Ns + (Ns >> 3);
  const int tid = pass_tid();
  const int total = nseq << lnb;
  double2 v[BPT][R];
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    const int id = tid + b * NT;
    if (id < total) {
      const int s = id >> lnb, j = id & (nb - 1);
      const double2* p = buf + pad(s * N + j);
#pragma unroll
      for (int r = 0; r < R; ++r) v[b][r] = p[r * rd_step];
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    const int id = tid + b * NT;
    if (id < total) {
      const int s = id >> lnb, j = id & (nb - 1);
      const int jm = j & (Ns - 1);
      if constexpr (Ns > 1) {
        const double2 w1 = twiddle(tw, jm << (LOGN - LOGNS - LR));
        double2 wr = w1;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          v[b][r] = cmul2(v[b][r], wr);
          if (r + 1 < R) wr = cmul2(wr, w1);
        }
      }
      dft<R>(v[b]);
      double2* p = buf + pad(s * N + (j - jm) * R + jm);
#pragma unroll
      for (int r = 0; r < R; ++r) p[r * (Ns == 1 ? 1 : wr_step)] = v[b][r];
    }
  }
  __syncthreads();
}

template <int NT, int EMAX, int LRMAX, int LOGN, int LOGNS = 0>
__device__ __forceinline__ void lds_ifft_n(double2* buf, int nseq, const double2* tw) {
  if constexpr (LOGNS < LOGN) {
    constexpr int left = LOGN - LOGNS;
    constexpr int lr0 = left < LRMAX ? left : LRMAX;
    constexpr int LR = (left - lr0 == 1 && lr0 > 2) ? lr0 - 1 : lr0;
    constexpr int R = 1 << LR;
    stockham_pass_n<NT, R, (EMAX / R > 0 ? EMAX / R : 1), LOGN, LOGNS>(buf, nseq, tw);
    lds_ifft_n<NT, EMAX, LRMAX, LOGN, LOGNS + LR>(buf, nseq, tw);
  }
}

// lds_ifft of sequences N apart, taking the compile-time path for N = 512 ..
// 8192 (as far as a workgroup holds one sequence: EMAX NT >= N).
template <int NT, int EMAX, int LRMAX, bool SHORT, int L = 13>
__device__ __forceinline__ void lds_ifft_any(double2* buf, int N, int logN, int nseq,
                                             const double2* tw) {
  if constexpr (L >= 9) {
    if constexpr (EMAX * NT >= (1 << L)) {
      if (logN == L) {
        lds_ifft_n<NT, EMAX, LRMAX, L>(buf, nseq, tw);
        return;
      }
    }
    lds_ifft_any<NT, EMAX, LRMAX, SHORT, L - 1>(buf, N, logN, nseq, tw);
  } else {
    lds_ifft<NT, EMAX, LRMAX, SHORT>(buf, N, logN, nseq, N, tw);
  }
}

}  // namespace
