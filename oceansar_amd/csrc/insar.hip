// InSAR L1b processing of focused SLC data on the device (gfx950):
//
// osar_insar_coregister: the sub-sample azimuth shift of one channel's SLC
//   planes (oceansar/radarsim/insar_processor.py:165-174), y = IFFT(FFT(x) S)
//   along azimuth for lines of any length N, computed as the circular
//   convolution of x with h = IFFT_N(S). That circular convolution is a linear
//   one of the zero-padded column with the wrapped kernel g (g[j] = h[j],
//   g[M - j] = h[N - j]) over a power-of-two length M >= 2N - 1, so one
//   forward and one inverse in-LDS Stockham FFT per column do the work
//   (fft_lds.h) with the spectrum G = FFT_M(g) / M given by the host. For
//   M = 16384 (N up to 8192) the column does not fit LDS as one sequence: the
//   first radix-2 stage of the forward transform is taken by hand (the upper
//   half of the padded column is zero, so the even spectrum lines are the
//   8192-point FFT of x and the odd lines that of x exp(-2 pi i n / M)), both
//   halves go through LDS in turn, the even half's result is held in
//   registers, and the last radix-2 stage of the inverse combines them:
//   z[m] = U[m] + exp(2 pi i m / M) V[m] for m < N. The shifted values are
//   rounded to complex64 (the reference stores them into the complex64 array
//   it read from the SLC file) and multiplied by the flat-earth phasor of
//   their range sample (:176-185) in FP64, as complex128 output.
//
// osar_insar_multilook: intensities |s_c|^2 and interferograms s_c2 conj(s_c1)
//   smoothed along range, then azimuth (insar_processor.py:192-214 through
//   oceansar/utils/misc.py:108-197), cropped to the region of interest, and
//   the coherence t / sqrt(I_c1 I_c2) as magnitude and phase. Flat windows
//   are the reference's edge-normalised moving average, summed in the same
//   offset order in FP64 (bit for bit the reference's sums); other windows
//   are the normalised taps with zeros beyond the edges (the reference's
//   fftconvolve 'same'). Real images (the ATI processor's velocity fields)
//   are smoothed as they are.
#include "common.h"
#include "fft_lds.h"

namespace {

constexpr int kCoThreads = 1024;
constexpr int kCoElems = 8192;                       // complex values per workgroup in LDS
constexpr int kCoEmax = kCoElems / kCoThreads;       // per thread and pass

struct CoregArgs {
  int npol, naz, nrg;
  int H, logH;                // LDS transform length (M, or M / 2 when split)
  int M;
  int nseq;                   // columns per workgroup
  int nblk;                   // column blocks per plane
  long long in_pol_stride;    // complex64 elements between input planes
  const osar_c32* in;
  const double2* G;           // [M] spectrum of the wrapped kernel / M
  const double2* flat;        // [nrg] or null
  double2* out;               // [npol][naz][nrg]
  const double2* tw;
};

__device__ __forceinline__ double2 conj2(double2 a) { return make_double2(a.x, -a.y); }

// exp(i sign 2 pi m / M) for a power-of-two M (the argument 2 m / M is exact)
__device__ __forceinline__ double2 cis_frac(int m, int M, double sign) {
  double s, c;
  sincospi(2.0 * (double)m / (double)M, &s, &c);
  return make_double2(c, sign * s);
}

// complex product with every product rounded before the sum, as numpy's
// complex multiply: (a.x b.x - a.y b.y, a.x b.y + a.y b.x)
__device__ __forceinline__ double2 cmul_rn(double2 a, double2 b) {
  return make_double2(__dadd_rn(__dmul_rn(a.x, b.x), -__dmul_rn(a.y, b.y)),
                      __dadd_rn(__dmul_rn(a.x, b.y), __dmul_rn(a.y, b.x)));
}

template <bool SPLIT>
__global__ __launch_bounds__(kCoThreads) void coreg_kernel(CoregArgs a) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const int H = a.H, ns = a.nseq;
  double2* buf = (double2*)lds_raw;                    // [ns][H] padded
  double2* tw = buf + padded((size_t)ns * H);
  load_twiddles(tw, a.tw, kTwEntries);
  // neighbouring column blocks of a plane on one XCD: they share input lines in L2
  const int item = osar::xcd_item(blockIdx.x, a.nblk * a.npol);
  const int blk = item % a.nblk, pol = item / a.nblk;
  const int c0 = blk * ns;
  const osar_c32* in = a.in + (size_t)pol * a.in_pol_stride;
  double2* out = a.out + (size_t)pol * a.naz * a.nrg;
  const int total = ns * H;
  constexpr int P = SPLIT ? 2 : 1;
  double2 keep[SPLIT ? kCoEmax : 1];
  const bool lshort = H >= 64;
#pragma unroll
  for (int ph = 0; ph < P; ++ph) {
    // load (conjugated: the forward FFT is the conjugate of the inverse of the conjugate)
    for (int e = threadIdx.x; e < total; e += kCoThreads) {
      const int j = e % ns, n = e / ns, col = c0 + j;
      double2 v = make_double2(0., 0.);
      if (n < a.naz && col < a.nrg) {
        const osar_c32 c = in[(size_t)n * a.nrg + col];
        v = make_double2((double)c.re, (double)c.im);
        if (ph == 1) v = cmul2(v, cis_frac(n, a.M, -1.));
      }
      buf[pad(j * H + n)] = conj2(v);
    }
    __syncthreads();
    if (lshort)
      lds_ifft<kCoThreads, kCoEmax, 3, true>(buf, H, a.logH, ns, H, tw);
    else
      lds_ifft<kCoThreads, kCoEmax, 3, false>(buf, H, a.logH, ns, H, tw);
    for (int e = threadIdx.x; e < total; e += kCoThreads) {
      const int k = e % H;
      buf[pad(e)] = cmul2(conj2(buf[pad(e)]), a.G[(size_t)k * P + ph]);
    }
    __syncthreads();
    if (lshort)
      lds_ifft<kCoThreads, kCoEmax, 3, true>(buf, H, a.logH, ns, H, tw);
    else
      lds_ifft<kCoThreads, kCoEmax, 3, false>(buf, H, a.logH, ns, H, tw);
    // unrolled over the fixed slots so that the even half's results stay in registers
#pragma unroll
    for (int i = 0; i < kCoEmax; ++i) {
      const int e = threadIdx.x + i * kCoThreads;
      if (e >= total) break;
      const int j = e % ns, m = e / ns, col = c0 + j;
      double2 z = buf[pad(j * H + m)];
      if (SPLIT && ph == 0) {
        keep[i] = z;
        continue;
      }
      if (SPLIT) z = cadd2(keep[i], cmul2(cis_frac(m, a.M, 1.), z));
      if (m < a.naz && col < a.nrg) {
        // stored into the complex64 array of the SLC, then flattened in FP64
        double2 r = make_double2((double)(float)z.x, (double)(float)z.y);
        if (a.flat) r = cmul_rn(r, a.flat[col]);
        out[(size_t)m * a.nrg + col] = r;
      }
    }
    if (SPLIT && ph == 0) __syncthreads();
  }
}

// ---- multilook --------------------------------------------------------

struct MlArgs {
  int n_img, n_pairs, n_az, n_rg, is_complex;
  int rg_len, az_len, rg_flat, az_flat;
  int row_lo, rows;           // range-smoothed rows kept in the workspace
  int az_lo, az_hi, rg_lo, rg_hi;
  long long img_stride;       // elements between images
  const void* img;
  const double* rg_taps;
  const double* az_taps;
  double* wi;                 // [n_img][rows][cols]
  double2* wx;                // [n_pairs][rows][cols]
  double* intensity;          // [n_img][az_hi - az_lo][cols]
  double* coherence;          // [n_pairs][..]
  double* phase;              // [n_pairs][..]
};

// image pair index p -> (c1, c2), c1 < c2, in the reference's loop order
__device__ __forceinline__ void pair_of(int p, int C, int& c1, int& c2) {
  c1 = 0;
  while (p >= C - 1 - c1) {
    p -= C - 1 - c1;
    ++c1;
  }
  c2 = c1 + 1 + p;
}

// offsets of a window: out[n] = sum_{d = lo .. hi} w(d) data[n + d]; flat windows
// take d = -(L / 2) .. L - L / 2 - 1 with unit weights over the valid samples,
// tap windows the taps in correlation order (tap t at d = (L - 1) / 2 - t)
__device__ __forceinline__ int win_lo(int L, int flat) { return flat ? -(L / 2) : (L - 1) / 2 - (L - 1); }

__device__ __forceinline__ double intensity_at(const MlArgs& a, int c, size_t idx) {
  if (a.is_complex) {
    const double2 s = ((const double2*)a.img)[(size_t)c * a.img_stride + idx];
    const double h = hypot(s.x, s.y);       // numpy: abs(s) ** 2
    return h * h;
  }
  return ((const double*)a.img)[(size_t)c * a.img_stride + idx];
}

// range pass: one thread per (kept row, kept column, product)
__global__ __launch_bounds__(256) void ml_range_kernel(MlArgs a) {
  const int cols = a.rg_hi - a.rg_lo;
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  const int rr = blockIdx.y;
  const int prod = blockIdx.z;
  if (col >= cols) return;
  const int row = a.row_lo + rr;
  const int r = a.rg_lo + col;
  const int L = a.rg_len, dlo = win_lo(L, a.rg_flat);
  const size_t line = (size_t)row * a.n_rg;
  const size_t o = ((size_t)prod * a.rows + rr) * cols + col;
  if (prod < a.n_img) {
    double acc = 0.;
    int cnt = 0;
    for (int t = 0; t < L; ++t) {
      const int d = a.rg_flat ? dlo + t : (L - 1) / 2 - t;
      const int x = r + d;
      if (x < 0 || x >= a.n_rg) continue;
      const double v = intensity_at(a, prod, line + x);
      if (a.rg_flat) {
        acc = __dadd_rn(acc, v);
        ++cnt;
      } else {
        acc = fma(a.rg_taps[t], v, acc);
      }
    }
    a.wi[o] = a.rg_flat ? acc / (double)cnt : acc;
  } else {
    const int p = prod - a.n_img;
    int c1, c2;
    pair_of(p, a.n_img, c1, c2);
    const double2* s1 = (const double2*)a.img + (size_t)c1 * a.img_stride + line;
    const double2* s2 = (const double2*)a.img + (size_t)c2 * a.img_stride + line;
    double2 acc = make_double2(0., 0.);
    int cnt = 0;
    for (int t = 0; t < L; ++t) {
      const int d = a.rg_flat ? dlo + t : (L - 1) / 2 - t;
      const int x = r + d;
      if (x < 0 || x >= a.n_rg) continue;
      const double2 u = s2[x], v = s1[x];
      const double2 prd = cmul_rn(u, make_double2(v.x, -v.y));   // s2 conj(s1)
      if (a.rg_flat) {
        acc = make_double2(__dadd_rn(acc.x, prd.x), __dadd_rn(acc.y, prd.y));
        ++cnt;
      } else {
        const double w = a.rg_taps[t];
        acc = make_double2(fma(w, prd.x, acc.x), fma(w, prd.y, acc.y));
      }
    }
    if (a.rg_flat) acc = make_double2(acc.x / (double)cnt, acc.y / (double)cnt);
    a.wx[((size_t)p * a.rows + rr) * cols + col] = acc;
  }
}

// azimuth pass of the intensities: one thread per (output row, column, image)
__global__ __launch_bounds__(256) void ml_az_intensity_kernel(MlArgs a) {
  const int cols = a.rg_hi - a.rg_lo;
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  const int orow = blockIdx.y;
  const int c = blockIdx.z;
  if (col >= cols) return;
  const int row = a.az_lo + orow;
  const int L = a.az_len, dlo = win_lo(L, a.az_flat);
  double acc = 0.;
  int cnt = 0;
  for (int t = 0; t < L; ++t) {
    const int d = a.az_flat ? dlo + t : (L - 1) / 2 - t;
    const int y = row + d;
    if (y < 0 || y >= a.n_az) continue;
    const double v = a.wi[((size_t)c * a.rows + (y - a.row_lo)) * cols + col];
    if (a.az_flat) {
      acc = __dadd_rn(acc, v);
      ++cnt;
    } else {
      acc = fma(a.az_taps[t], v, acc);
    }
  }
  const int orows = a.az_hi - a.az_lo;
  a.intensity[((size_t)c * orows + orow) * cols + col] = a.az_flat ? acc / (double)cnt : acc;
}

// azimuth pass of the interferograms and their coherence (after the intensities)
__global__ __launch_bounds__(256) void ml_az_coherence_kernel(MlArgs a) {
  const int cols = a.rg_hi - a.rg_lo;
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  const int orow = blockIdx.y;
  const int p = blockIdx.z;
  if (col >= cols) return;
  const int row = a.az_lo + orow;
  const int L = a.az_len, dlo = win_lo(L, a.az_flat);
  double2 acc = make_double2(0., 0.);
  int cnt = 0;
  for (int t = 0; t < L; ++t) {
    const int d = a.az_flat ? dlo + t : (L - 1) / 2 - t;
    const int y = row + d;
    if (y < 0 || y >= a.n_az) continue;
    const double2 v = a.wx[((size_t)p * a.rows + (y - a.row_lo)) * cols + col];
    if (a.az_flat) {
      acc = make_double2(__dadd_rn(acc.x, v.x), __dadd_rn(acc.y, v.y));
      ++cnt;
    } else {
      const double w = a.az_taps[t];
      acc = make_double2(fma(w, v.x, acc.x), fma(w, v.y, acc.y));
    }
  }
  if (a.az_flat) acc = make_double2(acc.x / (double)cnt, acc.y / (double)cnt);
  int c1, c2;
  pair_of(p, a.n_img, c1, c2);
  const int orows = a.az_hi - a.az_lo;
  const size_t pix = (size_t)orow * cols + col;
  const double i1 = a.intensity[(size_t)c1 * orows * cols + pix];
  const double i2 = a.intensity[(size_t)c2 * orows * cols + pix];
  // numpy divides a complex by a real as a complex with zero imaginary part:
  // both parts times the reciprocal of the divisor
  const double inv = 1.0 / sqrt(__dmul_rn(i1, i2));
  const double re = __dmul_rn(acc.x, inv), im = __dmul_rn(acc.y, inv);
  const size_t o = (size_t)p * orows * cols + pix;
  a.coherence[o] = hypot(re, im);
  a.phase[o] = atan2(im, re);
}

int log2_exact(int n) {
  int l = 0;
  while ((1 << l) < n) ++l;
  return (1 << l) == n ? l : -1;
}

}  // namespace

extern "C" int osar_insar_coregister(osar_ctx* ctx, const osar_coreg_desc* d, void* stream) {
  OSAR_PTR(ctx);
  OSAR_PTR(d);
  OSAR_REQUIRE(d->n_pol >= 1 && d->n_az >= 1 && d->n_rg >= 1, OSAR_EINVAL,
               "empty SLC planes (%d x %d x %d)", d->n_pol, d->n_az, d->n_rg);
  const int lm = log2_exact(d->n_fft);
  OSAR_REQUIRE(lm >= 4 && d->n_fft <= 2 * kCoElems, OSAR_EUNSUPPORTED,
               "n_fft %d: a power of two in [16, %d]", d->n_fft, 2 * kCoElems);
  OSAR_REQUIRE(d->n_fft >= 2 * d->n_az - 1, OSAR_EINVAL,
               "n_fft %d shorter than 2 n_az - 1 = %d", d->n_fft, 2 * d->n_az - 1);
  OSAR_REQUIRE(d->in_pol_stride >= (int64_t)d->n_az * d->n_rg || d->n_pol == 1, OSAR_EINVAL,
               "in_pol_stride smaller than a plane");
  OSAR_PTR(d->in);
  OSAR_PTR(d->kernel);
  OSAR_PTR(d->out);
  OSAR_HIP(hipSetDevice(ctx->device));
  const bool split = d->n_fft > kCoElems;
  CoregArgs a{};
  a.npol = d->n_pol;
  a.naz = d->n_az;
  a.nrg = d->n_rg;
  a.M = d->n_fft;
  a.H = split ? d->n_fft / 2 : d->n_fft;
  a.logH = split ? lm - 1 : lm;
  a.nseq = kCoElems / a.H;
  if (a.nseq > d->n_rg) a.nseq = d->n_rg;
  a.nblk = (d->n_rg + a.nseq - 1) / a.nseq;
  a.in_pol_stride = d->in_pol_stride;
  a.in = d->in;
  a.G = (const double2*)d->kernel;
  a.flat = (const double2*)d->flat;
  a.out = (double2*)d->out;
  int rc = osar_twiddles(ctx, a.H, &a.tw);
  if (rc) return rc;
  const size_t lds = (padded((size_t)a.nseq * a.H) + kTwEntries) * sizeof(double2);
  OSAR_REQUIRE(lds <= ctx->max_lds, OSAR_EUNSUPPORTED, "coregistration needs %zu B of LDS (limit %zu)",
               lds, ctx->max_lds);
  const void* k = split ? (const void*)coreg_kernel<true> : (const void*)coreg_kernel<false>;
  rc = osar_reserve_lds(k, lds);
  if (rc) return rc;
  const dim3 grid((unsigned)((int64_t)a.nblk * a.npol));
  if (split)
    hipLaunchKernelGGL(coreg_kernel<true>, grid, dim3(kCoThreads), lds, (hipStream_t)stream, a);
  else
    hipLaunchKernelGGL(coreg_kernel<false>, grid, dim3(kCoThreads), lds, (hipStream_t)stream, a);
  OSAR_LAUNCHED("coreg_kernel");
  return OSAR_OK;
}

static void ml_rows(const osar_multilook_desc* d, int* row_lo, int* rows) {
  // rows of the range-smoothed images the azimuth windows of the kept rows reach
  const int L = d->az_len;
  const int before = d->az_flat ? L / 2 : L - 1 - (L - 1) / 2;
  const int after = d->az_flat ? L - L / 2 - 1 : (L - 1) / 2;
  int lo = d->az_lo - before, hi = d->az_hi + after;
  if (lo < 0) lo = 0;
  if (hi > d->n_az) hi = d->n_az;
  *row_lo = lo;
  *rows = hi - lo;
}

extern "C" int64_t osar_multilook_work_bytes(const osar_multilook_desc* d) {
  if (!d || d->n_img < 1 || d->az_hi <= d->az_lo || d->rg_hi <= d->rg_lo) return -1;
  int lo, rows;
  ml_rows(d, &lo, &rows);
  const int64_t cols = d->rg_hi - d->rg_lo;
  const int64_t pairs = d->is_complex ? (int64_t)d->n_img * (d->n_img - 1) / 2 : 0;
  // the interferogram sums start on a 16-byte boundary
  const int64_t real_words = ((int64_t)rows * cols * d->n_img + 1) & ~(int64_t)1;
  return 8 * real_words + 16 * (int64_t)rows * cols * pairs;
}

extern "C" int osar_insar_multilook(osar_ctx* ctx, const osar_multilook_desc* d, void* stream) {
  OSAR_PTR(ctx);
  OSAR_PTR(d);
  OSAR_REQUIRE(d->n_img >= 1 && d->n_img <= OSAR_ML_MAX_IMG, OSAR_EINVAL,
               "n_img %d outside [1, %d]", d->n_img, OSAR_ML_MAX_IMG);
  OSAR_REQUIRE(d->n_az >= 1 && d->n_rg >= 1, OSAR_EINVAL, "empty images");
  OSAR_REQUIRE(d->rg_len >= 1 && d->az_len >= 1, OSAR_EINVAL, "window lengths %d x %d",
               d->az_len, d->rg_len);
  OSAR_REQUIRE(0 <= d->az_lo && d->az_lo < d->az_hi && d->az_hi <= d->n_az && 0 <= d->rg_lo &&
                   d->rg_lo < d->rg_hi && d->rg_hi <= d->n_rg,
               OSAR_EINVAL, "crop [%d, %d) x [%d, %d) outside %d x %d", d->az_lo, d->az_hi,
               d->rg_lo, d->rg_hi, d->n_az, d->n_rg);
  OSAR_REQUIRE(d->img_stride >= (int64_t)d->n_az * d->n_rg || d->n_img == 1, OSAR_EINVAL,
               "img_stride smaller than an image");
  OSAR_REQUIRE(d->rg_flat || d->rg_taps, OSAR_EINVAL, "range taps missing");
  OSAR_REQUIRE(d->az_flat || d->az_taps, OSAR_EINVAL, "azimuth taps missing");
  OSAR_PTR(d->img);
  OSAR_PTR(d->work);
  OSAR_PTR(d->intensity);
  const int pairs = d->is_complex ? d->n_img * (d->n_img - 1) / 2 : 0;
  if (pairs) {
    OSAR_PTR(d->coherence);
    OSAR_PTR(d->phase);
  }
  OSAR_HIP(hipSetDevice(ctx->device));
  MlArgs a{};
  a.n_img = d->n_img;
  a.n_pairs = pairs;
  a.n_az = d->n_az;
  a.n_rg = d->n_rg;
  a.is_complex = d->is_complex ? 1 : 0;
  a.rg_len = d->rg_len;
  a.az_len = d->az_len;
  a.rg_flat = d->rg_flat;
  a.az_flat = d->az_flat;
  ml_rows(d, &a.row_lo, &a.rows);
  a.az_lo = d->az_lo;
  a.az_hi = d->az_hi;
  a.rg_lo = d->rg_lo;
  a.rg_hi = d->rg_hi;
  a.img_stride = d->img_stride;
  a.img = d->img;
  a.rg_taps = d->rg_taps;
  a.az_taps = d->az_taps;
  const int64_t cols = d->rg_hi - d->rg_lo;
  a.wi = d->work;
  a.wx = (double2*)(d->work + (((size_t)d->n_img * a.rows * cols + 1) & ~(size_t)1));
  a.intensity = d->intensity;
  a.coherence = d->coherence;
  a.phase = d->phase;
  hipStream_t s = (hipStream_t)stream;
  const unsigned gx = (unsigned)((cols + 255) / 256);
  hipLaunchKernelGGL(ml_range_kernel, dim3(gx, (unsigned)a.rows, (unsigned)(d->n_img + pairs)),
                     dim3(256), 0, s, a);
  OSAR_LAUNCHED("ml_range_kernel");
  const unsigned orows = (unsigned)(d->az_hi - d->az_lo);
  hipLaunchKernelGGL(ml_az_intensity_kernel, dim3(gx, orows, (unsigned)d->n_img), dim3(256), 0, s,
                     a);
  OSAR_LAUNCHED("ml_az_intensity_kernel");
  if (pairs) {
    hipLaunchKernelGGL(ml_az_coherence_kernel, dim3(gx, orows, (unsigned)pairs), dim3(256), 0, s,
                       a);
    OSAR_LAUNCHED("ml_az_coherence_kernel");
  }
  return OSAR_OK;
}
