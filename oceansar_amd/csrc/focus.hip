// SAR focusing of one channel's raw planes: the processing loop of sar_focus
// (oceansar/radarsim/sar_processor.py:131-225) as four FP64 in-LDS FFT passes
// over a complex128 workspace [pol][n_az][n_rg] (powers of two up to 8192):
//   1. rows, forward FFT along range of the zero-padded complex64 raw lines
//   2. columns, forward FFT along azimuth, times exp(4 pi i rcmc(fa) fr / c)
//   3. rows, inverse FFT along range (1 / n_rg), times the azimuth filter h(fa)
//   4. columns, inverse FFT along azimuth (1 / n_az), cropped to the kept
//      lines and the original range samples
// numpy.fft's forward transform is the conjugate of the inverse of the
// conjugate; every transform is the shared Stockham core (fft_lds.h). A
// workgroup transforms as many sequences as make 8192 elements (147 KB of
// LDS, 1024 threads): 8192 / n_rg lines or 8192 / n_az columns, the columns
// read as runs of adjacent elements of each line, the blocks of neighbouring
// columns on one XCD.
#include "common.h"
#include "fft_lds.h"

namespace {

constexpr int kFocThreads = 1024;
constexpr int kFocElems = 8192;                        // elements per workgroup
constexpr int kFocEmax = kFocElems / kFocThreads;      // elements per thread and pass

struct FocusArgs {
  int npol, az0, rg0, naz, nrg, lognaz, lognrg;
  int nseq_r, nseq_c;        // lines per row workgroup, columns per column workgroup
  int az_lo, az_hi;
  long long raw_pol_stride;
  const osar_c32* raw;
  const double* fr;
  const double* rcmc;
  const double2* h;
  double2* work;
  double2* slc;
  const double2* tw_rg;
  const double2* tw_az;
};

__device__ __forceinline__ double2 conj2(double2 a) { return make_double2(a.x, -a.y); }

// Pass 1 (FWD, rows): FFT along range of raw lines (zero beyond rg0 and az0).
// Pass 3 (!FWD, rows): inverse FFT along range of workspace lines, / n_rg, x h[a].
template <bool FWD>
__global__ __launch_bounds__(kFocThreads) void focus_rows_kernel(FocusArgs a) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  double2* buf = (double2*)lds_raw;                      // [nseq_r][nrg] padded
  double2* tw = buf + padded((size_t)a.nseq_r * a.nrg);  // twiddles
  load_twiddles(tw, a.tw_rg, kTwEntries);
  const int n = a.nrg, ns = a.nseq_r;
  const int line0 = blockIdx.x * ns;                     // flattened (pol, az) line index
  const int total = ns * n;
  // Every load of the block is issued before the first is used (unrolled,
  // unconditional: indices clamped into the block and the raw plane, values
  // outside selected away afterwards), so the loads overlap instead of each
  // waiting for the one before.
  double2 v[kFocEmax];
  if (FWD) {
    osar_c32 c[kFocEmax];
#pragma unroll
    for (int k = 0; k < kFocEmax; ++k) {
      const int e = min((int)threadIdx.x + k * kFocThreads, total - 1);
      const int s = e >> a.lognrg, r = e & (n - 1);
      const int line = line0 + s;
      const int pol = line / a.naz, az = line - pol * a.naz;
      c[k] = a.raw[(size_t)pol * a.raw_pol_stride + (size_t)min(az, a.az0 - 1) * a.rg0 +
                   min(r, a.rg0 - 1)];
    }
#pragma unroll
    for (int k = 0; k < kFocEmax; ++k) {
      const int e = min((int)threadIdx.x + k * kFocThreads, total - 1);
      const int s = e >> a.lognrg, r = e & (n - 1);
      const int line = line0 + s;
      const int az = line % a.naz;
      // zero beyond the raw lines and samples; the conjugate in
      v[k] = (az < a.az0 && r < a.rg0) ? make_double2((double)c[k].re, -(double)c[k].im)
                                       : make_double2(0., 0.);
    }
  } else {
#pragma unroll
    for (int k = 0; k < kFocEmax; ++k) {
      const int e = min((int)threadIdx.x + k * kFocThreads, total - 1);
      v[k] = a.work[(size_t)(line0 + (e >> a.lognrg)) * n + (e & (n - 1))];
    }
  }
  // a clamped index writes the block's last element again, with its own value
#pragma unroll
  for (int k = 0; k < kFocEmax; ++k)
    buf[pad(min((int)threadIdx.x + k * kFocThreads, total - 1))] = v[k];
  __syncthreads();
  if (n >= 64)
    lds_ifft_any<kFocThreads, kFocEmax, 3, true>(buf, n, a.lognrg, ns, tw);
  else
    lds_ifft<kFocThreads, kFocEmax, 3, false>(buf, n, a.lognrg, ns, n, tw);
  const double scale = 1.0 / (double)n;
  double2 hv[kFocEmax];   // the azimuth filter of each element's line, loaded up front
  if (!FWD) {
#pragma unroll
    for (int k = 0; k < kFocEmax; ++k) {
      const int e = min((int)threadIdx.x + k * kFocThreads, total - 1);
      hv[k] = a.h[(line0 + (e >> a.lognrg)) % a.naz];
    }
  }
#pragma unroll
  for (int k = 0; k < kFocEmax; ++k) {
    const int e = threadIdx.x + k * kFocThreads;
    if (e >= total) break;
    const int s = e >> a.lognrg, r = e & (n - 1);
    const int line = line0 + s;
    double2 w = buf[pad(e)];
    if (FWD)
      w = conj2(w);
    else
      w = cmul2(make_double2(w.x * scale, w.y * scale), hv[k]);
    a.work[(size_t)line * n + r] = w;
  }
}

// Pass 2 (FWD, columns): FFT along azimuth, x exp(i (4 pi rcmc[a]) / c fr[r]).
// Pass 4 (!FWD, columns): inverse FFT along azimuth, / n_az, kept lines to slc.
template <bool FWD>
__global__ __launch_bounds__(kFocThreads) void focus_cols_kernel(FocusArgs a) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const int n = a.naz, cb = a.nseq_c;
  double2* buf = (double2*)lds_raw;                 // [cb][naz] padded
  double2* tw = buf + padded((size_t)cb * n);
  load_twiddles(tw, a.tw_az, kTwEntries);
  const int nblk = a.nrg / cb;
  const int item = osar::xcd_item(blockIdx.x, nblk * a.npol);
  const int blk = item % nblk, pol = item / nblk;
  const int c0 = blk * cb;
  double2* plane = a.work + (size_t)pol * a.naz * a.nrg;
  const int total = cb * n;
  // every load issued before the first is used (see focus_rows_kernel)
  double2 v[kFocEmax];
#pragma unroll
  for (int k = 0; k < kFocEmax; ++k) {
    const int e = min((int)threadIdx.x + k * kFocThreads, total - 1);
    const int j = e % cb, az = e / cb;               // runs of cb adjacent columns per line
    v[k] = plane[(size_t)az * a.nrg + c0 + j];
  }
  // a clamped index writes the block's last element again, with its own value
#pragma unroll
  for (int k = 0; k < kFocEmax; ++k) {
    const int e = min((int)threadIdx.x + k * kFocThreads, total - 1);
    buf[pad((e % cb) * n + e / cb)] = FWD ? conj2(v[k]) : v[k];
  }
  __syncthreads();
  if (n >= 64)
    lds_ifft_any<kFocThreads, kFocEmax, 3, true>(buf, n, a.lognaz, cb, tw);
  else
    lds_ifft<kFocThreads, kFocEmax, 3, false>(buf, n, a.lognaz, cb, n, tw);
  const double scale = 1.0 / (double)n;
  const int nkeep = a.az_hi - a.az_lo;
  double rc[kFocEmax], fq[kFocEmax];   // RCMC terms of each element, loaded up front
  if (FWD) {
#pragma unroll
    for (int k = 0; k < kFocEmax; ++k) {
      const int e = min((int)threadIdx.x + k * kFocThreads, total - 1);
      rc[k] = a.rcmc[e / cb];
      fq[k] = a.fr[c0 + e % cb];
    }
  }
#pragma unroll
  for (int k = 0; k < kFocEmax; ++k) {
    const int e = threadIdx.x + k * kFocThreads;
    if (e >= total) break;
    const int j = e % cb, az = e / cb;
    const int r = c0 + j;
    double2 w = buf[pad(j * n + az)];
    if (FWD) {
      w = conj2(w);
      // numpy's order: exp(((4 pi i) rcmc / c) fr), a purely imaginary argument
      const double t = (4. * M_PI * rc[k]) / 299792458.0;
      double sn, cs;
      sincos(t * fq[k], &sn, &cs);
      plane[(size_t)az * a.nrg + r] = cmul2(w, make_double2(cs, sn));
    } else if (az >= a.az_lo && az < a.az_hi && r < a.rg0) {
      a.slc[((size_t)pol * nkeep + (az - a.az_lo)) * a.rg0 + r] =
          make_double2(w.x * scale, w.y * scale);
    }
  }
}

int ilog2f(int n) {
  int l = 0;
  while ((1 << l) < n) ++l;
  return (1 << l) == n ? l : -1;
}

}  // namespace

extern "C" int osar_sar_focus(osar_ctx* ctx, const osar_focus_desc* d, void* stream) {
  OSAR_PTR(ctx);
  OSAR_PTR(d);
  const int lr = ilog2f(d->n_rg), la = ilog2f(d->n_az);
  OSAR_REQUIRE(lr >= 4 && la >= 4 && d->n_rg <= kFocElems && d->n_az <= kFocElems,
               OSAR_EUNSUPPORTED, "padded sizes %d x %d: powers of two in [16, %d]", d->n_az,
               d->n_rg, kFocElems);
  OSAR_REQUIRE(d->n_pol >= 1 && d->n_az0 >= 1 && d->n_rg0 >= 1 && d->n_az0 <= d->n_az &&
                   d->n_rg0 <= d->n_rg,
               OSAR_EINVAL, "raw %d x %d does not fit the padded %d x %d", d->n_az0, d->n_rg0,
               d->n_az, d->n_rg);
  OSAR_REQUIRE(d->az_lo >= 0 && d->az_lo < d->az_hi && d->az_hi <= d->n_az0, OSAR_EINVAL,
               "kept lines [%d, %d) outside [0, %d)", d->az_lo, d->az_hi, d->n_az0);
  OSAR_REQUIRE(d->raw_pol_stride >= (int64_t)d->n_az0 * d->n_rg0 || d->n_pol == 1, OSAR_EINVAL,
               "raw_pol_stride smaller than a plane");
  OSAR_PTR(d->raw);
  OSAR_PTR(d->fr);
  OSAR_PTR(d->rcmc);
  OSAR_PTR(d->h);
  OSAR_PTR(d->work);
  OSAR_PTR(d->slc);
  OSAR_HIP(hipSetDevice(ctx->device));
  FocusArgs a{};
  a.npol = d->n_pol;
  a.az0 = d->n_az0;
  a.rg0 = d->n_rg0;
  a.naz = d->n_az;
  a.nrg = d->n_rg;
  a.lognaz = la;
  a.lognrg = lr;
  a.nseq_r = kFocElems / d->n_rg;
  a.nseq_c = kFocElems / d->n_az;
  if (a.nseq_c > d->n_rg) a.nseq_c = d->n_rg;
  if (a.nseq_r > d->n_az) a.nseq_r = d->n_az;
  a.az_lo = d->az_lo;
  a.az_hi = d->az_hi;
  a.raw_pol_stride = d->raw_pol_stride;
  a.raw = d->raw;
  a.fr = d->fr;
  a.rcmc = d->rcmc;
  a.h = (const double2*)d->h;
  a.work = (double2*)d->work;
  a.slc = (double2*)d->slc;
  int rc = osar_twiddles(ctx, d->n_rg, &a.tw_rg);
  if (rc) return rc;
  rc = osar_twiddles(ctx, d->n_az, &a.tw_az);
  if (rc) return rc;
  const size_t lds_r = (padded((size_t)a.nseq_r * d->n_rg) + kTwEntries) * sizeof(double2);
  const size_t lds_c = (padded((size_t)a.nseq_c * d->n_az) + kTwEntries) * sizeof(double2);
  OSAR_REQUIRE(lds_r <= ctx->max_lds && lds_c <= ctx->max_lds, OSAR_EUNSUPPORTED,
               "focusing needs %zu / %zu B of LDS (limit %zu)", lds_r, lds_c, ctx->max_lds);
  const void* kr[2] = {(const void*)focus_rows_kernel<true>, (const void*)focus_rows_kernel<false>};
  const void* kc[2] = {(const void*)focus_cols_kernel<true>, (const void*)focus_cols_kernel<false>};
  for (int i = 0; i < 2; ++i) {
    rc = osar_reserve_lds(kr[i], lds_r);
    if (rc) return rc;
    rc = osar_reserve_lds(kc[i], lds_c);
    if (rc) return rc;
  }
  hipStream_t s = (hipStream_t)stream;
  const dim3 grid_r((unsigned)((int64_t)d->n_pol * d->n_az / a.nseq_r));
  const dim3 grid_c((unsigned)((int64_t)d->n_pol * (d->n_rg / a.nseq_c)));
  hipLaunchKernelGGL(focus_rows_kernel<true>, grid_r, dim3(kFocThreads), lds_r, s, a);
  OSAR_LAUNCHED("focus_rows_kernel");
  hipLaunchKernelGGL(focus_cols_kernel<true>, grid_c, dim3(kFocThreads), lds_c, s, a);
  OSAR_LAUNCHED("focus_cols_kernel");
  hipLaunchKernelGGL(focus_rows_kernel<false>, grid_r, dim3(kFocThreads), lds_r, s, a);
  OSAR_LAUNCHED("focus_rows_kernel");
  hipLaunchKernelGGL(focus_cols_kernel<false>, grid_c, dim3(kFocThreads), lds_c, s, a);
  OSAR_LAUNCHED("focus_cols_kernel");
  return OSAR_OK;
}
