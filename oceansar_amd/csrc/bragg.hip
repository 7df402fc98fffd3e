// Bragg facet kernel: Romeiser & Alpers (1997) composite-model RCS of every
// facet, in FP64, written as float32 amplitudes sqrt(RCS) for the echo
// accumulator, with the per-polarisation RCS sum for NRCS_avg.
//
// Replaces RCSRomeiser97.rcs_numba/_rcs_numba (oceansar/nrcs/romeiser97.py:
// 251-393) and the driver lines that take sqrt(RCS) and sum it
// (oceansar/radarsim/sar_raw_nompi.py:432-467). One thread per facet,
// coalesced struct-of-arrays reads of the two slope fields; the spectrum
// table lookup reproduces scipy interp1d(kind='linear', fill_value=1).
#include "common.h"

namespace {

struct cd {
  double re, im;
};
__device__ __forceinline__ cd cadd(cd a, cd b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cd cmulc(cd a, cd b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ cd cscale(cd a, double s) { return {a.re * s, a.im * s}; }
__device__ __forceinline__ cd cdiv(cd a, cd b) {
  // Smith's algorithm, as numpy divides complex128
  if (fabs(b.re) >= fabs(b.im)) {
    const double rat = b.im / b.re;
    const double scl = 1.0 / (b.re + b.im * rat);
    return {(a.re + a.im * rat) * scl, (a.im - a.re * rat) * scl};
  }
  const double rat = b.re / b.im;
  const double scl = 1.0 / (b.im + b.re * rat);
  return {(a.re * rat + a.im) * scl, (a.im * rat - a.re) * scl};
}
__device__ __forceinline__ double cabs2(cd a) {
  const double h = hypot(a.re, a.im);
  return h * h;
}

constexpr double kG = 9.80665;
constexpr double kRho = 1000.;
constexpr double kSurfT = 0.072;
constexpr double kX0 = 22e3;

// Delta(k) of the Elfouhaily spreading (oceansar/spread/elfouhaily.py:26-52).
__device__ double elf_delta(double k, double U, double fetch) {
  const double k_0 = kG / (U * U);
  const double X = k_0 * fetch;
  const double Omega_c = 0.84 * pow(tanh(pow(X / kX0, 0.4)), -0.75);
  const double cK = sqrt(kG / k + kSurfT / kRho * k);
  const double k_p = k_0 * Omega_c * Omega_c;
  const double cK_p = sqrt(kG / k_p + kSurfT / kRho * k_p);
  const double k_m = sqrt(kRho * kG / kSurfT);
  const double cK_m = sqrt(kG / k_m + kSurfT / kRho * k_m);
  const double C_10 = (0.8 + 0.065 * U) * 1e-3;
  const double ustar = sqrt(C_10) * U;
  const double a_0 = log(2.) / 4.;
  const double a_m = 0.13 * ustar / cK_m;
  return tanh(a_0 + 4. * pow(cK / cK_p, 2.5) + a_m * pow(cK_m / cK, 2.5));
}

__device__ double spread_g(int model, double k, double th, double delta) {
  if (model == 0) {
    return (th > -M_PI / 2. && th < M_PI / 2.) ? (1. + delta * cos(2. * th)) / M_PI : 0.;
  }
  const double c = cos(th);
  return (th < -M_PI / 2. || th > M_PI / 2.) ? 0. : 1. / M_PI * c * c;
}

// scipy interp1d linear with fill_value=1 (searchsorted side='left').
__device__ double interp_fill1(double x, const double* xp, const double* fp, int n) {
  if (x < xp[0] || x > xp[n - 1]) return 1.0;
  if (x != x) return x;  // NaN propagates, as in scipy
  int lo = 0, hi = n;  // first index with xp[i] >= x
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (xp[mid] < x) lo = mid + 1; else hi = mid;
  }
  int idx = lo < 1 ? 1 : (lo > n - 1 ? n - 1 : lo);
  const double xl = xp[idx - 1], xh = xp[idx];
  const double yl = fp[idx - 1], yh = fp[idx];
  const double slope = (yh - yl) / (xh - xl);
  return slope * (x - xl) + yl;
}

__global__ __launch_bounds__(256) void facet_bragg_kernel(osar_bragg_desc d) {
  const size_t n = (size_t)d.nx * d.ny;
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int npol = d.pol == 2 ? 2 : 1;
  double sums[2] = {0., 0.};
  if (idx < n) {
    // fields are x-major: idx = col * ny + row
    const int col = (int)(idx / d.ny);
    const int row = (int)(idx - (size_t)col * d.ny);
    const double t_now = d.pulse * d.t_step;
    const double az_now = (t_now - d.t_span / 2.) * d.v_ground;
    const double az = asin((d.y_row[row] - az_now) / d.gr0);
    const double inc = d.inc_col[col];
    const double diffx = (double)d.diffx[idx];
    const double diffy = (double)d.diffy[idx];
    // romeiser97.py:262-314
    const double cos_az = cos(az), sin_az = sin(az);
    const double s_p = atan(cos_az * diffx + sin_az * diffy);
    const double s_t = atan(-sin_az * diffx + cos_az * diffy);
    const double cos_s_t = cos(s_t);
    const double cos_s_t_2 = cos_s_t * cos_s_t;
    const double sin_s_t_2 = 1 - cos_s_t_2;
    const double sin_ip = sin(inc - s_p);
    const double cos_ip = cos(inc - s_p);
    const double cos_ip_2 = cos_ip * cos_ip;
    const double sin_ip_2 = sin_ip * sin_ip;
    const double cos_ti = cos_ip * cos_s_t;
    const double sin_ti_2 = 1 - cos_ti * cos_ti;
    const cd eps = {73., 18.};
    // sqrt(73 + 18j), principal branch
    const double eps_abs = hypot(73., 18.);
    const cd sq = {sqrt((eps_abs + 73.) / 2.), 18. / (2. * sqrt((eps_abs + 73.) / 2.))};
    const cd den_h = cadd({cos_ti, 0.}, sq);
    const cd b_hh = cdiv(eps, cmulc(den_h, den_h));
    const cd den_v = cadd(cscale(eps, cos_ti), sq);
    const cd b_vv = cdiv(cscale(cmulc(eps, eps), 1. + sin_ti_2), cmulc(den_v, den_v));
    const double wgt = cos_ip / (cos(inc) * cos(s_p));
    const double a = sin_s_t_2 / sin_ti_2;
    const double b = sin_ip_2 * cos_s_t_2 / sin_ti_2;
    const double kc = d.k0 * cos_ti;
    const double F0 = 8 * M_PI * (kc * kc) * (kc * kc);
    const double T_vv = wgt * (F0 * cabs2(cadd(cscale(b_vv, b), cscale(b_hh, a))));
    const double T_hh = wgt * (F0 * cabs2(cadd(cscale(b_hh, b), cscale(b_vv, a))));
    double k_b = 2 * d.k0 * sqrt(sin_ip_2 + cos_ip_2 * sin_s_t_2);
    if (k_b > 2 * d.k0) k_b = 0.;
    const double phi_b = az + atan((cos_ip * sin(-s_t)) / sin(inc - s_p));
    double th[2];
    th[0] = phi_b - d.wind_dir;
    th[1] = phi_b + M_PI - d.wind_dir;
    th[0] = atan2(sin(th[0]), cos(th[0]));  // np.angle(np.exp(1j * th))
    th[1] = atan2(sin(th[1]), cos(th[1]));
    const bool bad = 4 * sin_ip_2 < d.d * d.d;
    const double S1 = interp_fill1(k_b, d.spec_k, d.spec_S, d.n_spec);
    const double delta = d.spread == 0 ? elf_delta(k_b, d.wind_U, d.wind_fetch) : 0.;
    const double dxf = (double)(float)d.dx, dyf = (double)(float)d.dy;
    for (int c = 0; c < 2; ++c) {
      const double spectrum = S1 / k_b * spread_g(d.spread, k_b, th[c], delta);
      // pol 2 (DP): slot 0 = hh, slot 1 = vv
      for (int ip = 0; ip < npol; ++ip) {
        const bool is_hh = (d.pol == 1) || (d.pol == 2 && ip == 0);
        const double T = is_hh ? T_hh : T_vv;
        double rcs = T * spectrum * dxf * dyf;
        // romeiser97.py:366-392 (DP: hh masks only the plus component)
        const bool mask = bad && !(d.pol == 2 && is_hh && c == 0);
        if (mask) rcs = 0.;
        d.amp[((size_t)ip * 2 + c) * n + idx] = (float)sqrt(rcs);
        sums[ip] += rcs;
      }
    }
  }
  // block reduction of the RCS sums
  __shared__ double red[2][256];
  for (int ip = 0; ip < 2; ++ip) red[ip][threadIdx.x] = sums[ip];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0 && d.rcs_sum) {
    for (int ip = 0; ip < npol; ++ip) atomicAdd(&d.rcs_sum[ip], red[ip][0]);
  }
}

}  // namespace

extern "C" int osar_facet_bragg(osar_ctx* ctx, const osar_bragg_desc* d, void* stream) {
  OSAR_PTR(ctx);
  OSAR_PTR(d);
  OSAR_REQUIRE(d->nx > 0 && d->ny > 0, OSAR_EINVAL, "empty surface");
  OSAR_REQUIRE(d->pol >= 0 && d->pol <= 2, OSAR_EINVAL, "pol must be 0 (vv), 1 (hh), 2 (DP)");
  OSAR_REQUIRE(d->spread == 0 || d->spread == 1, OSAR_EUNSUPPORTED,
               "spread must be 0 (elfouhaily) or 1 (cos2s)");
  OSAR_REQUIRE(d->n_spec >= 2, OSAR_EINVAL, "spectrum table needs >= 2 points");
  OSAR_PTR(d->inc_col);
  OSAR_PTR(d->y_row);
  OSAR_PTR(d->spec_k);
  OSAR_PTR(d->spec_S);
  OSAR_PTR(d->diffx);
  OSAR_PTR(d->diffy);
  OSAR_PTR(d->amp);
  OSAR_HIP(hipSetDevice(ctx->device));
  const size_t n = (size_t)d->nx * d->ny;
  const unsigned blocks = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(facet_bragg_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, *d);
  OSAR_LAUNCHED("facet_bragg_kernel");
  return OSAR_OK;
}
