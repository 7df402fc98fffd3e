// Bragg facet kernel: Romeiser & Alpers (1997) composite-model RCS of every
// facet, in float32 from the float32 slopes (FP64 row geometry, table search
// and sums), written as float32 amplitudes sqrt(RCS) for the echo
// accumulator, with the per-polarisation RCS sum for NRCS_avg.
//
// Replaces RCSRomeiser97.rcs_numba/_rcs_numba (oceansar/nrcs/romeiser97.py:
// 251-393) and the driver lines that take sqrt(RCS) and sum it
// (oceansar/radarsim/sar_raw_nompi.py:432-467). One thread per facet,
// coalesced struct-of-arrays reads of the two slope fields; the spectrum
// table lookup reproduces scipy interp1d(kind='linear', fill_value=1).
#include <cmath>

#include "common.h"

namespace {

constexpr double kG = 9.80665;
constexpr double kRho = 1000.;
constexpr double kSurfT = 0.072;
constexpr double kX0 = 22e3;

// Facet-independent terms of the Elfouhaily Delta(k) (oceansar/spread/
// elfouhaily.py:26-52) and of the geometry, computed once per call on the host.
struct BraggConst {
  double cK_p, cK_m, a_0, a_m;   // Delta(k) = tanh(a_0 + 4 (cK/cK_p)^2.5 + a_m (cK_m/cK)^2.5)
  double cos_wd, sin_wd;         // wind direction
  double sq_re, sq_im;           // sqrt(epsilon_sw), principal branch
  double dxdy;                   // float32(dx) * float32(dy) as the reference multiplies
  double inv_cK_p;               // 1 / cK_p
  double inv_cK_p25, cK_m25;     // cK_p^-2.5, cK_m^2.5
  double inv_2k0;                // 1 / (2 k0)
};

BraggConst bragg_const(const osar_bragg_desc& d) {
  BraggConst c{};
  const double U = d.wind_U;
  const double k_0 = kG / (U * U);
  const double X = k_0 * d.wind_fetch;
  const double Omega_c = 0.84 * std::pow(std::tanh(std::pow(X / kX0, 0.4)), -0.75);
  const double k_p = k_0 * Omega_c * Omega_c;
  c.cK_p = std::sqrt(kG / k_p + kSurfT / kRho * k_p);
  const double k_m = std::sqrt(kRho * kG / kSurfT);
  c.cK_m = std::sqrt(kG / k_m + kSurfT / kRho * k_m);
  const double C_10 = (0.8 + 0.065 * U) * 1e-3;
  const double ustar = std::sqrt(C_10) * U;
  c.a_0 = std::log(2.) / 4.;
  c.a_m = 0.13 * ustar / c.cK_m;
  c.cos_wd = std::cos(d.wind_dir);
  c.sin_wd = std::sin(d.wind_dir);
  const double eps_abs = std::hypot(73., 18.);
  c.sq_re = std::sqrt((eps_abs + 73.) / 2.);
  c.sq_im = 18. / (2. * c.sq_re);
  c.dxdy = (double)(float)d.dx * (double)(float)d.dy;
  c.inv_cK_p = 1. / c.cK_p;
  c.inv_cK_p25 = std::pow(c.cK_p, -2.5);
  c.cK_m25 = std::pow(c.cK_m, 2.5);
  c.inv_2k0 = 1. / (2. * d.k0);
  return c;
}

// (cos, sin) of atan(t): (1, t) / sqrt(1 + t^2). Beyond |t| = 1e18 (where
// t^2 would overflow float32) cos < 1e-18 is returned as 0 and sin as sign t,
// the limits atan(+-inf) = +-pi/2 has.
__device__ __forceinline__ void cos_sin_atan(float t, float* c, float* s) {
  const float r = rsqrtf(1.f + t * t);
  const bool huge = fabsf(t) >= 1e18f;
  *c = huge ? 0.f : r;
  *s = huge ? copysignf(1.f, t) : t * r;
}

// Grid point i of a table built as numpy.arange(x0, ., dk): x0 + i * dk with
// the product rounded before the sum, as numpy fills it.
__device__ __forceinline__ double arange_point(double x0, double dk, int i) {
#pragma clang fp contract(off)
  return x0 + (double)i * dk;
}

// Table point i: computed when the table is an exact arange grid (no load on
// the search path), read otherwise.
template <bool UNIFORM>
__device__ __forceinline__ double table_x(const double* xp, double x0, double dk, int i) {
  return UNIFORM ? arange_point(x0, dk, i) : xp[i];
}

// scipy interp1d linear with fill_value=1. The first index with xp[i] >= x
// (searchsorted side='left') is found from a guess on the table's mean step and
// corrected by a local walk, so any ascending table gives the exact index. The
// interval slope comes from the precomputed table when there is one.
template <bool UNIFORM>
__device__ double interp_fill1(double x, const double* xp, const double* fp, const double* sl,
                               int n, double x0, double dk, double inv_dk) {
  if (x < table_x<UNIFORM>(xp, x0, dk, 0) || x > table_x<UNIFORM>(xp, x0, dk, n - 1)) return 1.0;
  if (x != x) return x;  // NaN propagates, as in scipy
  int i = (int)fmin(fmax(ceil((x - x0) * inv_dk), 0.), (double)(n - 1));
  while (i < n - 1 && table_x<UNIFORM>(xp, x0, dk, i) < x) ++i;
  while (i > 0 && table_x<UNIFORM>(xp, x0, dk, i - 1) >= x) --i;
  const int idx = i < 1 ? 1 : i;
  const double xl = table_x<UNIFORM>(xp, x0, dk, idx - 1);
  const double yl = fp[idx - 1];
  const double slope =
      sl ? sl[idx] : (fp[idx] - yl) / (table_x<UNIFORM>(xp, x0, dk, idx) - xl);
  return slope * (x - xl) + yl;
}

constexpr int kThreads = 256;
constexpr int kMaxColsPerBlock = 256;   // columns one workgroup walks (column terms in LDS)

struct cf {
  float re, im;
};
__device__ __forceinline__ cf caddf(cf a, cf b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cf cmulf(cf a, cf b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ cf cscalef(cf a, float s) { return {a.re * s, a.im * s}; }
__device__ __forceinline__ float cabs2f(cf a) { return a.re * a.re + a.im * a.im; }

// Every angle of romeiser97.py:262-314 enters only
// through its sine and cosine, so atan/asin/sin/cos are replaced by the exact
// identities cos(atan t) = 1/sqrt(1+t^2), cos(asin x) = sqrt(1-x^2) and angle
// sums, and the wrapped spreading angle th of the two Bragg components by
// cos(th) (th and th + pi have opposite cosines). The half-plane test
// |th| < pi/2 becomes cos(th) > 0 (cos(th) < 0 for the cos^2 model's zero).
// The facet model runs in float32 (its inputs are float32 slopes; its result
// is a float32 amplitude): relative error a few 1e-7 against the FP64
// evaluation, except within that of the bad-facet threshold. Two terms take
// cancellation-free forms of the reference's expressions: sin^2(s_t) as
// v^2 / (1 + v^2) instead of 1 - cos^2(s_t), and sin^2(theta_i) = 1 -
// cos^2(theta_i) as sin^2(inc - s_p) + cos^2(inc - s_p) sin^2(s_t), the
// radicand of k_b. The azimuth terms of a row, the spectrum table search and
// the RCS sums stay FP64; the reference's k_b > 2 k0 rule is decided in FP64
// where the float32 radicand comes within 1e-5 of 1.
// Grid (workgroups per refresh, refreshes): refresh i = blockIdx.y reads the
// slopes of snapshot slot refresh_snap[i], writes amplitude slot i and adds
// its RCS sums to rcs_sum[i][pol] (one atomic per workgroup and pol). POL
// (0 vv, 1 hh, 2 DP) and SPREAD (0 Elfouhaily, 1 cos^2) are compile-time.
template <bool UNIFORM, int POL, int SPREAD>
// Seven waves per SIMD (at most 72 VGPRs, no spills): the facet loop waits on
// its slope loads for about 40 % of the time, and one more wave hides more of
// it than the register budget of six costs (15.3 against 15.5 ms per
// 512-refresh launch; eight waves need spills and take 17.9 ms).
__global__ __launch_bounds__(kThreads, 7) void facet_bragg_kernel(osar_bragg_desc d, BraggConst c,
                                                               int cols_per_block) {
  const size_t n = (size_t)d.nx * d.ny;
  constexpr int npol = POL == 2 ? 2 : 1;
  const int ri = blockIdx.y;
  const int pulse = d.n_refresh > 0 ? d.refresh_pulse[ri] : d.pulse;
  const size_t snap = d.n_refresh > 0 ? (size_t)(d.refresh_snap[ri] - d.snap0) : 0;
  const float* diffx_s = d.diffx + snap * d.diff_stride;
  const float* diffy_s = d.diffy + snap * d.diff_stride;
  const float* hmtf_s = d.hmtf ? d.hmtf + snap * d.diff_stride : nullptr;
  float* amp_s = d.amp + (size_t)ri * d.amp_stride;
  const double az_now = (pulse * d.t_step - d.t_span / 2.) * d.v_ground;
  double sums[2] = {0., 0.};
  // Thread = one facet row; the workgroup walks a range of columns of 256
  // consecutive rows (fields are x-major, idx = col * ny + row, so every
  // column is one contiguous read per wave). The row's azimuth terms are
  // computed once in FP64: az = asin(arg) (sar_raw_nompi.py:363) as (sin az,
  // cos az), and A = az - wind_dir as (cos A, sin A).
  const int nrb = (d.ny + kThreads - 1) / kThreads;
  const int row = (int)(blockIdx.x % nrb) * kThreads + threadIdx.x;
  const int cs = (int)(blockIdx.x / nrb);
  const int col_lo = cs * cols_per_block;
  const int col_hi = row < d.ny ? min(d.nx, col_lo + cols_per_block) : col_lo;
  const double arg = row < d.ny ? (d.y_row[row] - az_now) / d.gr0 : 0.;
  const double sin_az64 = fabs(arg) <= 1. ? arg : NAN;
  const double cos_az64 = sqrt(1. - arg * arg);
  const float sin_az = (float)sin_az64, cos_az = (float)cos_az64;
  const float cos_A = (float)(cos_az64 * c.cos_wd + sin_az64 * c.sin_wd);
  const float sin_A = (float)(sin_az64 * c.cos_wd - cos_az64 * c.sin_wd);
  // column terms of the workgroup's columns: cos, sin and 1 / cos of the incidence
  __shared__ float col_cos[kMaxColsPerBlock], col_sin[kMaxColsPerBlock],
      col_sec[kMaxColsPerBlock];
  for (int i = threadIdx.x; i < min(d.nx - col_lo, cols_per_block); i += kThreads) {
    const double ci = d.cos_inc[col_lo + i];
    col_cos[i] = (float)ci;
    col_sin[i] = (float)d.sin_inc[col_lo + i];
    col_sec[i] = (float)(1. / ci);
  }
  __syncthreads();
  const float k0 = (float)d.k0, two_k0 = (float)(2. * d.k0);
  const float inv_2k0 = (float)c.inv_2k0;
  const float dd4 = (float)(d.d * d.d) * 0.25f;
  const cf eps = {73.f, 18.f};
  const cf sq = {(float)c.sq_re, (float)c.sq_im};
  const cf eps2 = cmulf(eps, eps);
  const float a_0 = (float)c.a_0, a_m = (float)c.a_m;
  const float inv_cK_p25 = (float)c.inv_cK_p25, cK_m25 = (float)c.cK_m25;
  const float dxdy = (float)c.dxdy;
  for (int col = col_lo; col < col_hi; ++col) {
    const size_t idx = (size_t)col * d.ny + row;
    const float cos_inc = col_cos[col - col_lo], sin_inc = col_sin[col - col_lo];
    const float diffx = diffx_s[idx];
    const float diffy = diffy_s[idx];
    // parallel and orthogonal slopes s_p = atan(u), s_t = atan(v), with
    // 1 / cos(s_p) = (1 + u^2) cos(s_p)
    const float u = cos_az * diffx + sin_az * diffy;
    const float v = -sin_az * diffx + cos_az * diffy;
    const float uu = 1.f + u * u;
    const float r_sp = rsqrtf(uu);
    const float cos_sp = r_sp, sin_sp = u * r_sp, sec_sp = uu * r_sp;
    float cos_st, sin_st;
    cos_sin_atan(v, &cos_st, &sin_st);
    const float cos_s_t_2 = cos_st * cos_st;
    const float sin_s_t_2 = sin_st * sin_st;
    const float sin_ip = sin_inc * cos_sp - cos_inc * sin_sp;   // sin(inc - s_p)
    const float cos_ip = cos_inc * cos_sp + sin_inc * sin_sp;   // cos(inc - s_p)
    const float cos_ip_2 = cos_ip * cos_ip;
    const float sin_ip_2 = sin_ip * sin_ip;
    const float cos_ti = cos_ip * cos_st;
    // sin^2(theta_i), the Bragg radicand: k_b = 2 k0 sqrt(zb)
    const float zb = sin_ip_2 + cos_ip_2 * sin_s_t_2;
    const float sin_ti_2 = zb;
    // b_hh = eps / d1^2 and b_vv = eps^2 (1 + sin_ti^2) / d2^2 as u conj(q) / |q|^2
    // (q = d^2); the reciprocals are taken below, with the facet's others
    const cf d1 = caddf({cos_ti, 0.f}, sq), d2 = caddf(cscalef(eps, cos_ti), sq);
    const cf q1 = cmulf(d1, d1), q2 = cmulf(d2, d2);
    const float n1 = cabs2f(q1), n2 = cabs2f(q2);
    const cf u_hh = cmulf(eps, {q1.re, -q1.im});
    const cf u_vv = cmulf(cscalef(eps2, 1.f + sin_ti_2), {q2.re, -q2.im});
    const float wgt = cos_ip * (col_sec[col - col_lo] * sec_sp);   // cos_ip / (cos_inc cos_sp)
    const float kc = k0 * cos_ti;
    const float F0 = 8.f * (float)M_PI * (kc * kc) * (kc * kc);
    // k_b and 1 / k_b from one reciprocal square root
    const float rz = rsqrtf(zb);   // +inf at zb = 0
    float k_b = zb > 0.f ? two_k0 * (zb * rz) : 0.f;
    float inv_kb = rz * inv_2k0;
    if (zb > 0.99999f) {
      // the reference's k_b > 2 k0 rule (romeiser97.py:307), decided in FP64
      const double u64 = (double)u, v64 = (double)v;
      const double rs = 1. / sqrt(1. + u64 * u64), rt = 1. / sqrt(1. + v64 * v64);
      const double ci = d.cos_inc[col], si = d.sin_inc[col];
      const double sip = si * rs - ci * (u64 * rs), cip = ci * rs + si * (u64 * rs);
      const double st2 = fabs(v64) >= 1e150 ? 1. : (v64 * rt) * (v64 * rt);
      if (2. * d.k0 * sqrt(sip * sip + cip * cip * st2) > 2. * d.k0) {
        k_b = 0.f;
        inv_kb = INFINITY;
      }
    }
    // phi_b = az + atan(q), q = -cos_ip sin_st / sin_ip; th0 = phi_b - wind_dir;
    // cos(th0) from angle sums. 1 + q^2 = zb / sin_ip^2, so cos(atan q) =
    // |sin_ip| / sqrt(zb) and sin(atan q) = -cos_ip sin_st sign(sin_ip) / sqrt(zb)
    const float cos_q = fabsf(sin_ip) * rz;
    const float sin_q = -cos_ip * sin_st * copysignf(rz, sin_ip);
    const float cth0 = cos_A * cos_q - sin_A * sin_q;
    const bool bad = sin_ip_2 < dd4;
    const float S1 = (float)interp_fill1<UNIFORM>((double)k_b, d.spec_k, d.spec_S, d.spec_slope,
                                                  d.n_spec, d.spec_k0, d.spec_dk, d.spec_inv_dk);
    float x = 0.f, e = 0.f;
    if (SPREAD == 0) {
      const float y = (float)kG * inv_kb + (float)(kSurfT / kRho) * k_b;
      const float cK = sqrtf(y);
      // (cK / cK_p)^2.5 = y sqrt(cK) cK_p^-2.5 and (cK_m / cK)^2.5 = cK_m^2.5
      // cK^-2.5, from r = 1 / sqrt(cK); k_b = 0 gives x = +inf
      const float r = rsqrtf(cK);
      const float r2 = r * r;
      x = a_0 + 4.f * (y * (cK * r) * inv_cK_p25) + a_m * (cK_m25 * (r2 * r2 * r));
      if (!(y < INFINITY)) x = INFINITY;
      // tanh(x) = sign(x) (1 - e) / (1 + e), e = exp(-2 |x|) (x >= ln(2) / 4 here)
      e = expf(-2.f * fabsf(x));
    }
    // One division gives 1 / |q1|^2, 1 / |q2|^2, 1 / sin_ti^2 and (Elfouhaily)
    // 1 / (1 + e): |q| is between 8^2 and 100^2, 1 + e in [1, 2], so the
    // product stays within float32 range for sin_ti^2 above 1e-20
    const float ope = SPREAD == 0 ? 1.f + e : 1.f;
    const float n12 = n1 * n2;
    const float inv = 1.f / (n12 * (sin_ti_2 * ope));
    const float inv_ti = (n12 * ope) * inv;
    const float inv_n12 = (sin_ti_2 * ope) * inv;
    const cf b_hh = cscalef(u_hh, n2 * inv_n12), b_vv = cscalef(u_vv, n1 * inv_n12);
    const float delta = SPREAD == 0 ? copysignf((1.f - e) * ((n12 * sin_ti_2) * inv), x) : 0.f;
    const float a = sin_s_t_2 * inv_ti;
    const float b = sin_ip_2 * cos_s_t_2 * inv_ti;
    const float T_vv = wgt * (F0 * cabs2f(caddf(cscalef(b_vv, b), cscalef(b_hh, a))));
    const float T_hh = wgt * (F0 * cabs2f(caddf(cscalef(b_hh, b), cscalef(b_vv, a))));
    const float base = S1 * inv_kb;
    // hydrodynamic modulation: 1 + hMTF with hMTF below -1 set to -1
    float mod = 1.f;
    if (hmtf_s) {
      const float h = hmtf_s[idx];
      mod = 1.f + (h < -1.f ? -1.f : h);
    }
#pragma unroll
    for (int comp = 0; comp < 2; ++comp) {
      const float cth = comp == 0 ? cth0 : -cth0;
      float g;
      constexpr float kInvPi = 0.31830988618379067154f;
      if (SPREAD == 0) g = cth > 0.f ? (1.f + delta * (2.f * cth * cth - 1.f)) * kInvPi : 0.f;
      else g = cth < 0.f ? 0.f : kInvPi * cth * cth;
      const float spectrum = base * g;
      // pol 2 (DP): slot 0 = hh, slot 1 = vv
#pragma unroll
      for (int ip = 0; ip < npol; ++ip) {
        const bool is_hh = (POL == 1) || (POL == 2 && ip == 0);
        const float T = is_hh ? T_hh : T_vv;
        float rcs = T * spectrum * dxdy;
        // romeiser97.py:366-392 (DP: hh masks only the plus component)
        const bool mask = bad && !(POL == 2 && is_hh && comp == 0);
        if (mask) rcs = 0.f;
        rcs = rcs * mod;
        amp_s[((size_t)ip * 2 + comp) * n + idx] = sqrtf(rcs);
        sums[ip] += (double)rcs;
      }
    }
  }
  // wave, then workgroup reduction of the RCS sums
  __shared__ double red[kThreads / 64][2];
  for (int ip = 0; ip < 2; ++ip)
    for (int off = 32; off > 0; off >>= 1) sums[ip] += __shfl_xor(sums[ip], off);
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[wave][0] = sums[0];
    red[wave][1] = sums[1];
  }
  __syncthreads();
  if (threadIdx.x == 0 && d.rcs_sum) {
    for (int ip = 0; ip < npol; ++ip) {
      double v = 0.;
      for (int w = 0; w < kThreads / 64; ++w) v += red[w][ip];
      atomicAdd(&d.rcs_sum[(size_t)ri * npol + ip], v);
    }
  }
}

// ---- FP64 facet model (osar_facet_bragg with fp64 = 1): the same
// identities as the float32 kernel, every value FP64, amplitudes written as
// FP64 square roots; the reference-precision path the float32 one is measured
// against.
struct cd {
  double re, im;
};
__device__ __forceinline__ cd cadd(cd a, cd b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cd cmulc(cd a, cd b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ cd cscale(cd a, double s) { return {a.re * s, a.im * s}; }
__device__ __forceinline__ double cabs2(cd a) { return a.re * a.re + a.im * a.im; }

// (cos, sin) of atan(t) in FP64; beyond |t| = 1e150 the limits of atan(+-inf).
__device__ __forceinline__ void cos_sin_atan64(double t, double* c, double* s) {
  const double r = rsqrt(1. + t * t);
  const bool huge = fabs(t) >= 1e150;
  *c = huge ? 0. : r;
  *s = huge ? copysign(1., t) : t * r;
}

template <bool UNIFORM, int POL, int SPREAD>
__global__ __launch_bounds__(kThreads) void facet_bragg_fp64_kernel(osar_bragg_desc d, BraggConst c,
                                                               int cols_per_block) {
  const size_t n = (size_t)d.nx * d.ny;
  constexpr int npol = POL == 2 ? 2 : 1;
  const int ri = blockIdx.y;
  const int pulse = d.n_refresh > 0 ? d.refresh_pulse[ri] : d.pulse;
  const size_t snap = d.n_refresh > 0 ? (size_t)(d.refresh_snap[ri] - d.snap0) : 0;
  const float* diffx_s = d.diffx + snap * d.diff_stride;
  const float* diffy_s = d.diffy + snap * d.diff_stride;
  const float* hmtf_s = d.hmtf ? d.hmtf + snap * d.diff_stride : nullptr;
  double* amp_s = d.amp64 + (size_t)ri * d.amp_stride;
  const double az_now = (pulse * d.t_step - d.t_span / 2.) * d.v_ground;
  double sums[2] = {0., 0.};
  // Thread = one facet row; the workgroup walks a range of columns of 256
  // consecutive rows (fields are x-major, idx = col * ny + row, so every
  // column is one contiguous read per wave). The row's azimuth terms are
  // computed once: az = asin(arg) (sar_raw_nompi.py:363) as (sin az, cos az),
  // and A = az - wind_dir as (cos A, sin A).
  const int nrb = (d.ny + kThreads - 1) / kThreads;
  const int row = (int)(blockIdx.x % nrb) * kThreads + threadIdx.x;
  const int cs = (int)(blockIdx.x / nrb);
  const int col_lo = cs * cols_per_block;
  const int col_hi = row < d.ny ? min(d.nx, col_lo + cols_per_block) : col_lo;
  const double arg = row < d.ny ? (d.y_row[row] - az_now) / d.gr0 : 0.;
  const double sin_az = fabs(arg) <= 1. ? arg : NAN;
  const double cos_az = sqrt(1. - arg * arg);
  const double cos_A = cos_az * c.cos_wd + sin_az * c.sin_wd;
  const double sin_A = sin_az * c.cos_wd - cos_az * c.sin_wd;
  __shared__ double sec_inc[kMaxColsPerBlock];   // 1 / cos(inc) of the workgroup's columns
  for (int i = threadIdx.x; i < min(d.nx - col_lo, cols_per_block); i += kThreads)
    sec_inc[i] = 1. / d.cos_inc[col_lo + i];
  __syncthreads();
  for (int col = col_lo; col < col_hi; ++col) {
    const size_t idx = (size_t)col * d.ny + row;
    const double cos_inc = d.cos_inc[col], sin_inc = d.sin_inc[col];
    const double diffx = (double)diffx_s[idx];
    const double diffy = (double)diffy_s[idx];
    // parallel and orthogonal slopes s_p = atan(u), s_t = atan(v) (|u|, |v| are
    // below 1e39: float32 slopes), with 1 / cos(s_p) = (1 + u^2) cos(s_p)
    const double u = cos_az * diffx + sin_az * diffy;
    const double r_sp = rsqrt(1. + u * u);
    const double cos_sp = r_sp, sin_sp = u * r_sp, sec_sp = (1. + u * u) * r_sp;
    double cos_st, sin_st;
    cos_sin_atan64(-sin_az * diffx + cos_az * diffy, &cos_st, &sin_st);
    const double cos_s_t_2 = cos_st * cos_st;
    const double sin_s_t_2 = 1 - cos_s_t_2;
    const double sin_ip = sin_inc * cos_sp - cos_inc * sin_sp;   // sin(inc - s_p)
    const double cos_ip = cos_inc * cos_sp + sin_inc * sin_sp;   // cos(inc - s_p)
    const double cos_ip_2 = cos_ip * cos_ip;
    const double sin_ip_2 = sin_ip * sin_ip;
    const double cos_ti = cos_ip * cos_st;
    const double sin_ti_2 = 1 - cos_ti * cos_ti;
    const cd eps = {73., 18.};
    const cd sq = {c.sq_re, c.sq_im};
    // b_hh = eps / d1^2 and b_vv = eps^2 (1 + sin_ti^2) / d2^2 as u conj(q) / |q|^2
    // (q = d^2); the reciprocals are taken below, with the facet's others
    const cd d1 = cadd({cos_ti, 0.}, sq), d2 = cadd(cscale(eps, cos_ti), sq);
    const cd q1 = cmulc(d1, d1), q2 = cmulc(d2, d2);
    const double n1 = cabs2(q1), n2 = cabs2(q2);
    const cd u_hh = cmulc(eps, {q1.re, -q1.im});
    const cd u_vv = cmulc(cscale(cmulc(eps, eps), 1. + sin_ti_2), {q2.re, -q2.im});
    const double wgt = cos_ip * (sec_inc[col - col_lo] * sec_sp);   // cos_ip / (cos_inc cos_sp)
    const double kc = d.k0 * cos_ti;
    const double F0 = 8 * M_PI * (kc * kc) * (kc * kc);
    // k_b = 2 k0 sqrt(zb) and 1 / k_b from one reciprocal square root
    const double zb = sin_ip_2 + cos_ip_2 * sin_s_t_2;
    const double rz = rsqrt(zb);   // +inf at zb = 0
    double k_b = zb > 0. ? 2 * d.k0 * (zb * rz) : 0.;
    double inv_kb = rz * c.inv_2k0;
    if (k_b > 2 * d.k0) {
      k_b = 0.;
      inv_kb = INFINITY;
    }
    // phi_b = az + atan(q), q = -cos_ip sin_st / sin_ip; th0 = phi_b - wind_dir;
    // cos(th0) from angle sums. 1 + q^2 = zb / sin_ip^2, so cos(atan q) =
    // |sin_ip| / sqrt(zb) and sin(atan q) = -cos_ip sin_st sign(sin_ip) / sqrt(zb)
    const double cos_q = fabs(sin_ip) * rz;
    const double sin_q = -cos_ip * sin_st * copysign(rz, sin_ip);
    const double cth0 = cos_A * cos_q - sin_A * sin_q;
    const bool bad = 4 * sin_ip_2 < d.d * d.d;
    const double S1 = interp_fill1<UNIFORM>(k_b, d.spec_k, d.spec_S, d.spec_slope, d.n_spec,
                                            d.spec_k0, d.spec_dk, d.spec_inv_dk);
    double x = 0., e = 0.;
    if (SPREAD == 0) {
      const double y = kG * inv_kb + kSurfT / kRho * k_b;
      const double cK = sqrt(y);
      // (cK / cK_p)^2.5 = y sqrt(cK) cK_p^-2.5 and (cK_m / cK)^2.5 = cK_m^2.5
      // cK^-2.5, from r = 1 / sqrt(cK); k_b = 0 gives x = +inf as before
      const double r = rsqrt(cK);
      const double r2 = r * r;
      x = c.a_0 + 4. * (y * (cK * r) * c.inv_cK_p25) + c.a_m * (c.cK_m25 * (r2 * r2 * r));
      if (!(y < INFINITY)) x = INFINITY;
      // tanh(x) = sign(x) (1 - e) / (1 + e), e = exp(-2 |x|): one exponential,
      // no branches, within 3 ulp of the library tanh (x >= ln(2) / 4 here)
      e = exp(-2. * fabs(x));
    }
    // One division gives 1 / |q1|^2, 1 / |q2|^2, 1 / sin_ti^2 and (Elfouhaily)
    // 1 / (1 + e): |q| is between 8^4 and 100^4, 1 + e in [1, 2], so the
    // product neither overflows nor underflows; each reciprocal is within a few
    // ulp of its own division
    const double ope = SPREAD == 0 ? 1. + e : 1.;
    const double n12 = n1 * n2;
    const double inv = 1. / (n12 * (sin_ti_2 * ope));
    const double inv_ti = (n12 * ope) * inv;
    const double inv_n12 = (sin_ti_2 * ope) * inv;
    const cd b_hh = cscale(u_hh, n2 * inv_n12), b_vv = cscale(u_vv, n1 * inv_n12);
    const double delta = SPREAD == 0 ? copysign((1. - e) * ((n12 * sin_ti_2) * inv), x) : 0.;
    const double a = sin_s_t_2 * inv_ti;
    const double b = sin_ip_2 * cos_s_t_2 * inv_ti;
    const double T_vv = wgt * (F0 * cabs2(cadd(cscale(b_vv, b), cscale(b_hh, a))));
    const double T_hh = wgt * (F0 * cabs2(cadd(cscale(b_hh, b), cscale(b_vv, a))));
    const double base = S1 * inv_kb;
    // hydrodynamic modulation: 1 + hMTF (float32) with hMTF below -1 set to -1
    double mod = 1.;
    if (hmtf_s) {
      const float h = hmtf_s[idx];
      mod = (double)(1.f + (h < -1.f ? -1.f : h));
    }
    for (int comp = 0; comp < 2; ++comp) {
      const double cth = comp == 0 ? cth0 : -cth0;
      double g;
      constexpr double kInvPi = 0.31830988618379067154;
      if (SPREAD == 0) g = cth > 0. ? (1. + delta * (2. * cth * cth - 1.)) * kInvPi : 0.;
      else g = cth < 0. ? 0. : kInvPi * cth * cth;
      const double spectrum = base * g;
      // pol 2 (DP): slot 0 = hh, slot 1 = vv
      for (int ip = 0; ip < npol; ++ip) {
        const bool is_hh = (POL == 1) || (POL == 2 && ip == 0);
        const double T = is_hh ? T_hh : T_vv;
        double rcs = T * spectrum * c.dxdy;
        // romeiser97.py:366-392 (DP: hh masks only the plus component)
        const bool mask = bad && !(POL == 2 && is_hh && comp == 0);
        if (mask) rcs = 0.;
        rcs = rcs * mod;
        amp_s[((size_t)ip * 2 + comp) * n + idx] = sqrt(rcs);
        sums[ip] += rcs;
      }
    }
  }
  // wave, then workgroup reduction of the RCS sums
  __shared__ double red[kThreads / 64][2];
  for (int ip = 0; ip < 2; ++ip)
    for (int off = 32; off > 0; off >>= 1) sums[ip] += __shfl_xor(sums[ip], off);
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[wave][0] = sums[0];
    red[wave][1] = sums[1];
  }
  __syncthreads();
  if (threadIdx.x == 0 && d.rcs_sum) {
    for (int ip = 0; ip < npol; ++ip) {
      double v = 0.;
      for (int w = 0; w < kThreads / 64; ++w) v += red[w][ip];
      atomicAdd(&d.rcs_sum[(size_t)ri * npol + ip], v);
    }
  }
}

template <bool FP64, bool UNIFORM, int POL>
void launch_bragg_pol(int spread, dim3 grid, hipStream_t s, const osar_bragg_desc& d,
                      const BraggConst& c, int cpb) {
  if (FP64) {
    if (spread == 0)
      hipLaunchKernelGGL((facet_bragg_fp64_kernel<UNIFORM, POL, 0>), grid, dim3(kThreads), 0, s,
                         d, c, cpb);
    else
      hipLaunchKernelGGL((facet_bragg_fp64_kernel<UNIFORM, POL, 1>), grid, dim3(kThreads), 0, s,
                         d, c, cpb);
    return;
  }
  if (spread == 0)
    hipLaunchKernelGGL((facet_bragg_kernel<UNIFORM, POL, 0>), grid, dim3(kThreads), 0, s, d, c,
                       cpb);
  else
    hipLaunchKernelGGL((facet_bragg_kernel<UNIFORM, POL, 1>), grid, dim3(kThreads), 0, s, d, c,
                       cpb);
}

template <bool FP64, bool UNIFORM>
void launch_bragg_u(int pol, int spread, dim3 grid, hipStream_t s, const osar_bragg_desc& d,
                    const BraggConst& c, int cpb) {
  if (pol == 0) launch_bragg_pol<FP64, UNIFORM, 0>(spread, grid, s, d, c, cpb);
  else if (pol == 1) launch_bragg_pol<FP64, UNIFORM, 1>(spread, grid, s, d, c, cpb);
  else launch_bragg_pol<FP64, UNIFORM, 2>(spread, grid, s, d, c, cpb);
}

void launch_bragg(bool fp64, bool uniform, int pol, int spread, dim3 grid, hipStream_t s,
                  const osar_bragg_desc& d, const BraggConst& c, int cpb) {
  if (fp64) {
    if (uniform) launch_bragg_u<true, true>(pol, spread, grid, s, d, c, cpb);
    else launch_bragg_u<true, false>(pol, spread, grid, s, d, c, cpb);
  } else {
    if (uniform) launch_bragg_u<false, true>(pol, spread, grid, s, d, c, cpb);
    else launch_bragg_u<false, false>(pol, spread, grid, s, d, c, cpb);
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
  OSAR_PTR(d->cos_inc);
  OSAR_PTR(d->sin_inc);
  OSAR_PTR(d->y_row);
  OSAR_PTR(d->spec_k);
  OSAR_PTR(d->spec_S);
  OSAR_PTR(d->diffx);
  OSAR_PTR(d->diffy);
  if (d->fp64) OSAR_PTR(d->amp64);
  else OSAR_PTR(d->amp);
  OSAR_HIP(hipSetDevice(ctx->device));
  OSAR_REQUIRE(!d->spec_uniform || d->spec_dk > 0., OSAR_EINVAL,
               "spec_uniform needs spec_dk > 0");
  const int n_ref = d->n_refresh > 0 ? d->n_refresh : 1;
  OSAR_REQUIRE(n_ref <= 65535, OSAR_EINVAL, "n_refresh %d exceeds 65535", n_ref);
  if (d->n_refresh > 0) {
    OSAR_PTR(d->refresh_pulse);
    OSAR_PTR(d->refresh_snap);
    OSAR_REQUIRE(d->diff_stride >= (int64_t)d->nx * d->ny || d->n_refresh == 1, OSAR_EINVAL,
                 "diff_stride smaller than a field");
    OSAR_REQUIRE(d->amp_stride >= 2 * (int64_t)d->nx * d->ny * (d->pol == 2 ? 2 : 1) ||
                     d->n_refresh == 1,
                 OSAR_EINVAL, "amp_stride smaller than one refresh's amplitudes");
  }
  const size_t n = (size_t)d->nx * d->ny;
  // about 8 workgroups per CU in total, at least 16 per refresh, none idle
  size_t per = ((size_t)ctx->n_cu * 8 + n_ref - 1) / n_ref;
  if (per < 16) per = 16;
  const size_t want = (n + kThreads - 1) / kThreads;
  if (per > want) per = want;
  const BraggConst c = bragg_const(*d);
  // workgroups per refresh: row blocks x column ranges, about 8 per CU in all
  const size_t nrb = ((size_t)d->ny + kThreads - 1) / kThreads;
  size_t ncs = (per + nrb - 1) / nrb;
  if (ncs > (size_t)d->nx) ncs = d->nx;
  const size_t ncs_min = ((size_t)d->nx + kMaxColsPerBlock - 1) / kMaxColsPerBlock;
  if (ncs < ncs_min) ncs = ncs_min;
  const int cpb = (int)(((size_t)d->nx + ncs - 1) / ncs);
  ncs = ((size_t)d->nx + cpb - 1) / cpb;
  const dim3 grid((unsigned)(nrb * ncs), (unsigned)n_ref);
  launch_bragg(d->fp64 != 0, d->spec_uniform != 0, d->pol, d->spread, grid, (hipStream_t)stream,
               *d, c, cpb);
  OSAR_LAUNCHED("facet_bragg_kernel");
  return OSAR_OK;
}
