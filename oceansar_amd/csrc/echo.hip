// Echo accumulation: per-pulse facet geometry, Bragg scene, beam, channel
// phases and the 8-tap range deposit, fused in one kernel.
//
// Replaces the body of the sar_raw pulse loop (oceansar/radarsim/
// sar_raw_nompi.py:349-535) from the surface fields to the range lines,
// including chan_profile_numba / profile_integrator
// (oceansar/radarsim/range_profile.py:22-49,83-116).
//
// Work decomposition: a workgroup owns P consecutive pulses and R rows of
// facets. Each wavefront lane is one (facet, pulse) pair: 64/P consecutive
// rows of one column times P pulses (fields are stored x-major, so those
// facets are contiguous), so the fields of a facet are loaded once for P
// pulses, per-column quantities are wave-uniform, and every lane of an
// instruction deposits into a different (pulse) line or, within a line, into
// a tap rotated by the facet's lane index: facets of one column sit in nearly
// the same range bin, and the rotation keeps their LDS atomics on different
// addresses.
// Range lines of the P pulses (all channels and polarisations) live in an LDS
// window of W bins; they are added to the raw cube with one atomic pass per
// workgroup. Taps that fall outside the window (but inside the line) go
// straight to global memory; taps outside [0, nbins) are dropped as in the
// reference.
//
// Precision: slant range, bin position, phases and the lerp of the
// displacement fields are FP64 in the reference's operation order (FP
// contraction off), so bin and sub-bin indices match an FP64 evaluation; the
// phase is reduced to revolutions in FP64 before FP32 sin/cos; scene values,
// channel factors and the accumulation are FP32.
#include "common.h"

namespace {

constexpr int kThreads = 256;
constexpr int kTaps = 8;
constexpr int kOvs = 20;
constexpr int kTab = kTaps * kOvs + 1;  // 161

__device__ __forceinline__ double np_sinc(double x) {
  // numpy.sinc: sin(pi x)/(pi x) with x = 0 replaced by 1e-20
  double y = M_PI * (x == 0.0 ? 1.0e-20 : x);
  return sin(y) / y;
}

struct Geo {
  double srs;   // slant range of the facet minus min(sr)
  double dinc;  // wave_dinc (for the cross-track phase)
};

// sr_surface and wave_dinc (sar_raw_nompi.py:366-369), in the reference's order.
__device__ __forceinline__ Geo facet_range(double sr, double cosi, double sini, double mig,
                                           double sin_az, double Dx, double Dy, double Dz,
                                           double sr0) {
#pragma clang fp contract(off)
  Geo g;
  double a = sr - cosi * Dz;
  a = a + mig;
  a = a + Dx * sini;
  a = a + Dy * sin_az;
  g.srs = a;
  g.dinc = (Dz * sini + Dx * sini) / sr0;
  return g;
}

__device__ __forceinline__ double lerp_field(const float* f, size_t o0, size_t step, int rk,
                                             double w) {
#pragma clang fp contract(off)
  double a = (double)f[o0];
  if (rk == 0) return a;
  double b = (double)f[o0 + step];
  return (1. - w) * a + w * b;
}

// ---- Kirchhoff stationary-phase specular field (RCSKA.field, 'spa', monostatic;
// oceansar/nrcs/ka.py:131-270), FP64, in the reference's operation order.
struct zd {
  double re, im;
};
__device__ __forceinline__ zd zmul(zd a, zd b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ zd zscl(zd a, double s) { return {a.re * s, a.im * s}; }
__device__ __forceinline__ zd zadd(zd a, zd b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ zd zsub(zd a, zd b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ zd zdiv(zd a, zd b) {
  if (fabs(b.re) >= fabs(b.im)) {
    const double rat = b.im / b.re, scl = 1.0 / (b.re + b.im * rat);
    return {(a.re + a.im * rat) * scl, (a.im - a.re * rat) * scl};
  }
  const double rat = b.re / b.im, scl = 1.0 / (b.im + b.re * rat);
  return {(a.re * rat + a.im) * scl, (a.im * rat - a.re) * scl};
}
__device__ __forceinline__ zd zsqrt(zd z) {
  // principal square root
  const double m = hypot(z.re, z.im);
  if (z.re >= 0.) {
    const double t = sqrt(0.5 * (m + z.re));
    return {t, z.im / (2. * t)};
  }
  const double t = sqrt(0.5 * (m - z.re));
  return {fabs(z.im) / (2. * t), z.im >= 0. ? t : -t};
}
struct v3 {
  double x, y, z;
};
__device__ __forceinline__ v3 cross3(v3 a, v3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double dot3(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
struct zv3 {
  zd x, y, z;
};
__device__ __forceinline__ zv3 zv(zd s, v3 v) {
  return {zscl(s, v.x), zscl(s, v.y), zscl(s, v.z)};
}
__device__ __forceinline__ zv3 zvsub(zv3 a, zv3 b) {
  return {zsub(a.x, b.x), zsub(a.y, b.y), zsub(a.z, b.z)};
}
__device__ __forceinline__ zv3 zvadd(zv3 a, zv3 b) {
  return {zadd(a.x, b.x), zadd(a.y, b.y), zadd(a.z, b.z)};
}
// cross(real a, complex b)
__device__ __forceinline__ zv3 rzcross(v3 a, zv3 b) {
  return {zsub(zscl(b.z, a.y), zscl(b.y, a.z)), zsub(zscl(b.x, a.z), zscl(b.z, a.x)),
          zsub(zscl(b.y, a.x), zscl(b.x, a.y))};
}

// Returns true for a stationary-phase facet and writes E for the requested
// polarisations (pol_code 0 = vv, 1 = hh, 2 = both: e[0] = hh, e[1] = vv).
__device__ bool ka_spa(double k0, int pol_code, double th, double ph, double Dx, double Dy,
                       double Dxx, double Dyy, double Dxy, double fdx, double fdy, zd* e) {
#pragma clang fp contract(off)
  const double sin_t = sin(th), cos_t = cos(th), tan_t = tan(th);
  const double sin_p = sin(ph), cos_p = cos(ph);
  double J = Dxx * Dyy - Dxy * Dxy;
  if (J == 0.) return false;                           // reference: J -> NaN, never sps
  const double J_abs = fabs(J);
  const double delta_x = (1. / J_abs) * (Dxy * (Dy - tan_t * sin_p) - Dyy * (Dx - tan_t * cos_p));
  const double delta_y = (1. / J_abs) * (Dxy * (Dx - tan_t * cos_p) - Dxx * (Dy - tan_t * sin_p));
  if (!((-0.5 * fdx < delta_x) && (delta_x < 0.5 * fdx) && (-0.5 * fdy < delta_y) &&
        (delta_y < 0.5 * fdy)))
    return false;
  const double delta_z = delta_x * Dx + delta_y * Dy;
  const zd eps_sgn = J > 0. ? zd{Dxx > 0. ? 1. : (Dxx < 0. ? -1. : 0.), 0.} : zd{0., 1.};
  const double n_norm = sqrt(Dx * Dx + Dy * Dy + 1.);
  const v3 n = {-Dx / n_norm, -Dy / n_norm, 1. / n_norm};
  const v3 n_i = {sin_t * cos_p, sin_t * sin_p, -cos_t};
  const v3 n_s = {-sin_t * cos_p, -sin_t * sin_p, cos_t};
  const v3 q = {k0 * (n_s.x - n_i.x), k0 * (n_s.y - n_i.y), k0 * (n_s.z - n_i.z)};
  v3 t = cross3(n_i, n);
  const double tn = sqrt((t.x * t.x + t.y * t.y) + t.z * t.z);
  t = {t.x / tn, t.y / tn, t.z / tn};
  const v3 dd = cross3(n_i, t);
  const double cos_l = -dot3(n, n_i);
  const zd eps = {73., 18.};
  const zd se = zsqrt({73. + cos_l * cos_l - 1., 18.});
  const zd ec = zscl(eps, cos_l);
  const zd r_v = zdiv(zsub(ec, se), zadd(ec, se));
  const zd r_h = zdiv(zsub({cos_l, 0.}, se), zadd({cos_l, 0.}, se));
  const double n_p_ni = -cos_l;
  const v3 n_x_t = cross3(n, t);
  const v3 h_i = {-sin_p, cos_p, 0.};
  const v3 v_i = {-cos_t * cos_p, -cos_t * sin_p, -sin_t};
  const zd K = zdiv(zmul({0., -k0}, {cos(-k0 * 2.), sin(-k0 * 2.)}),
                    {(4. * M_PI) * (4. * M_PI), 0.});
  const double qd = (q.x * delta_x + q.y * delta_y) + q.z * delta_z;
  const zd eq = {cos(qd), sin(qd)};
  const double tail = 2. * M_PI / q.z * n_norm / sqrt(J_abs);
  const int first = pol_code == 2 ? 0 : pol_code;   // 0 = vv, 1 = hh in single-pol mode
  const int npol = pol_code == 2 ? 2 : 1;
  for (int ip = 0; ip < npol; ++ip) {
    const bool is_h = pol_code == 2 ? (ip == 0) : (first == 1);
    const v3 a = is_h ? h_i : v_i;
    const double a_t = dot3(a, t), a_d = dot3(a, dd);
    // n x Es = ((1 + r_h) a_t) n_x_t - ((1 - r_v) n_p_ni a_d) t
    const zv3 nxEs = zvsub(zv(zscl(zadd({1., 0.}, r_h), a_t), n_x_t),
                           zv(zscl(zscl(zsub({1., 0.}, r_v), n_p_ni), a_d), t));
    // eta n x Hs = -(((1 - r_h) n_p_ni a_t) t + ((1 + r_v) a_d) n_x_t)
    const zv3 eHs0 = zvadd(zv(zscl(zscl(zsub({1., 0.}, r_h), n_p_ni), a_t), t),
                           zv(zscl(zadd({1., 0.}, r_v), a_d), n_x_t));
    const zv3 eHs = {zscl(eHs0.x, -1.), zscl(eHs0.y, -1.), zscl(eHs0.z, -1.)};
    const zv3 inner = zvsub(nxEs, rzcross(n_s, eHs));
    const zv3 c = rzcross(n_s, inner);
    const zd pp = zadd(zadd(zscl(c.x, a.x), zscl(c.y, a.y)), zscl(c.z, a.z));
    zd E = zmul(zmul(zmul(eps_sgn, pp), K), eq);
    E = zscl(E, tail);
    e[ip] = E;
  }
  return true;
}

template <int P, bool SPEC>
__global__ __launch_bounds__(kThreads) void echo_accumulate_kernel(osar_echo_desc d, int W,
                                                                   int R) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const int npol = d.n_pol, nch = d.n_ch;
  const int nlines = npol * nch * P;
  double* geo = (double*)lds_raw;                        // [P][R][3] sin_az, mig, az_proj
  float2* F = (float2*)(geo + 3 * P * R);                // [P][nch][R]
  float* acc = (float*)(F + P * nch * R);                // [nlines][2][W]
  float* tab = acc + (size_t)nlines * 2 * W;             // [kTab]
  int* wlo = (int*)(tab + 164);                          // [P]

  const int tid = threadIdx.x;
  const int pstart = d.pulse0 + blockIdx.x * P;
  const int pend = d.pulse0 + d.npulses;
  const int y0 = blockIdx.y * R;
  const int rows = min(R, d.ny - y0);

  for (int i = tid; i < nlines * 2 * W; i += kThreads) acc[i] = 0.f;
  for (int i = tid; i < kTab; i += kThreads) tab[i] = d.sinc_tab[i];
  for (int i = tid; i < P * R; i += kThreads) {
    const int pi = i / R, r = i - pi * R;
    const int p = pstart + pi;
    double sin_az = 0., mig = 0., az_proj = 0.;
    double beam = 0.;
    if (p < pend && r < rows) {
      const double t_now = p * d.t_step;
      const double az_now = (t_now - d.t_span / 2.) * d.v_ground;
      const double az = d.y_row[y0 + r] - az_now;
      sin_az = az / d.sr0;
      mig = (az / 2) * sin_az;
      if (SPEC) az_proj = asin(az / d.gr0);
      beam = np_sinc(d.ltx_over_l * sin_az) * np_sinc(d.lrx_over_l * sin_az);
    }
    geo[3 * i] = sin_az;
    geo[3 * i + 1] = mig;
    geo[3 * i + 2] = az_proj;
    for (int ch = 0; ch < nch; ++ch) {
      const double ph = -(d.k0 * d.b_ati[ch]) * sin_az;
      double s, c;
      sincos(ph, &s, &c);
      F[(pi * nch + ch) * R + r] = make_float2((float)(beam * c), (float)(beam * s));
    }
  }
  __syncthreads();
  if (tid < P) {
    double mmin = 1e300, mmax = -1e300;
    for (int r = 0; r < rows; ++r) {
      const double m = geo[3 * (tid * R + r) + 1];
      mmin = fmin(mmin, m);
      mmax = fmax(mmax, m);
    }
    const double lo = (d.sr_col_min + mmin - d.resid_max - d.min_sr) / d.binsize;
    wlo[tid] = max((int)floor(lo) - 5, 0);
  }
  __syncthreads();

  // Lane = (facet, pulse): FPW consecutive rows of one column x times P pulses.
  // Fields are x-major (field[x][y]), so the FPW facets of a wave are contiguous.
  const int lane = tid & 63;
  const int pi = lane & (P - 1);
  const int fsub = lane / P;
  constexpr int FPW = 64 / P;
  constexpr int FPB = (kThreads / 64) * FPW;
  const int wave = tid >> 6;
  const int p = pstart + pi;
  const bool active = p < pend;
  const int pc = active ? p : pstart;
  const int m = pc / d.K;
  const int rk = pc - m * d.K;
  const double w = (double)rk / (double)d.K;
  const size_t plane = (size_t)d.nx * d.ny;
  const size_t snap_off = (size_t)(m - d.snap0) * plane;
  const int rslot = (d.flags & OSAR_ECHO_BRAGG) ? d.rcs_slot[pc] - d.rcs0 : 0;
  const int wb = wlo[pi];
  const int wend = min(wb + W, d.nbins);
  const double two_k0 = 2 * d.k0;
  const bool bragg = (d.flags & OSAR_ECHO_BRAGG) != 0;
  const bool xti = (d.flags & OSAR_ECHO_XTI) != 0;
  // pol code of the specular field: 0 vv, 1 hh, 2 DP (e[0] hh, e[1] vv)
  const int spol = npol == 2 ? 2 : (d.pol == 1 ? 1 : 0);
  double spec_acc = 0.;

  for (int rg = 0; rg < rows; rg += FPB) {
    const int r = rg + wave * FPW + fsub;
    const bool row_ok = active && r < rows;
    const int rr = row_ok ? r : 0;
    const int yrow = y0 + rr;
    const double sin_az = geo[3 * (pi * R + rr)];
    const double mig = geo[3 * (pi * R + rr) + 1];
    const double phi = geo[3 * (pi * R + rr) + 2];
    float2 fch[OSAR_MAX_CH];
#pragma unroll
    for (int ch = 0; ch < OSAR_MAX_CH; ++ch)
      fch[ch] = ch < nch ? F[(pi * nch + ch) * R + rr] : make_float2(0.f, 0.f);
    for (int x = 0; x < d.nx; ++x) {
      if (!row_ok) continue;
      const size_t idx = (size_t)x * d.ny + yrow;
      const size_t so = snap_off + idx;
      const double Dx = lerp_field(d.dx, so, plane, rk, w);
      const double Dy = lerp_field(d.dy, so, plane, rk, w);
      const double Dz = lerp_field(d.dz, so, plane, rk, w);
      const Geo g = facet_range(d.sr_col[x], d.cos_inc[x], d.sin_inc[x], mig, sin_az, Dx, Dy,
                                Dz, d.sr0);
      const double xb = (g.srs - d.min_sr) / d.binsize;
      if (!(fabs(xb) < 1.0e9)) continue;  // non-finite or absurd range: no valid tap
      const double fl = floor(xb);
      const int dbin = (int)fl;
      const int sb = min(max((int)rint(20.0 * (1.0 - (xb - fl))), 0), kOvs);

      float2 scene[2] = {make_float2(0.f, 0.f), make_float2(0.f, 0.f)};
      if (bragg && rslot >= 0 && rslot < d.n_rcs_slots) {
        const double surf = -two_k0 * g.srs;
        const double cap = d.cap_rate[x] * (double)(p + 1);
        const float2 e0 = osar::cis_fp64(surf - cap);
        const float2 e1 = osar::cis_fp64(surf + cap);
        const osar_c32 sm = d.scat_m[d.ind0_m[idx] + p];
        const osar_c32 sp = d.scat_p[d.ind0_p[idx] + p];
        const float2 b0 = osar::cmul(e0, make_float2(sm.re, sm.im));
        const float2 b1 = osar::cmul(e1, make_float2(sp.re, sp.im));
        for (int pol = 0; pol < npol; ++pol) {
          const size_t ao = ((size_t)(rslot * npol + pol) * 2) * plane + idx;
          const float a0 = d.amp[ao];
          const float a1 = d.amp[ao + plane];
          scene[pol] = make_float2(a0 * b0.x + a1 * b1.x, a0 * b0.y + a1 * b1.y);
        }
      }
      if (SPEC) {
        // Esn = exp(-i 2 k0 sr_surface) (4 pi)^1.5 E   (sar_raw_nompi.py:405-426)
        const double sxv = lerp_field(d.sx, so, plane, rk, w);
        const double syv = lerp_field(d.sy, so, plane, rk, w);
        const double sxxv = lerp_field(d.sxx, so, plane, rk, w);
        const double syyv = lerp_field(d.syy, so, plane, rk, w);
        const double sxyv = lerp_field(d.sxy, so, plane, rk, w);
        zd e[2];
        if (ka_spa(d.k0, spol, d.inc_col[x], phi, sxv, syv, sxxv, syyv, sxyv, d.facet_dx,
                   d.facet_dy, e)) {
          const double ph = -(2. * d.k0 * g.srs);
          double ps, pcs;
          sincos(ph, &ps, &pcs);
          const double c15 = 44.546623974653535;  // (4 pi)^1.5
          for (int pol = 0; pol < npol; ++pol) {
            const zd esn = zscl(zmul({pcs, ps}, e[pol]), c15);
            scene[pol].x += (float)esn.re;
            scene[pol].y += (float)esn.im;
            if (pol == npol - 1) spec_acc += esn.re * esn.re + esn.im * esn.im;
          }
        }
      }
      const bool inwin = (dbin - 4 >= wb) && (dbin + 4 <= wend);
      const int base = dbin - 4 - wb;
#pragma unroll
      for (int ch = 0; ch < OSAR_MAX_CH; ++ch) {
        if (ch >= nch) break;
        float2 f = fch[ch];
        if (xti && d.b_xti[ch] != 0.0) {
          const double tot = d.dinc_col[x] + g.dinc;
          f = osar::cmul(f, osar::cis_fp64(-(d.k0 * d.b_xti[ch]) * tot));
        }
        for (int pol = 0; pol < npol; ++pol) {
          const float2 v = osar::cmul(scene[pol], f);
          float* lre = acc + (size_t)(((pol * nch + ch) * P + pi) * 2) * W;
          float* lim = lre + W;
          if (inwin) {
#pragma unroll
            for (int j = 0; j < kTaps; ++j) {
              const int jj = (j + fsub) & (kTaps - 1);
              const float t = tab[sb + kOvs * jj];
              atomicAdd(&lre[base + jj], t * v.x);
              atomicAdd(&lim[base + jj], t * v.y);
            }
          } else {
            for (int j = 0; j < kTaps; ++j) {
              const int b = dbin - 4 + j;
              if (b < 0 || b >= d.nbins) continue;
              const float t = tab[sb + kOvs * j];
              if (b >= wb && b < wend) {
                atomicAdd(&lre[b - wb], t * v.x);
                atomicAdd(&lim[b - wb], t * v.y);
              } else {
                osar_c32* o = d.raw + ((size_t)(pol * nch + ch) * d.n_pulses_total + p) *
                                          d.nbins + b;
                unsafeAtomicAdd(&o->re, t * v.x);
                unsafeAtomicAdd(&o->im, t * v.y);
              }
            }
          }
        }
      }
    }
  }
  if (SPEC) {
    // sum |Esn|^2 over the lanes of this pulse (lanes pi, pi + P, ...), one atomic per wave
#pragma unroll
    for (int off = P; off < 64; off <<= 1) spec_acc += __shfl_xor(spec_acc, off);
    if (fsub == 0 && active && d.spec_energy) atomicAdd(&d.spec_energy[p], spec_acc);
  }
  __syncthreads();

  // Add the LDS windows to the raw cube.
  for (int i = tid; i < nlines * W; i += kThreads) {
    const int line = i / W, b = i - line * W;
    const int lp = line % P;
    const int pc2 = line / P;  // pol * nch + ch
    const int pp = pstart + lp;
    const int gb = wlo[lp] + b;
    if (pp >= pend || gb >= d.nbins) continue;
    const float re = acc[(size_t)line * 2 * W + b];
    const float im = acc[(size_t)line * 2 * W + W + b];
    if (re == 0.f && im == 0.f) continue;
    osar_c32* o = d.raw + ((size_t)pc2 * d.n_pulses_total + pp) * d.nbins + gb;
    unsafeAtomicAdd(&o->re, re);
    unsafeAtomicAdd(&o->im, im);
  }
}

template <int P, bool SPEC>
int launch_accumulate(osar_ctx* ctx, const osar_echo_desc* d, int W, int R, hipStream_t s) {
  const int nlines = d->n_pol * d->n_ch * P;
  const size_t lds = sizeof(double) * 3 * P * R + sizeof(float2) * P * d->n_ch * R +
                     sizeof(float) * (size_t)nlines * 2 * W + sizeof(float) * 164 +
                     sizeof(int) * P;
  OSAR_REQUIRE(lds <= ctx->max_lds, OSAR_EUNSUPPORTED,
               "echo window needs %zu B of LDS (limit %zu); lower window_bins or "
               "pulses_per_block", lds, ctx->max_lds);
  OSAR_HIP(hipFuncSetAttribute((const void*)echo_accumulate_kernel<P, SPEC>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  dim3 grid((d->npulses + P - 1) / P, (d->ny + R - 1) / R);
  hipLaunchKernelGGL((echo_accumulate_kernel<P, SPEC>), grid, dim3(kThreads), lds, s, *d, W,
                     R);
  OSAR_LAUNCHED("echo_accumulate_kernel");
  return OSAR_OK;
}

template <int P>
int launch_accumulate_p(osar_ctx* ctx, const osar_echo_desc* d, int W, int R, hipStream_t s) {
  if (d->flags & OSAR_ECHO_SPECULAR) return launch_accumulate<P, true>(ctx, d, W, R, s);
  return launch_accumulate<P, false>(ctx, d, W, R, s);
}

// chan_profile_numba: one thread per facet, taps added with global atomics.
__global__ void echo_deposit_kernel(const double* srg, const osar_c32* scene, int64_t n,
                                    double binsize, double min_sr, const float* tab, int nsamp,
                                    int ovs, osar_c32* out, int nbins) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double xb = (srg[i] - min_sr) / binsize;
  if (!(fabs(xb) < 1.0e9)) return;
  const double fl = floor(xb);
  const int dbin = (int)fl;
  const int sb = min(max((int)rint((double)ovs * (1.0 - (xb - fl))), 0), ovs);
  const osar_c32 v = scene[i];
  const int b0 = dbin - nsamp / 2;
  for (int j = 0; j < nsamp; ++j) {
    const int b = b0 + j;
    if (b < 0 || b >= nbins) continue;
    const float t = tab[sb + j * ovs];
    unsafeAtomicAdd(&out[b].re, t * v.re);
    unsafeAtomicAdd(&out[b].im, t * v.im);
  }
}

}  // namespace

extern "C" int osar_echo_accumulate(osar_ctx* ctx, const osar_echo_desc* d, void* stream) {
  OSAR_PTR(ctx);
  OSAR_PTR(d);
  OSAR_REQUIRE(d->nx > 0 && d->ny > 0 && d->nbins > 0, OSAR_EINVAL, "empty scene");
  OSAR_REQUIRE(d->n_ch >= 1 && d->n_ch <= OSAR_MAX_CH, OSAR_EINVAL, "n_ch must be 1..%d",
               OSAR_MAX_CH);
  OSAR_REQUIRE(d->n_pol == 1 || d->n_pol == 2, OSAR_EINVAL, "n_pol must be 1 or 2");
  OSAR_REQUIRE(d->K >= 1, OSAR_EINVAL, "K must be >= 1");
  OSAR_REQUIRE(d->pulse0 >= 0 && d->npulses >= 0 &&
                   d->pulse0 + d->npulses <= d->n_pulses_total,
               OSAR_EINVAL, "pulse window [%d, %d) outside [0, %d)", d->pulse0,
               d->pulse0 + d->npulses, d->n_pulses_total);
  if (d->npulses == 0) return OSAR_OK;
  // Every snapshot the window touches must be resident.
  const int m_first = d->pulse0 / d->K;
  const int p_last = d->pulse0 + d->npulses - 1;
  const int m_last = p_last / d->K + ((p_last % d->K) ? 1 : 0);
  OSAR_REQUIRE(m_first >= d->snap0 && m_last < d->snap0 + d->n_snap_slots, OSAR_EINVAL,
               "snapshots %d..%d not in slots [%d, %d)", m_first, m_last, d->snap0,
               d->snap0 + d->n_snap_slots);
  OSAR_PTR(d->sr_col);
  OSAR_PTR(d->sin_inc);
  OSAR_PTR(d->cos_inc);
  OSAR_PTR(d->y_row);
  OSAR_PTR(d->dx);
  OSAR_PTR(d->dy);
  OSAR_PTR(d->dz);
  OSAR_PTR(d->sinc_tab);
  OSAR_PTR(d->raw);
  if (d->flags & OSAR_ECHO_BRAGG) {
    OSAR_PTR(d->amp);
    OSAR_PTR(d->rcs_slot);
    OSAR_PTR(d->scat_m);
    OSAR_PTR(d->scat_p);
    OSAR_PTR(d->ind0_m);
    OSAR_PTR(d->ind0_p);
    OSAR_PTR(d->cap_rate);
  }
  if (d->flags & OSAR_ECHO_XTI) OSAR_PTR(d->dinc_col);
  if (d->flags & OSAR_ECHO_SPECULAR) {
    OSAR_PTR(d->sx);
    OSAR_PTR(d->sy);
    OSAR_PTR(d->sxx);
    OSAR_PTR(d->syy);
    OSAR_PTR(d->sxy);
    OSAR_PTR(d->inc_col);
  }
  OSAR_REQUIRE(d->pol >= 0 && d->pol <= 2 && (d->pol == 2) == (d->n_pol == 2), OSAR_EINVAL,
               "pol %d inconsistent with n_pol %d", d->pol, d->n_pol);
  int P = d->pulses_per_block ? d->pulses_per_block : 8;
  const int R = d->rows_per_block ? d->rows_per_block : 64;
  int W = d->window_bins ? d->window_bins : 512;
  OSAR_REQUIRE(R >= 1 && R <= 1024 && W >= 8 && W <= 65536, OSAR_EINVAL,
               "bad launch shape R=%d W=%d", R, W);
  hipStream_t s = (hipStream_t)stream;
  OSAR_HIP(hipSetDevice(ctx->device));
  switch (P) {
    case 1: return launch_accumulate_p<1>(ctx, d, W, R, s);
    case 2: return launch_accumulate_p<2>(ctx, d, W, R, s);
    case 4: return launch_accumulate_p<4>(ctx, d, W, R, s);
    case 8: return launch_accumulate_p<8>(ctx, d, W, R, s);
    default: break;
  }
  osar_set_error("pulses_per_block must be 1, 2, 4 or 8 (got %d)", P);
  return OSAR_EINVAL;
}

extern "C" int osar_echo_deposit(osar_ctx* ctx, const double* srg, const osar_c32* scene,
                                 int64_t n, double binsize, double min_sr,
                                 const float* sinc_tab, int32_t n_sinc_samples,
                                 int32_t sinc_ovs, osar_c32* out, int32_t nbins,
                                 void* stream) {
  OSAR_PTR(ctx);
  OSAR_REQUIRE(n >= 0 && nbins > 0, OSAR_EINVAL, "bad sizes");
  OSAR_REQUIRE(n_sinc_samples > 0 && n_sinc_samples <= 64 && sinc_ovs > 0, OSAR_EINVAL,
               "bad sinc table shape");
  OSAR_REQUIRE(binsize > 0, OSAR_EINVAL, "binsize must be positive");
  if (n == 0) return OSAR_OK;
  OSAR_PTR(srg);
  OSAR_PTR(scene);
  OSAR_PTR(sinc_tab);
  OSAR_PTR(out);
  OSAR_HIP(hipSetDevice(ctx->device));
  const int64_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(echo_deposit_kernel, dim3((unsigned)blocks), dim3(256), 0,
                     (hipStream_t)stream, srg, scene, n, binsize, min_sr, sinc_tab,
                     n_sinc_samples, sinc_ovs, out, nbins);
  OSAR_LAUNCHED("echo_deposit_kernel");
  return OSAR_OK;
}
