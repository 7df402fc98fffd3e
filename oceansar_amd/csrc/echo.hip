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
// Precision: the fields between snapshots are interpolated in float32 as the
// reference's balancer does; slant range, bin position and phases are FP64 in
// the reference's operation order (FP contraction off), so bin and sub-bin
// indices match an FP64 evaluation; the phase is reduced to revolutions in
// FP64 before the FP32 sin/cos; scene values, channel factors and the
// accumulation are FP32 (LDS partial sums FP64).
#include <limits.h>
#include <string.h>

#include "common.h"

namespace {

constexpr int kThreads = 512;
constexpr int kTaps = 8;
constexpr int kOvs = 20;
constexpr int kTab = kTaps * kOvs + 1;  // 161
// LDS copy of the table, sub-bin major: row sb holds the 8 taps
// sinc_tab[sb + 20 j], j = 0..7, so a facet reads its taps with two 16-byte loads.
constexpr int kTabRows = kOvs + 1;       // 21
constexpr int kTabLds = kTabRows * kTaps;  // 168 floats

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

// Read-only element at a wave-uniform index through the constant address space,
// which the compiler serves with scalar loads into scalar registers.
__device__ __forceinline__ double uniform_load(const double* p, int i) {
  typedef const __attribute__((address_space(4))) double cdouble;
  return ((cdouble*)p)[__builtin_amdgcn_readfirstlane(i)];
}

// Scatterer series offsets of one facet. They are loaded one column ahead of
// the facet's other inputs, so that the gather of the scatterer values (which
// depends on them) never waits for a load issued in the same column.
struct FacetIdx {
  int m, p;
};

// Element at a 32-bit byte offset from a wave-uniform base: the load takes the
// base in scalar registers and only the offset in a vector register, so loads
// that share an offset share its arithmetic.
template <typename T>
__device__ __forceinline__ T load_at(const T* base, uint32_t byte_off) {
  return *(const T*)((const char*)base + byte_off);
}

__device__ __forceinline__ void load_index(const osar_echo_desc& d, uint32_t idx4, bool bragg,
                                           FacetIdx& fi) {
  if (bragg) {
    fi.m = load_at(d.ind0_m, idx4);
    fi.p = load_at(d.ind0_p, idx4);
  }
}

// Wave-uniform bases of the per-facet inputs of a workgroup's pulse group: the
// snapshot fields at the group's first snapshot (a lane adds its own snapshot
// offset), and per lane the amplitude plane of its RCS refresh.
template <int NPOL>
struct FacetBases {
  const float* dx;
  const float* dy;
  const float* dz;
  const float* s[5];
  const float* amp[NPOL][2];   // planes of the group's first refresh slot
  uint32_t aoff;               // per lane: byte offset of the lane's slot from them
};

// Per-lane inputs of one facet (one column x, the lane's row and pulse),
// loaded one column ahead, so no load is waited for in the column that issues
// it. The column's geometry (the same for every lane) is read in the column's
// own iteration through the scalar cache: carrying it through the pipeline
// would cost vector register moves every column.
template <int NPOL, bool SPEC>
struct FacetIn {
  float dx0, dy0, dz0, dx1, dy1, dz1;
  float amp[NPOL][2];
  osar_c32 sm, sp;
  float s0[5], s1[5];  // specular slope / curvature fields at m and m+1
};

// idx4: byte offset of the facet in a field plane; f0, f1: the lane's byte
// offsets of its snapshots m and m + 1 (f1 = f0 at a snapshot pulse, where the
// second value is not used) from the group's first snapshot. The displacement
// fields are loaded as soon as the column before has used them; the Bragg and
// specular inputs once the column before has formed its scene value.
template <int NPOL, bool SPEC>
__device__ __forceinline__ void load_fields(const FacetBases<NPOL>& B, uint32_t idx4,
                                            uint32_t f0, uint32_t f1, FacetIn<NPOL, SPEC>& f) {
  const uint32_t o0 = idx4 + f0, o1 = idx4 + f1;
  f.dx0 = load_at(B.dx, o0);
  f.dy0 = load_at(B.dy, o0);
  f.dz0 = load_at(B.dz, o0);
  f.dx1 = load_at(B.dx, o1);
  f.dy1 = load_at(B.dy, o1);
  f.dz1 = load_at(B.dz, o1);
}

// bragg: the call's Bragg flag (uniform). Lanes whose pulse has no RCS refresh
// in the call's slots read slot 0 and do not use the values.
template <int NPOL, bool SPEC>
__device__ __forceinline__ void load_scat(const osar_echo_desc& d, const FacetBases<NPOL>& B,
                                          uint32_t idx4, uint32_t f0, uint32_t f1, int p,
                                          bool bragg, const FacetIdx& fi,
                                          FacetIn<NPOL, SPEC>& f) {
  if (bragg) {
    const uint32_t oa = idx4 + B.aoff;
#pragma unroll
    for (int pol = 0; pol < NPOL; ++pol) {
      f.amp[pol][0] = load_at(B.amp[pol][0], oa);
      f.amp[pol][1] = load_at(B.amp[pol][1], oa);
    }
    // byte offsets below 2^32: the host bounds the series to 2^29 entries
    f.sm = load_at(d.scat_m, (uint32_t)(fi.m + p) << 3);
    f.sp = load_at(d.scat_p, (uint32_t)(fi.p + p) << 3);
  }
  if (SPEC) {
    const uint32_t o0 = idx4 + f0, o1 = idx4 + f1;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      f.s0[k] = load_at(B.s[k], o0);
      f.s1[k] = load_at(B.s[k], o1);
    }
  }
}

// Field value between two snapshots as the reference's snapshot balancer
// interpolates (oceansar/surfaces/balancer.py:274-282): float32 products and
// sum, w_l * a + w_h * b (no fused multiply-add), widened to FP64 for the range.
// At a snapshot pulse (rk = 0: w_l = 1, w_h = 0) b is the same snapshot's
// value (load_facet), and 1 * a + 0 * a is exactly a.
__device__ __forceinline__ double lerp2(float a, float b, int rk, float wl, float wh) {
#pragma clang fp contract(off)
  (void)rk;
  const float x = wl * a;
  const float y = wh * b;
  return (double)(x + y);
}

// Range-line layout in LDS: line (pol, ch, pulse) at offset line * LS doubles,
// re in [0, W), im in [W, 2W); LS = 2W + pad keeps lines of different pulses
// on different LDS banks. Lines are FP64: LDS FP64 add atomics run ~9x the
// rate of FP32 add atomics on gfx950 and keep the partial sums exact enough.
struct Lines {
  double* acc;
  int LS, W, wb, wend, nbins, npt;
  osar_c32* raw;
};

// Add v to bin b of line (pol, ch) of this lane's pulse: LDS window when b is
// inside it, global memory when only inside the range line, dropped outside
// [0, nbins) (profile_integrator's edge rule, range_profile.py:40-45).
template <int P, int NCH>
__device__ __forceinline__ void deposit(const Lines& L, int pol, int ch, int nch, int pi, int p,
                                        int b, float2 v) {
  if (b >= L.wb && b < L.wend) {
    double* lre = L.acc + (size_t)((pol * NCH + ch) * P + pi) * L.LS;
    atomicAdd(&lre[b - L.wb], (double)v.x);
    atomicAdd(&lre[L.W + b - L.wb], (double)v.y);
  } else if (b >= 0 && b < L.nbins) {
    osar_c32* o = L.raw + ((size_t)(pol * nch + ch) * L.npt + p) * L.nbins + b;
    unsafeAtomicAdd(&o->re, v.x);
    unsafeAtomicAdd(&o->im, v.y);
  }
}

// Flush one channel-independent window slot: multiply by each channel's
// per-(row, pulse) factor and deposit. The window test is made once per slot;
// inside the window every (pol, ch) line takes its two LDS adds without further
// branches (lines of channels >= nch exist in LDS and receive zeros, as their
// factors are zero, and are never written back).
template <int P, int NCH, int NPOL>
__device__ __forceinline__ void flush_slot(const Lines& L, const float2 (&s)[NPOL],
                                           const float2 (&fch)[NCH], int nch, int pi, int p,
                                           int b) {
  if (b >= L.wb && b < L.wend) {
    double* base = L.acc + (size_t)pi * L.LS + (b - L.wb);
#pragma unroll
    for (int pol = 0; pol < NPOL; ++pol) {
#pragma unroll
      for (int ch = 0; ch < NCH; ++ch) {
        const float2 v = osar::cmul(s[pol], fch[ch]);
        double* l = base + (size_t)((pol * NCH + ch) * P) * L.LS;
        atomicAdd(l, (double)v.x);
        atomicAdd(l + L.W, (double)v.y);
      }
    }
  } else if (b >= 0 && b < L.nbins) {
#pragma unroll
    for (int pol = 0; pol < NPOL; ++pol) {
#pragma unroll
      for (int ch = 0; ch < NCH; ++ch) {
        if (ch >= nch) break;
        const float2 v = osar::cmul(s[pol], fch[ch]);
        osar_c32* o = L.raw + ((size_t)(pol * nch + ch) * L.npt + p) * L.nbins + b;
        unsafeAtomicAdd(&o->re, v.x);
        unsafeAtomicAdd(&o->im, v.y);
      }
    }
  }
}

// XTI: the cross-track phase varies per facet, so channels cannot share a
// window; this template variant deposits every tap of every channel.
// Items run so that each XCD takes a contiguous range: the pulse groups of one
// snapshot interval, which read the same surface rows, then share an L2.

// One work item of the accumulator: P pulses x R rows x a column chunk of
// descriptor d, in the LDS of the calling workgroup (every thread calls it).
template <int P, int NCH, int NPOL, bool SPEC, bool XTI>
__device__ __forceinline__ void echo_item(const osar_echo_desc& d, unsigned char* lds_raw,
                                          int item, int W, int R, int LS, int n_pgroups,
                                          int n_xchunks, int CX) {
  constexpr int NL = NPOL * NCH * P;
  double* acc = (double*)lds_raw;                  // [NL][LS]
  float* tab = (float*)(acc + (size_t)NL * LS);    // [kTabRows][kTaps]
  int* wlo = (int*)(tab + kTabLds);                // [P]

  const int tid = threadIdx.x;
  // item = (pulse group fastest, then column chunk, then row tile)
  const int pstart = d.pulse0 + (item % n_pgroups) * P;
  const int pend = d.pulse0 + d.npulses;
  const int rest = item / n_pgroups;
  const int x_lo = (rest % n_xchunks) * CX;
  const int x_hi = min(d.nx, x_lo + CX);
  const int y0 = (rest / n_xchunks) * R;
  const int rows = min(R, d.ny - y0);
  const int nch = d.n_ch;

  for (int i = tid; i < NL * LS; i += kThreads) acc[i] = 0.;
  for (int i = tid; i < kTabLds; i += kThreads) {
    const int sb = i / kTaps, j = i - sb * kTaps;
    tab[i] = d.sinc_tab[sb + kOvs * j];
  }
  if (tid < P) {
    // LDS window start for this pulse: smallest migration term over the tile's rows
    // (monotonic in |az| on either side of az = 0)
    const int p = min(pstart + tid, pend - 1);
    const double az_now = (p * d.t_step - d.t_span / 2.) * d.v_ground;
    const double a0 = d.y_row[y0] - az_now, a1 = d.y_row[y0 + rows - 1] - az_now;
    double mmin = fmin((a0 / 2) * (a0 / d.sr0), (a1 / 2) * (a1 / d.sr0));
    if (a0 <= 0. && a1 >= 0.) mmin = 0.;
    const double sr_lo = n_xchunks > 1 ? d.sr_col[x_lo] : d.sr_col_min;
    const double lo = (sr_lo + mmin - d.resid_max - d.min_sr) / d.binsize;
    wlo[tid] = max((int)floor(lo) - 5, 0);
  }
  __syncthreads();

  // Lane = (facet row, pulse); the lane walks the columns x of its row, along
  // which the range bin grows monotonically (by ~0.2 bin per column), so the
  // lane's 8 taps land in a window of 8 consecutive bins held in registers.
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
  const float wh = (float)((double)rk / (double)d.K);   // w_h, w_l rounded to float32
  const float wl = (float)(1.0 - (double)wh);
  const size_t plane = (size_t)d.nx * d.ny;
  // The plain variants (no specular, no cross-track phase) run only with the
  // Bragg flag set (their launches skip calls without it), so their column
  // loop has no branch around the Bragg loads.
  const bool bragg = (SPEC || XTI) ? (d.flags & OSAR_ECHO_BRAGG) != 0 : true;
  int rslot = bragg ? d.rcs_slot[pc] - d.rcs0 : 0;
  const bool bragg_ok = bragg && rslot >= 0 && rslot < d.n_rcs_slots;
  // Amplitude planes by 32-bit byte offsets from the refresh slot in force at
  // the group's first pulse (wave-uniform): the slot index does not decrease
  // with the pulse and grows by at most one per pulse, so a lane is 0 .. P - 1
  // slots above it (clamped, so that any slot table stays in bounds; the host
  // bounds P slots of planes to 32 bits). Lanes without a refresh in the
  // call's slots read the group's slot and do not use the values.
  const int rs0 = bragg ? min(max(d.rcs_slot[pstart] - d.rcs0, 0), max(d.n_rcs_slots - 1, 0)) : 0;
  if (!bragg_ok) rslot = rs0;
  const uint32_t aoff =
      (uint32_t)(min(max(rslot - rs0, 0), P - 1) * NPOL * 2) * (uint32_t)(plane * 4);
  // Snapshot fields from the group's first snapshot m_lo = pstart / K (every
  // lane's m is at most P / K + 1 above it; the host bounds the byte offsets to
  // 32 bits); at a snapshot pulse (rk = 0) the second value repeats the first.
  FacetBases<NPOL> B;
  {
    const size_t fb = (size_t)(pstart / d.K - d.snap0) * plane;
    B.dx = d.dx + fb;
    B.dy = d.dy + fb;
    B.dz = d.dz + fb;
    if (SPEC) {
      B.s[0] = d.sx + fb;
      B.s[1] = d.sy + fb;
      B.s[2] = d.sxx + fb;
      B.s[3] = d.syy + fb;
      B.s[4] = d.sxy + fb;
    }
#pragma unroll
    for (int pol = 0; pol < NPOL; ++pol) {
      const float* a0 = d.amp + ((size_t)(rs0 * NPOL + pol) * 2) * plane;
      B.amp[pol][0] = a0;
      B.amp[pol][1] = a0 + plane;
    }
    B.aoff = aoff;
  }
  const uint32_t f0 = (uint32_t)((m - pstart / d.K) * plane * 4);
  const uint32_t f1 = f0 + (rk != 0 ? (uint32_t)(plane * 4) : 0u);
  Lines L;
  L.acc = acc;
  L.LS = LS;
  L.W = W;
  L.wb = wlo[pi];
  L.wend = min(L.wb + W, d.nbins);
  L.nbins = d.nbins;
  L.npt = d.n_pulses_total;
  L.raw = d.raw;
  // Bragg phases in revolutions: -2 k0 sr / (2 pi) and cap_rate (p + 1) / (2 pi)
  const double surf_rev = -(2 * d.k0) * osar::kInvTwoPi;
  const double cap_rev = (double)(p + 1) * osar::kInvTwoPi;
  const double inv_binsize = 1.0 / d.binsize;
  const int spol = NPOL == 2 ? 2 : (d.pol == 1 ? 1 : 0);
  double spec_acc = 0.;

  for (int rg = 0; rg < rows; rg += FPB) {
    const int r = rg + wave * FPW + fsub;
    if (!(active && r < rows)) continue;
    const int yrow = y0 + r;
    // per-row quantities of this lane's pulse (sar_raw_nompi.py:352-363, 485-493)
    const double az_now = (p * d.t_step - d.t_span / 2.) * d.v_ground;
    const double az = d.y_row[yrow] - az_now;
    const double sin_az = az / d.sr0;
    const double mig = (az / 2) * sin_az;
    const double phi = SPEC ? asin(az / d.gr0) : 0.;
    const double beam = np_sinc(d.ltx_over_l * sin_az) * np_sinc(d.lrx_over_l * sin_az);
    float2 fch[NCH];
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) {
      double s = 0., c = 0.;
      if (ch < nch) sincos(-(d.k0 * d.b_ati[ch]) * sin_az, &s, &c);
      fch[ch] = make_float2((float)(beam * c), (float)(beam * s));
    }
    // register window: win[pol][j] accumulates bin wbase + j
    float2 win[NPOL][kTaps];
#pragma unroll
    for (int pol = 0; pol < NPOL; ++pol)
#pragma unroll
      for (int j = 0; j < kTaps; ++j) win[pol][j] = make_float2(0.f, 0.f);
    int wbase = INT_MIN;

    // software pipeline: each input of column x + 1 is loaded into the
    // register that held column x's value as soon as column x has used it
    // (the last column reloads itself), the scatterer offsets two columns
    // ahead, so the loop carries no register copies
    FacetIn<NPOL, SPEC> cur;
    FacetIdx ia = {0, 0};
    const uint32_t col4 = (uint32_t)d.ny * 4u;   // bytes between columns of a field plane
    const uint32_t idx4_lo = ((uint32_t)x_lo * d.ny + yrow) * 4u;
    load_index(d, idx4_lo, bragg, ia);
    load_fields<NPOL, SPEC>(B, idx4_lo, f0, f1, cur);
    load_scat<NPOL, SPEC>(d, B, idx4_lo, f0, f1, p, bragg, ia, cur);
    load_index(d, idx4_lo + (uint32_t)(min(x_lo + 1, x_hi - 1) - x_lo) * col4, bragg, ia);
    // s_waitcnt vmcnt(0): the row's first loads complete here, so that inside
    // the column loop (which issues its loads in a fixed order, unconditionally
    // in the plain variants) each wait covers only the loads of the column
    // before, not the ones just issued for the next columns
    __builtin_amdgcn_s_waitcnt(0x0F70);
    for (int x = x_lo; x < x_hi; ++x) {
      const uint32_t idx4n = idx4_lo + (uint32_t)(min(x + 1, x_hi - 1) - x_lo) * col4;
      const uint32_t idx4nn = idx4_lo + (uint32_t)(min(x + 2, x_hi - 1) - x_lo) * col4;
      const double Dx = lerp2(cur.dx0, cur.dx1, rk, wl, wh);
      const double Dy = lerp2(cur.dy0, cur.dy1, rk, wl, wh);
      const double Dz = lerp2(cur.dz0, cur.dz1, rk, wl, wh);
      load_fields<NPOL, SPEC>(B, idx4n, f0, f1, cur);
      const double c_sr = uniform_load(d.sr_col, x), c_ci = uniform_load(d.cos_inc, x),
                   c_si = uniform_load(d.sin_inc, x);
      const Geo g = facet_range(c_sr, c_ci, c_si, mig, sin_az, Dx, Dy,
                                Dz, d.sr0);
      const double xb = osar::div_by(g.srs - d.min_sr, d.binsize, inv_binsize);
      const bool finite = fabs(xb) < 1.0e9;   // non-finite or absurd range: no valid tap
      const double fl = finite ? floor(xb) : 0.;
      const int dbin = (int)fl;
      const int sb = min(max((int)rint(20.0 * (1.0 - (xb - fl))), 0), kOvs);

      float2 scene[NPOL];
#pragma unroll
      for (int pol = 0; pol < NPOL; ++pol) scene[pol] = make_float2(0.f, 0.f);
      if (bragg_ok) {
        // exp(i (-2 k0 sr_surface -+ cap_phase)) (sar_raw_nompi.py:470-480),
        // the FP64 phases formed in revolutions
        const double surf = g.srs * surf_rev;
        const double cap = uniform_load(d.cap_rate, x) * cap_rev;
        const float2 e0 = osar::cis_rev(surf - cap);
        const float2 e1 = osar::cis_rev(surf + cap);
        const float2 b0 = osar::cmul(e0, make_float2(cur.sm.re, cur.sm.im));
        const float2 b1 = osar::cmul(e1, make_float2(cur.sp.re, cur.sp.im));
#pragma unroll
        for (int pol = 0; pol < NPOL; ++pol) {
          const float a0 = cur.amp[pol][0], a1 = cur.amp[pol][1];
          scene[pol] = make_float2(a0 * b0.x + a1 * b1.x, a0 * b0.y + a1 * b1.y);
        }
      }
      if (SPEC) {
        // Esn = exp(-i 2 k0 sr_surface) (4 pi)^1.5 E   (sar_raw_nompi.py:405-426)
        zd e[2];
        if (finite && ka_spa(d.k0, spol, uniform_load(d.inc_col, x), phi, lerp2(cur.s0[0], cur.s1[0], rk, wl, wh),
                             lerp2(cur.s0[1], cur.s1[1], rk, wl, wh),
                             lerp2(cur.s0[2], cur.s1[2], rk, wl, wh),
                             lerp2(cur.s0[3], cur.s1[3], rk, wl, wh),
                             lerp2(cur.s0[4], cur.s1[4], rk, wl, wh), d.facet_dx, d.facet_dy, e)) {
          double ps, pcs;
          sincos(-(2. * d.k0 * g.srs), &ps, &pcs);
          const double c15 = 44.546623974653535;  // (4 pi)^1.5
#pragma unroll
          for (int pol = 0; pol < NPOL; ++pol) {
            const zd esn = zscl(zmul({pcs, ps}, e[pol]), c15);
            scene[pol].x += (float)esn.re;
            scene[pol].y += (float)esn.im;
            if (pol == NPOL - 1) spec_acc += esn.re * esn.re + esn.im * esn.im;
          }
        }
      }
      load_scat<NPOL, SPEC>(d, B, idx4n, f0, f1, p, bragg, ia, cur);
      load_index(d, idx4nn, bragg, ia);
      if (finite) {
        if (!XTI) {
          // slide the window so that it starts at bin dbin - 4
          const int want = dbin - 4;
          if (want != wbase) {
            const int shift = (wbase == INT_MIN) ? -1 : want - wbase;
            if (shift < 0 || shift > kTaps) {
              if (wbase != INT_MIN) {
#pragma unroll
                for (int j = 0; j < kTaps; ++j) {
                  float2 s[NPOL];
#pragma unroll
                  for (int pol = 0; pol < NPOL; ++pol) s[pol] = win[pol][j];
                  flush_slot<P, NCH, NPOL>(L, s, fch, nch, pi, p, wbase + j);
                }
              }
#pragma unroll
              for (int pol = 0; pol < NPOL; ++pol)
#pragma unroll
                for (int j = 0; j < kTaps; ++j) win[pol][j] = make_float2(0.f, 0.f);
            } else {
              for (int k = 0; k < shift; ++k) {
                float2 s[NPOL];
#pragma unroll
                for (int pol = 0; pol < NPOL; ++pol) s[pol] = win[pol][0];
                flush_slot<P, NCH, NPOL>(L, s, fch, nch, pi, p, wbase + k);
#pragma unroll
                for (int pol = 0; pol < NPOL; ++pol) {
#pragma unroll
                  for (int j = 0; j < kTaps - 1; ++j) win[pol][j] = win[pol][j + 1];
                  win[pol][kTaps - 1] = make_float2(0.f, 0.f);
                }
              }
            }
            wbase = want;
          }
          const float4 t0 = *(const float4*)(tab + sb * kTaps);
          const float4 t1 = *(const float4*)(tab + sb * kTaps + 4);
          const float t[kTaps] = {t0.x, t0.y, t0.z, t0.w, t1.x, t1.y, t1.z, t1.w};
#pragma unroll
          for (int j = 0; j < kTaps; ++j) {
#pragma unroll
            for (int pol = 0; pol < NPOL; ++pol) {
              win[pol][j].x = fmaf(t[j], scene[pol].x, win[pol][j].x);
              win[pol][j].y = fmaf(t[j], scene[pol].y, win[pol][j].y);
            }
          }
        } else {
#pragma unroll
          for (int ch = 0; ch < NCH; ++ch) {
            if (ch >= nch) break;
            float2 f = fch[ch];
            if (d.b_xti[ch] != 0.0) {
              const double tot = uniform_load(d.dinc_col, x) + g.dinc;
              f = osar::cmul(f, osar::cis_fp64(-(d.k0 * d.b_xti[ch]) * tot));
            }
#pragma unroll
            for (int pol = 0; pol < NPOL; ++pol) {
              const float2 v = osar::cmul(scene[pol], f);
              for (int j = 0; j < kTaps; ++j) {
                const float t = tab[sb * kTaps + j];
                deposit<P, NCH>(L, pol, ch, nch, pi, p, dbin - 4 + j,
                                make_float2(t * v.x, t * v.y));
              }
            }
          }
        }
      }
    }
    if (!XTI && wbase != INT_MIN) {
#pragma unroll
      for (int j = 0; j < kTaps; ++j) {
        float2 s[NPOL];
#pragma unroll
        for (int pol = 0; pol < NPOL; ++pol) s[pol] = win[pol][j];
        flush_slot<P, NCH, NPOL>(L, s, fch, nch, pi, p, wbase + j);
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
  for (int i = tid; i < NL * W; i += kThreads) {
    const int line = i / W, b = i - line * W;
    const int lp = line % P;
    const int pc2 = line / P;  // pol * NCH + ch
    const int pol = pc2 / NCH, ch = pc2 - pol * NCH;
    const int pp = pstart + lp;
    const int gb = wlo[lp] + b;
    if (ch >= nch || pp >= pend || gb >= d.nbins) continue;
    const float re = (float)acc[(size_t)line * LS + b];
    const float im = (float)acc[(size_t)line * LS + W + b];
    if (re == 0.f && im == 0.f) continue;
    osar_c32* o = d.raw + ((size_t)(pol * nch + ch) * d.n_pulses_total + pp) * d.nbins + gb;
    unsafeAtomicAdd(&o->re, re);
    unsafeAtomicAdd(&o->im, im);
  }
}

// Two workgroups per CU (4 waves per SIMD, at most 128 VGPRs) for the plain
// Bragg kernels when the LDS lines fit twice (column chunks make them fit);
// the specular and XTI variants keep their registers (2 waves per SIMD).
// Work item w = (pulse group fastest, then column chunk, then row tile) is
// remapped with xcd_item so that each XCD runs a contiguous range of items.
template <int P, int NCH, int NPOL, bool SPEC, bool XTI>
__global__ __launch_bounds__(kThreads, (SPEC || XTI) ? 2 : 4) void echo_accumulate_kernel(
    osar_echo_desc d, int W, int R, int LS, int n_pgroups, int n_xchunks, int CX, int n_items) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  echo_item<P, NCH, NPOL, SPEC, XTI>(d, lds_raw, osar::xcd_item(blockIdx.x, n_items), W, R, LS,
                                     n_pgroups, n_xchunks, CX);
}

// Persistent batched accumulation (osar_echo_accumulate_batch): the launch
// parameters of one descriptor of the batch.
struct BatchParams {
  int W, R, LS, n_pgroups, n_xchunks, CX, n_items, reserved;
};
constexpr int kBatchCounterStride = 32;   // ints between XCD counters: one 128-byte line each

// The grid holds as many workgroups as fit the device at once; workgroup b
// runs on XCD b mod 8 and takes items from that XCD's queue: the concatenation,
// over the descriptors in plan order, of the contiguous eighth of each
// descriptor's items that xcd_item gives XCD x, so items sharing surface rows
// still run behind one L2. qstart[x][k] is the queue position of descriptor
// k's part; a workgroup exits when its queue is empty.
// Each item runs the accumulator variant of its own descriptor's channel
// count; NCHF > 0: every descriptor has NCHF (rounded) channels.
template <int P, int NPOL, int NCHF>
__global__ __launch_bounds__(kThreads, 4) void echo_batch_kernel(
    const osar_echo_desc* __restrict__ descs, const BatchParams* __restrict__ prm,
    const int* __restrict__ qstart, int* counters, int n) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  __shared__ int s_item;
  const int x = blockIdx.x % osar::kXcds;
  const int* qs = qstart + (size_t)x * (n + 1);
  const int total = qs[n];
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(&counters[x * kBatchCounterStride], 1);
    __syncthreads();
    const int c = s_item;
    __syncthreads();   // every thread has read s_item before thread 0 takes the next
    if (c >= total) return;
    int lo = 0, hi = n - 1;   // the descriptor whose part holds queue position c
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (qs[mid] <= c) lo = mid;
      else hi = mid - 1;
    }
    const int k = __builtin_amdgcn_readfirstlane(lo);
    const BatchParams p = prm[k];
    const int q = p.n_items / osar::kXcds, r = p.n_items % osar::kXcds;
    const int item = x * q + min(x, r) + (c - qs[k]);
    const osar_echo_desc& d = descs[k];
    if (!(d.flags & OSAR_ECHO_BRAGG)) continue;   // no scattering: nothing to add
    if (NCHF > 0)
      echo_item<P, (NCHF > 0 ? NCHF : 1), NPOL, false, false>(d, lds_raw, item, p.W, p.R, p.LS,
                                                              p.n_pgroups, p.n_xchunks, p.CX);
    else if (d.n_ch <= 1)
      echo_item<P, 1, NPOL, false, false>(d, lds_raw, item, p.W, p.R, p.LS, p.n_pgroups,
                                          p.n_xchunks, p.CX);
    else if (d.n_ch <= 2)
      echo_item<P, 2, NPOL, false, false>(d, lds_raw, item, p.W, p.R, p.LS, p.n_pgroups,
                                          p.n_xchunks, p.CX);
    else
      echo_item<P, 4, NPOL, false, false>(d, lds_raw, item, p.W, p.R, p.LS, p.n_pgroups,
                                          p.n_xchunks, p.CX);
    __syncthreads();   // the item's LDS lines are flushed before the next item clears them
  }
}

size_t echo_lds_bytes(int P, int nch_c, int npol, int LS) {
  return sizeof(double) * (size_t)npol * nch_c * P * LS + sizeof(float) * kTabLds +
         sizeof(int) * P;
}

int line_stride(int W) {
  // 2W rounded up to a multiple of 32 doubles, plus 4: consecutive lines start
  // 8 banks apart
  return ((2 * W + 31) / 32) * 32 + 4;
}

template <int P, int NCH, int NPOL, bool SPEC, bool XTI>
int launch_accumulate(osar_ctx* ctx, const osar_echo_desc* d, int W, int R, hipStream_t s) {
  const int LS = line_stride(W);
  const size_t lds = echo_lds_bytes(P, NCH, NPOL, LS);
  OSAR_REQUIRE(lds <= ctx->max_lds, OSAR_EUNSUPPORTED,
               "echo window needs %zu B of LDS (limit %zu); lower window_bins or "
               "pulses_per_block", lds, ctx->max_lds);
  {
    const int rc = osar_reserve_lds((const void*)echo_accumulate_kernel<P, NCH, NPOL, SPEC, XTI>,
                                    lds);
    if (rc) return rc;
  }
  const int n_pgroups = (d->npulses + P - 1) / P;
  const int CX = (d->cols_per_block > 0 && d->cols_per_block < d->nx) ? d->cols_per_block : d->nx;
  const int n_xchunks = (d->nx + CX - 1) / CX;
  const int64_t n_items = (int64_t)n_pgroups * n_xchunks * ((d->ny + R - 1) / R);
  OSAR_REQUIRE(n_items < INT_MAX, OSAR_EINVAL, "too many workgroups (%lld)", (long long)n_items);
  // neither Bragg nor specular scattering: the call adds nothing to the lines
  if (!SPEC && !XTI && !(d->flags & OSAR_ECHO_BRAGG)) return OSAR_OK;
  hipLaunchKernelGGL((echo_accumulate_kernel<P, NCH, NPOL, SPEC, XTI>), dim3((unsigned)n_items),
                     dim3(kThreads), lds, s, *d, W, R, LS, n_pgroups, n_xchunks, CX,
                     (int)n_items);
  OSAR_LAUNCHED("echo_accumulate_kernel");
  return OSAR_OK;
}

template <int P, int NCH, int NPOL, bool SPEC>
int launch_x(osar_ctx* ctx, const osar_echo_desc* d, int W, int R, hipStream_t s) {
  if (d->flags & OSAR_ECHO_XTI) return launch_accumulate<P, NCH, NPOL, SPEC, true>(ctx, d, W, R, s);
  return launch_accumulate<P, NCH, NPOL, SPEC, false>(ctx, d, W, R, s);
}

template <int P, int NCH, int NPOL>
int launch_s(osar_ctx* ctx, const osar_echo_desc* d, int W, int R, hipStream_t s) {
  if (d->flags & OSAR_ECHO_SPECULAR) return launch_x<P, NCH, NPOL, true>(ctx, d, W, R, s);
  return launch_x<P, NCH, NPOL, false>(ctx, d, W, R, s);
}

template <int P, int NCH>
int launch_pol(osar_ctx* ctx, const osar_echo_desc* d, int W, int R, hipStream_t s) {
  if (d->n_pol == 2) return launch_s<P, NCH, 2>(ctx, d, W, R, s);
  return launch_s<P, NCH, 1>(ctx, d, W, R, s);
}

template <int P>
int launch_ch(osar_ctx* ctx, const osar_echo_desc* d, int W, int R, hipStream_t s) {
  if (d->n_ch <= 1) return launch_pol<P, 1>(ctx, d, W, R, s);
  if (d->n_ch <= 2) return launch_pol<P, 2>(ctx, d, W, R, s);
  if (d->n_ch <= 4) return launch_pol<P, 4>(ctx, d, W, R, s);
  return launch_pol<P, 8>(ctx, d, W, R, s);
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


// ---- All-FP64 reference accumulator (osar_echo_accumulate_fp64) ----------
// The same pulse body in FP64 from end to end: FP64 Bragg amplitudes and
// complex128 scatterer series, FP64 sin/cos of the full phases, FP64 beam,
// channel and cross-track factors, FP64 taps deposits into FP64 LDS range
// lines and a complex128 raw cube. Only the inputs the reference itself holds
// in float32 stay float32: the surface fields (interpolated between snapshots
// in float32 as the balancer does) and the range-response table. It is the
// precision reference of the fused kernel, not a fast path: a workgroup takes
// one pulse and kFp64Rows rows of facets, every facet deposits its 8 taps per
// channel and polarisation with LDS atomics, and the workgroup's lines are
// added to the cube once.
constexpr int kFp64Threads = 256;
constexpr int kFp64Rows = 16;

template <int NPOL, bool SPEC>
__global__ __launch_bounds__(kFp64Threads) void echo_fp64_kernel(osar_echo_desc d, int n_rowtiles) {
  // the reference's rounding of every product and sum (phases of ~1e4 rad
  // carry ulps of ~1e-12 rad, which a fused multiply-add would move)
#pragma clang fp contract(off)
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const int nch = d.n_ch, nb = d.nbins;
  double* lines = (double*)lds_raw;                  // [NPOL][nch][nb][2]
  const int p = d.pulse0 + (int)(blockIdx.x / n_rowtiles);
  const int y0 = (int)(blockIdx.x % n_rowtiles) * kFp64Rows;
  const int rows = min(kFp64Rows, d.ny - y0);
  for (int i = threadIdx.x; i < NPOL * nch * nb * 2; i += kFp64Threads) lines[i] = 0.;
  __syncthreads();
  const int m = p / d.K, rk = p - m * d.K;
  const float wh = (float)((double)rk / (double)d.K);
  const float wl = (float)(1.0 - (double)wh);
  const size_t plane = (size_t)d.nx * d.ny;
  const size_t s0 = (size_t)(m - d.snap0) * plane;
  const size_t s1 = rk != 0 ? s0 + plane : s0;
  const bool bragg = (d.flags & OSAR_ECHO_BRAGG) != 0;
  const bool xti = (d.flags & OSAR_ECHO_XTI) != 0;
  const int rslot = bragg ? d.rcs_slot[p] - d.rcs0 : 0;
  const bool bragg_ok = bragg && rslot >= 0 && rslot < d.n_rcs_slots;
  const double az_now = (p * d.t_step - d.t_span / 2.) * d.v_ground;
  const double binsize = d.binsize;
  const int spol = NPOL == 2 ? 2 : (d.pol == 1 ? 1 : 0);
  double spec_acc = 0.;
  const int nf = rows * d.nx;
  for (int f = threadIdx.x; f < nf; f += kFp64Threads) {
    const int x = f / rows, yrow = y0 + (f - x * rows);
    const size_t idx = (size_t)x * d.ny + yrow;
    const double az = d.y_row[yrow] - az_now;
    const double sin_az = az / d.sr0;
    const double mig = (az / 2) * sin_az;
    const double Dx = lerp2(d.dx[s0 + idx], d.dx[s1 + idx], rk, wl, wh);
    const double Dy = lerp2(d.dy[s0 + idx], d.dy[s1 + idx], rk, wl, wh);
    const double Dz = lerp2(d.dz[s0 + idx], d.dz[s1 + idx], rk, wl, wh);
    const Geo g = facet_range(d.sr_col[x], d.cos_inc[x], d.sin_inc[x], mig, sin_az, Dx, Dy, Dz,
                              d.sr0);
    const double xb = (g.srs - d.min_sr) / binsize;
    if (!(fabs(xb) < 1.0e9)) continue;
    const double fl = floor(xb);
    const int dbin = (int)fl;
    const int sb = min(max((int)rint(20.0 * (1.0 - (xb - fl))), 0), kOvs);
    zd scene[NPOL];
#pragma unroll
    for (int pol = 0; pol < NPOL; ++pol) scene[pol] = {0., 0.};
    if (bragg_ok) {
      const double surf = -(2 * d.k0) * g.srs;
      const double cap = d.cap_rate[x] * (double)(p + 1);
      double s_m, c_m, s_p, c_p;
      sincos(surf - cap, &s_m, &c_m);
      sincos(surf + cap, &s_p, &c_p);
      const int im = d.ind0_m[idx] + p, ip = d.ind0_p[idx] + p;
      const zd sm = {d.scat64_m[2 * (size_t)im], d.scat64_m[2 * (size_t)im + 1]};
      const zd sp = {d.scat64_p[2 * (size_t)ip], d.scat64_p[2 * (size_t)ip + 1]};
      const zd b0 = zmul({c_m, s_m}, sm), b1 = zmul({c_p, s_p}, sp);
#pragma unroll
      for (int pol = 0; pol < NPOL; ++pol) {
        const double* a = d.amp64 + ((size_t)(rslot * NPOL + pol) * 2) * plane + idx;
        scene[pol] = zadd(zscl(b0, a[0]), zscl(b1, a[plane]));
      }
    }
    if (SPEC) {
      const double phi = asin(az / d.gr0);
      zd e[2];
      if (ka_spa(d.k0, spol, d.inc_col[x], phi, lerp2(d.sx[s0 + idx], d.sx[s1 + idx], rk, wl, wh),
                 lerp2(d.sy[s0 + idx], d.sy[s1 + idx], rk, wl, wh),
                 lerp2(d.sxx[s0 + idx], d.sxx[s1 + idx], rk, wl, wh),
                 lerp2(d.syy[s0 + idx], d.syy[s1 + idx], rk, wl, wh),
                 lerp2(d.sxy[s0 + idx], d.sxy[s1 + idx], rk, wl, wh), d.facet_dx, d.facet_dy,
                 e)) {
        double ps, pcs;
        sincos(-(2. * d.k0 * g.srs), &ps, &pcs);
        const double c15 = 44.546623974653535;  // (4 pi)^1.5
#pragma unroll
        for (int pol = 0; pol < NPOL; ++pol) {
          const zd esn = zscl(zmul({pcs, ps}, e[pol]), c15);
          scene[pol] = zadd(scene[pol], esn);
          if (pol == NPOL - 1) spec_acc += esn.re * esn.re + esn.im * esn.im;
        }
      }
    }
    const double beam = np_sinc(d.ltx_over_l * sin_az) * np_sinc(d.lrx_over_l * sin_az);
    for (int ch = 0; ch < nch; ++ch) {
      double sn, cs;
      sincos(-(d.k0 * d.b_ati[ch]) * sin_az, &sn, &cs);
      zd fch = {beam * cs, beam * sn};
      if (xti && d.b_xti[ch] != 0.0) {
        double xs, xc;
        sincos(-(d.k0 * d.b_xti[ch]) * (d.dinc_col[x] + g.dinc), &xs, &xc);
        fch = zmul(fch, {xc, xs});
      }
#pragma unroll
      for (int pol = 0; pol < NPOL; ++pol) {
        const zd v = zmul(scene[pol], fch);
        double* l = lines + (size_t)(pol * nch + ch) * nb * 2;
        for (int j = 0; j < kTaps; ++j) {
          const int b = dbin - kTaps / 2 + j;
          if (b < 0 || b >= nb) continue;
          const double t = (double)d.sinc_tab[sb + kOvs * j];
          atomicAdd(&l[2 * b], t * v.re);
          atomicAdd(&l[2 * b + 1], t * v.im);
        }
      }
    }
  }
  if (SPEC) {
    for (int off = 32; off > 0; off >>= 1) spec_acc += __shfl_xor(spec_acc, off);
    if ((threadIdx.x & 63) == 0 && d.spec_energy) atomicAdd(&d.spec_energy[p], spec_acc);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < NPOL * nch * nb; i += kFp64Threads) {
    const int line = i / nb, b = i - line * nb;
    const double re = lines[2 * (size_t)i], im = lines[2 * (size_t)i + 1];
    if (re == 0. && im == 0.) continue;
    double* o = d.raw64 + (((size_t)line * d.n_pulses_total + p) * nb + b) * 2;
    atomicAdd(&o[0], re);
    atomicAdd(&o[1], im);
  }
}

}  // namespace

namespace {

// The checks of osar_echo_accumulate; P, W, R: the launch shape with defaults.
int check_echo_desc(const osar_echo_desc* d, int* P_out, int* W_out, int* R_out) {
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
  int P = d->pulses_per_block ? d->pulses_per_block : 8;
  const int R = d->rows_per_block ? d->rows_per_block : 64;
  int W = d->window_bins ? d->window_bins : 512;
  *P_out = P;
  *W_out = W;
  *R_out = R;
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
    // scatterer values are gathered at 32-bit byte offsets
    OSAR_REQUIRE(d->scat_len > d->n_pulses_total && d->scat_len <= (1 << 29), OSAR_EINVAL,
                 "scat_len %d: the scatterer series need n_pulses_total < entries <= 2^29",
                 d->scat_len);
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
  OSAR_REQUIRE(R >= 1 && R <= 1024 && W >= 8 && W <= 65536 && d->cols_per_block >= 0,
               OSAR_EINVAL, "bad launch shape R=%d W=%d cols=%d", R, W, d->cols_per_block);
  OSAR_REQUIRE(P == 2 || P == 4 || P == 8 || P == 16, OSAR_EINVAL,
               "pulses_per_block must be 2, 4, 8 or 16 (got %d)", P);
  // field inputs are addressed by 32-bit byte offsets from a pulse group's
  // first snapshot (its lanes are at most P / K + 2 snapshots further on)
  OSAR_REQUIRE((uint64_t)d->nx * d->ny * 4 * (uint64_t)(P / d->K + 3) < (1ull << 32),
               OSAR_EUNSUPPORTED, "surface of %d x %d facets too large for 32-bit field offsets",
               d->nx, d->ny);
  // Bragg amplitudes likewise, from the refresh slot of a group's first pulse
  // (its lanes are at most P - 1 slots further on, n_pol x 2 planes per slot)
  OSAR_REQUIRE(!(d->flags & OSAR_ECHO_BRAGG) ||
                   (uint64_t)d->nx * d->ny * 4 * (uint64_t)(P * d->n_pol * 2) < (1ull << 32),
               OSAR_EUNSUPPORTED,
               "surface of %d x %d facets too large for 32-bit amplitude offsets", d->nx, d->ny);
  return OSAR_OK;
}

// Byte layout of a batch plan: descriptors, parameters, XCD queue starts, counters.
struct PlanLayout {
  size_t descs, prm, qs, cnt, total;
};
PlanLayout plan_layout(int n) {
  auto up = [](size_t v) { return (v + 255) & ~(size_t)255; };
  PlanLayout L;
  L.descs = 0;
  L.prm = up(sizeof(osar_echo_desc) * (size_t)n);
  L.qs = L.prm + up(sizeof(BatchParams) * (size_t)n);
  L.cnt = L.qs + up(sizeof(int) * (size_t)osar::kXcds * (n + 1));
  L.total = L.cnt + sizeof(int) * osar::kXcds * kBatchCounterStride;
  return L;
}

template <int P, int NPOL>
const void* batch_kernel_ch(int nchf) {
  switch (nchf) {
    case 1: return (const void*)echo_batch_kernel<P, NPOL, 1>;
    case 2: return (const void*)echo_batch_kernel<P, NPOL, 2>;
    case 4: return (const void*)echo_batch_kernel<P, NPOL, 4>;
    default: return (const void*)echo_batch_kernel<P, NPOL, 0>;
  }
}

// nchf: the rounded channel count every descriptor shares, 0 when they differ
const void* batch_kernel_for(int P, int npol, int nchf) {
  if (P == 8) return npol == 1 ? batch_kernel_ch<8, 1>(nchf) : batch_kernel_ch<8, 2>(nchf);
  return npol == 1 ? batch_kernel_ch<4, 1>(nchf) : batch_kernel_ch<4, 2>(nchf);
}

}  // namespace

extern "C" int osar_echo_accumulate(osar_ctx* ctx, const osar_echo_desc* d, void* stream) {
  OSAR_PTR(ctx);
  int P = 0, W = 0, R = 0;
  const int rc = check_echo_desc(d, &P, &W, &R);
  if (rc) return rc;
  if (d->npulses == 0) return OSAR_OK;
  hipStream_t s = (hipStream_t)stream;
  OSAR_HIP(hipSetDevice(ctx->device));
  switch (P) {
    case 2: return launch_ch<2>(ctx, d, W, R, s);
    case 4: return launch_ch<4>(ctx, d, W, R, s);
    case 8: return launch_ch<8>(ctx, d, W, R, s);
    default: return launch_ch<16>(ctx, d, W, R, s);
  }
}

extern "C" int64_t osar_echo_batch_plan_bytes(int32_t n) {
  if (n < 1 || n > OSAR_ECHO_BATCH_MAX) return -1;
  return (int64_t)plan_layout(n).total;
}

extern "C" int osar_echo_accumulate_batch(osar_ctx* ctx, const osar_echo_batch_desc* b,
                                          void* stream) {
  OSAR_PTR(ctx);
  OSAR_PTR(b);
  OSAR_REQUIRE(b->n >= 1 && b->n <= OSAR_ECHO_BATCH_MAX, OSAR_EINVAL,
               "batch of %d descriptors: 1 .. %d", b->n, OSAR_ECHO_BATCH_MAX);
  OSAR_PTR(b->descs);
  OSAR_PTR(b->plan_host);
  OSAR_PTR(b->plan_dev);
  OSAR_REQUIRE(b->workgroups >= 0, OSAR_EINVAL, "negative workgroup count");
  const int n = b->n;
  const PlanLayout L = plan_layout(n);
  unsigned char* h = (unsigned char*)b->plan_host;
  BatchParams* prm = (BatchParams*)(h + L.prm);
  int* qs = (int*)(h + L.qs);
  int P0 = 0, npol = 0, pol = -1;
  for (int k = 0; k < n; ++k) {
    const osar_echo_desc* d = &b->descs[k];
    int P = 0, W = 0, R = 0;
    const int rc = check_echo_desc(d, &P, &W, &R);
    if (rc) return rc;
    OSAR_REQUIRE(!(d->flags & (OSAR_ECHO_SPECULAR | OSAR_ECHO_XTI)), OSAR_EUNSUPPORTED,
                 "descriptor %d: specular and cross-track batches are not supported", k);
    OSAR_REQUIRE(P == 4 || P == 8, OSAR_EUNSUPPORTED,
                 "descriptor %d: batches take pulses_per_block 4 or 8 (got %d)", k, P);
    OSAR_REQUIRE(d->n_ch <= 4, OSAR_EUNSUPPORTED, "descriptor %d: batches take at most 4 "
                 "channels (got %d)", k, d->n_ch);
    if (k == 0) {
      P0 = P;
      npol = d->n_pol;
      pol = d->pol;
    }
    OSAR_REQUIRE(P == P0 && d->n_pol == npol && d->pol == pol, OSAR_EINVAL,
                 "descriptor %d: every descriptor of a batch needs the same pulses_per_block "
                 "and polarisation", k);
    BatchParams& p = prm[k];
    p.W = W;
    p.R = R;
    p.LS = line_stride(W);
    p.n_pgroups = (d->npulses + P - 1) / P;
    p.CX = (d->cols_per_block > 0 && d->cols_per_block < d->nx) ? d->cols_per_block : d->nx;
    p.n_xchunks = (d->nx + p.CX - 1) / p.CX;
    const int64_t items = d->npulses == 0 ? 0
                                          : (int64_t)p.n_pgroups * p.n_xchunks *
                                                ((d->ny + R - 1) / R);
    OSAR_REQUIRE(items < INT_MAX / 2, OSAR_EINVAL, "descriptor %d: too many items", k);
    p.n_items = (int)items;
    p.reserved = 0;
  }
  size_t lds = 0;   // the largest item's lines (each item runs its own channel count)
  int nchf = -1;
  for (int k = 0; k < n; ++k) {
    const int c = b->descs[k].n_ch, cr = c <= 1 ? 1 : c <= 2 ? 2 : 4;
    const size_t l = echo_lds_bytes(P0, cr, npol, prm[k].LS);
    if (l > lds) lds = l;
    nchf = (nchf == -1 || nchf == cr) ? cr : 0;
  }
  OSAR_REQUIRE(lds <= ctx->max_lds, OSAR_EUNSUPPORTED,
               "batch window needs %zu B of LDS (limit %zu): lower window_bins", lds,
               ctx->max_lds);
  // queue starts: XCD x takes the eighth [x q + min(x, r), ...) of every descriptor
  int64_t total = 0;
  for (int x = 0; x < osar::kXcds; ++x) {
    int64_t pos = 0;
    for (int k = 0; k < n; ++k) {
      qs[x * (n + 1) + k] = (int)pos;
      const int q = prm[k].n_items / osar::kXcds, r = prm[k].n_items % osar::kXcds;
      pos += q + (x < r ? 1 : 0);
    }
    OSAR_REQUIRE(pos < INT_MAX / 2, OSAR_EINVAL, "batch too large");
    qs[x * (n + 1) + n] = (int)pos;
    total += pos;
  }
  memcpy(h + L.descs, b->descs, sizeof(osar_echo_desc) * (size_t)n);
  memset(h + L.cnt, 0, L.total - L.cnt);
  if (total == 0) return OSAR_OK;
  hipStream_t s = (hipStream_t)stream;
  OSAR_HIP(hipSetDevice(ctx->device));
  // the plan (zeroed counters included) reaches the device on the stream
  OSAR_HIP(hipMemcpyAsync(b->plan_dev, b->plan_host, L.total, hipMemcpyHostToDevice, s));
  const void* kern = batch_kernel_for(P0, npol, nchf);
  {
    const int rc = osar_reserve_lds(kern, lds);
    if (rc) return rc;
  }
  // launch bounds of 4 waves per SIMD: two 8-wave workgroups per CU when their LDS fits
  const int per_cu = lds * 2 <= ctx->max_lds ? 2 : 1;
  int G = b->workgroups > 0 ? b->workgroups : per_cu * ctx->n_cu;
  G = G < osar::kXcds ? osar::kXcds : (G / osar::kXcds) * osar::kXcds;
  unsigned char* dv = (unsigned char*)b->plan_dev;
  const osar_echo_desc* descs = (const osar_echo_desc*)(dv + L.descs);
  const BatchParams* prm_d = (const BatchParams*)(dv + L.prm);
  const int* qs_d = (const int*)(dv + L.qs);
  int* cnt_d = (int*)(dv + L.cnt);
  int nn = n;
  void* args[] = {&descs, &prm_d, &qs_d, &cnt_d, &nn};
  OSAR_HIP(hipLaunchKernel(kern, dim3(G), dim3(kThreads), args, lds, s));
  OSAR_LAUNCHED("echo_batch_kernel");
  return OSAR_OK;
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

extern "C" int osar_echo_accumulate_fp64(osar_ctx* ctx, const osar_echo_desc* d, void* stream) {
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
  OSAR_PTR(d->raw64);
  if (d->flags & OSAR_ECHO_BRAGG) {
    OSAR_PTR(d->amp64);
    OSAR_PTR(d->rcs_slot);
    OSAR_PTR(d->scat64_m);
    OSAR_PTR(d->scat64_p);
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
  const size_t lds = sizeof(double) * 2 * (size_t)d->n_pol * d->n_ch * d->nbins;
  OSAR_REQUIRE(lds <= ctx->max_lds, OSAR_EUNSUPPORTED,
               "FP64 range lines need %zu B of LDS (limit %zu)", lds, ctx->max_lds);
  OSAR_HIP(hipSetDevice(ctx->device));
  const int n_rowtiles = (d->ny + kFp64Rows - 1) / kFp64Rows;
  const int64_t n_blocks = (int64_t)d->npulses * n_rowtiles;
  OSAR_REQUIRE(n_blocks < INT_MAX, OSAR_EINVAL, "too many workgroups (%lld)",
               (long long)n_blocks);
  const bool spec = (d->flags & OSAR_ECHO_SPECULAR) != 0;
  const void* k = d->n_pol == 2 ? (spec ? (const void*)echo_fp64_kernel<2, true>
                                        : (const void*)echo_fp64_kernel<2, false>)
                                : (spec ? (const void*)echo_fp64_kernel<1, true>
                                        : (const void*)echo_fp64_kernel<1, false>);
  {
    const int rc = osar_reserve_lds(k, lds);
    if (rc) return rc;
  }
  hipStream_t s = (hipStream_t)stream;
  const dim3 grid((unsigned)n_blocks), block(kFp64Threads);
  if (d->n_pol == 2) {
    if (spec) hipLaunchKernelGGL((echo_fp64_kernel<2, true>), grid, block, lds, s, *d, n_rowtiles);
    else hipLaunchKernelGGL((echo_fp64_kernel<2, false>), grid, block, lds, s, *d, n_rowtiles);
  } else {
    if (spec) hipLaunchKernelGGL((echo_fp64_kernel<1, true>), grid, block, lds, s, *d, n_rowtiles);
    else hipLaunchKernelGGL((echo_fp64_kernel<1, false>), grid, block, lds, s, *d, n_rowtiles);
  }
  OSAR_LAUNCHED("echo_fp64_kernel");
  return OSAR_OK;
}
