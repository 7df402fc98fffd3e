// Surface synthesis: wave coefficients x dispersion phase -> inverse 2-D FFTs
// -> float32 displacement / slope / curvature fields.
//
// Replaces the OceanSurface.t setter (oceansar/surfaces/ocean.py:418-526) for
// the fields the raw-data path consumes ('D', 'Diff', 'Diff2'; choppy slopes).
// Everything is FP64 (the dispersion phase omega(k) t included); outputs are
// rounded to float32, the type OceanSurface stores them in (ocean.py:384-403).
//
// Two real outputs share one complex inverse FFT: for real fields
// A = Re ifft2(mA a) and B = Re ifft2(mB a) the packed spectrum is
// S[k] = H(mA a)[k] + i H(mB a)[k], H(g)[k] = (g[k] + conj(g[-k]))/2, and
// ifft2(S) = A + iB. A choppy slope uses the pair (numerator, denominator) and
// divides in the epilogue.
//
// Pass 1 (rows): one workgroup per row pair (j, -j): H(a) of both rows in
// registers (each thread owns columns and their mirrors), packed spectra built
// into LDS, radix-8/4 Stockham FFT along x in LDS, rows written to an FP64
// workspace W[q][x / cbw][y][x % cbw] (cbw = 8: each row writes full 128-byte
// lines; neighbouring rows run on one XCD).
// Pass 2 (columns): one workgroup per (block of cb = 2 columns, packed
// spectrum): the block is read from its workspace runs (the blocks of one run
// on one XCD), FFT along y in LDS, scaled by 1/(Nx Ny) and written as float32
// fields in x-major layout (field[x][y]) so that every store is a contiguous
// run along y. Both passes fit two workgroups per CU (LDS and registers), so
// one workgroup's HBM traffic overlaps the other's FFT.
#include <math.h>

#include <mutex>
#include <vector>

#include "common.h"
#include "fft_lds.h"

namespace {

constexpr double kG = 9.80665;
constexpr int kMaxQ = 10;
constexpr int kMaxN = 2048;

enum Mult : int {
  M_NONE = 0,
  M_ONE,
  M_IKXN,    // i kx/k       (Dx)
  M_IKYN,    // i ky/k       (Dy)
  M_IKX,     // i kx         (Diffx)
  M_IKY,     // i ky         (Diffy)
  M_NKX2,    // -kx^2        (Diffxx)
  M_NKY2,    // -ky^2        (Diffyy)
  M_NKXKY,   // -kx ky       (Diffxy)
  M_NKX2KI,  // -kx^2 / k    (choppy denominator x)
  M_NKY2KI,  // -ky^2 / k    (choppy denominator y)
  M_NIKX3KI,   // -i kx^3 / k       (choppy curvature terms, ocean.py:466-468)
  M_NIKY3KI,   // -i ky^3 / k
  M_NIKX2KYKI, // -i kx^2 ky / k
  M_WKXN,    // omega kx/k         (Vx: -Im ifft2(kx/k (-i omega) a))
  M_WKYN,    // omega ky/k         (Vy)
  M_NIW,     // -i omega           (Vz)
  M_NIW2KXN, // -i omega^2 kx/k    (Ax)
  M_NIW2KYN, // -i omega^2 ky/k    (Ay)
  M_NW2,     // -omega^2           (Az)
  M_HMTF,    // the hydrodynamic modulation transfer function M_h(k) (hMTF)
};

struct QPlan {
  int ma, mb;      // multipliers of the real and imaginary component
  int outa, outb;  // output field index (0..7) or -1
  int choppy;      // 1: single output = Re / (1 + Im) into outa
  int keep;        // 1: the column pass writes the FP64 result back in place
                   // (terms of the choppy curvatures, combined by synth_choppy_kernel)
  int ca, cb;      // the multipliers as monomial codes (mono_code)
  int f32;         // 1: the row pass stores this spectrum as complex64 (else complex128)
  int tpair;       // 1: one field at snapshots ta (component a) and ta + 1 (component b),
                   // ta even; the plan's last spectrum
  int cbw, lcbw;   // columns per contiguous run of a workspace row (128 bytes) and log2
  long long off;   // byte offset of this spectrum's plane within a snapshot's workspace
  float* pa;       // output planes of outa / outb (NULL for -1)
  float* pb;
};

// Every multiplier but M_h is a monomial s i^d kx^px ky^py |k|^-pk omega^po.
// Code bits: px (0-1), py (2-3), pk (4), d (5), s = -1 (6), no multiplier (7),
// po (8-9); M_h (10) is even in k with complex values (see hmtf_mult).
constexpr int kMonoImag = 32, kMonoNeg = 64, kMonoNone = 128, kMonoOmega = 256,
              kMonoHmtf = 1024;
constexpr int kNumFields = 15;

struct SynthArgs {
  int nx, ny, nq, logx, logy, cb, lcb, n_times;
  int tpair;            // 1: the plan's last spectrum is time-paired (QPlan::tpair)
  double depth, cur_x, cur_y, thr;
  const float* kx;
  const float* ky;
  const osar_c32* coefs;
  const double* times;
  const double2* twx;   // [128]: exp(2 pi i k / nx) for k < 64, then for k = 64 m
  const double2* twy;
  unsigned char* work;  // workspace: per snapshot, the planes of the nq spectra
  long long snap_bytes; // workspace bytes per snapshot
  double* omega;        // [ny][nx] omega(|k|), filled by omega_kernel
  double* ikt;          // [ny][nx] 1/|k| (0 off the good-k set), filled by omega_kernel
  QPlan q[kMaxQ];
  // choppy curvatures: the keep spectra A = (a_x, a_xx), B = (a_y, a_yy),
  // C = (Dxx, a_xxx), D = (Dyy, a_yyy), E = (Dxy, a_xxy) (plan indices, -1 = none)
  // and the outputs Diffx, Diffy, Diffxx, Diffyy, Diffxy (NULL = not requested)
  int ch_q[5];
  float* ch_out[5];
};

// Plan entry at a wave-uniform runtime index. A switch keeps
// the by-value kernel argument in scalar registers (indexing its arrays with a
// runtime value would copy the whole argument to scratch memory).
__device__ __forceinline__ QPlan plan_at(const SynthArgs& a, int i) {
  switch (i) {
    case 0: return a.q[0];
    case 1: return a.q[1];
    case 2: return a.q[2];
    case 3: return a.q[3];
    case 4: return a.q[4];
    case 5: return a.q[5];
    case 6: return a.q[6];
    case 7: return a.q[7];
    case 8: return a.q[8];
    default: return a.q[9];
  }
}

// Spectral multiplier m(k) = s i^d kx^px ky^py |k|^-pk for wave vector
// (kx, ky), with 1/|k| = 0 off the good-k set (so k/|k| = 0 there, as
// ocean.py:141-148 does): the row-uniform factor s ky^py (0 for no
// multiplier) times the per-column factor kx^px |k|^-pk; a sign or a unit
// factor is exact, so m(-k) is bitwise conj(m(k)) on a mirror-symmetric grid.
__device__ __forceinline__ double mono_y(int code, double ky) {
  if (code & kMonoNone) return 0.;
  const int py = (code >> 2) & 3;
  double y = py >= 1 ? ky : 1.;
  if (py >= 2) y = y * ky;
  if (py == 3) y = y * ky;
  return (code & kMonoNeg) ? -y : y;
}

__device__ __forceinline__ double mono_x(int code, double kx, double ik, double om) {
  const int px = code & 3;
  double f = px >= 1 ? kx : 1.;
  if (px >= 2) f = f * kx;
  if (px == 3) f = f * kx;
  if (code & 16) f = f * ik;
  const int po = (code >> 8) & 3;
  if (po >= 1) f = f * om;
  if (po == 2) f = f * om;
  return f;
}

// The hydrodynamic MTF of ocean.py:508-526: M_h = i k omega (omega + i mu) /
// (omega^2 + mu^2) (dlnS/dlnk - Vg/c), mu = 0.04 / 16, dlnS/dlnk = -3, Vg/c =
// 0.5, zero for k > 2 pi / 20; even in k (M_h(-k) = M_h(k)). The cut compares
// |k| as the reference's float32 table holds it (sqrt(kx^2 + ky^2), each
// operation rounded to float32) with 2 pi / 20 rounded to float32, so the same
// wave numbers are cut; the value is FP64.
__device__ __forceinline__ double2 hmtf_mult(float kx, float ky, double om) {
  constexpr double kMu = 0.04 / 16., kFac = -3. - 0.5;
  constexpr float kMax32 = (float)(2. * M_PI / 20.);
  const float k32 = __fsqrt_rn(__fadd_rn(__fmul_rn(kx, kx), __fmul_rn(ky, ky)));
  if (k32 > kMax32) return make_double2(0., 0.);
  const double k = sqrt((double)kx * kx + (double)ky * ky);
  const double den = om * om + kMu * kMu;
  // i k om (om + i mu) = -k om mu + i k om^2
  return make_double2((-k * om * kMu) / den * kFac, (k * om * om) / den * kFac);
}

// A monomial m with m(-k) = -conj(m(k)) (kx^px ky^py i^d of odd px + py + d:
// the velocities' omega k/|k| and -i omega): then H(m a) = m G(a) with
// G(a) = (a(k) - conj(a(-k))) / 2, not m H(a).
__host__ __device__ __forceinline__ bool mono_anti(int code) {
  if (code & (kMonoNone | kMonoHmtf)) return false;
  return (((code & 3) + ((code >> 2) & 3) + ((code & kMonoImag) ? 1 : 0)) & 1) != 0;
}

// the multiplier value: v real, or i v when d
__device__ __forceinline__ double2 mono_c(bool d, double v) {
  return d ? make_double2(0., v) : make_double2(v, 0.);
}

// H(m a)[k] = (m(k) a(k) + conj(m(-k) a(-k))) / 2 with m(k) = mk, m(-k) = mm
__device__ __forceinline__ double2 herm(double2 mk, double2 a, double2 mm, double2 am) {
  const double2 g = cmul2(mk, a);
  const double2 gm = cmul2(mm, am);
  return make_double2(0.5 * (g.x + gm.x), 0.5 * (g.y - gm.y));
}

constexpr int kRowThreads = 512;
constexpr int kRowWaves = 4;                     // waves per SIMD: two row workgroups per CU
constexpr int kColThreads = 512;
constexpr int kColWaves = 4;                     // two column workgroups per CU
constexpr int kColElems = 4096;                  // cb * ny: elements of one column block
constexpr int kColRadixLog = 3;                  // radix-8 passes (one butterfly per thread)
constexpr int kWsBlock = 8;                      // workspace run: 8 columns = 128 bytes
constexpr int kRowPairs = kMaxN / kRowThreads;   // column pairs a row thread owns

// |k| and 1/|k| (0 off the good-k set |k| > thr, as ocean.py:141-148 masks
// k/|k|) of the wave vector (kx, ky): one expression for every caller, so the
// same bits.
__device__ __forceinline__ double knorm(double kx, double ky) { return sqrt(kx * kx + ky * ky); }
__device__ __forceinline__ double inv_knorm(double k, double thr) { return k > thr ? 1. / k : 0.; }

// Dispersion relation omega(|k|) and 1/|k| over the wave-number grid, once per
// call: the row pass reads them instead of evaluating sqrt, tanh and a
// division per snapshot (the same expressions, so the same bits).
__global__ void omega_kernel(SynthArgs a) {
  const int64_t n = (int64_t)a.nx * a.ny;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int j = (int)(e >> a.logx), i = (int)(e & (a.nx - 1));
    const double k = knorm(a.kx[i], a.ky[j]);
    a.omega[e] = a.depth > 0. ? sqrt(kG * k * tanh(k * a.depth)) : sqrt(kG * k);
    a.ikt[e] = inv_knorm(k, a.thr);
  }
}

// c exp(-i omega t) for the complex64 coefficient c
__device__ __forceinline__ double2 phased(osar_c32 c, double om, double t) {
  double sn, cs;
  sincos(-(om * t), &sn, &cs);
  return cmul2(make_double2((double)c.re, (double)c.im), make_double2(cs, sn));
}

// The phased coefficients A0 = a(k) at (row j, column i) and A1 = a(k') at the
// mirror (jm, im); k' has the same |k| on a 2 pi fftfreq grid (kx(-i) = -kx(i)
// except at the Nyquist column, where it equals kx(i)), so omega, the phase and
// its sincos are shared then (equal omega gives equal bits either way). MIRROR:
// the tables are such a grid, so the mirror's omega is not read.
template <bool MIRROR>
__device__ __forceinline__ void phased_pair(const SynthArgs& a, int j, int i, int jm, int im,
                                            double t, double2& A0, double2& A1) {
  const double om0 = a.omega[(size_t)j * a.nx + i];
  const double om1 = MIRROR ? om0 : a.omega[(size_t)jm * a.nx + im];
  const osar_c32 c0 = a.coefs[(size_t)j * a.nx + i];
  const osar_c32 c1 = a.coefs[(size_t)jm * a.nx + im];
  double sn, cs;
  sincos(-(om0 * t), &sn, &cs);
  A0 = cmul2(make_double2((double)c0.re, (double)c0.im), make_double2(cs, sn));
  if (om1 == om0)
    A1 = cmul2(make_double2((double)c1.re, (double)c1.im), make_double2(cs, sn));
  else
    A1 = phased(c1, om1, t);
}

// Pass 1: one workgroup per row pair (j, -j) and snapshot. Thread t owns the
// columns i = t, t + kRowThreads, ... of row j together with their mirrors -i of row
// -j, so a(k) and a(-k) of every packed spectrum it builds are its own; only
// the FFT buffer lives in LDS (2 rows, 64 KB at 2048), so two workgroups share
// a CU and one's HBM stores overlap the other's FFT.
// Away from the Nyquist row and column, -k is the true mirror, every
// multiplier of the raw-data fields has m(-k) = conj(m(k)) and H(m a) = m H(a):
// the thread keeps only H(a) = (a(k) + conj(a(-k))) / 2 per column pair in
// registers. The velocities' multipliers have m(-k) = -conj(m(k)) and M_h is
// even in k; those spectra (TIMED) also need G(a) = (a(k) - conj(a(-k))) / 2,
// recomputed from the coefficients. The Nyquist column's pair (one per row
// pair) is kept in LDS by its thread.
// MIRROR (the tables are a 2 pi fftfreq grid: kx(-i) = -kx(i) off the Nyquist
// column, kx(0) = 0, the same for ky): the only elements that need the general
// form are the Nyquist column and the self-paired rows 0 and ny / 2, and a
// self-paired row keeps its own a(k) in the second half of the FFT buffer (a
// single row is transformed), so no a(k) is recomputed inside the spectrum
// loop and H(a) stays in registers. Otherwise the general form recomputes a(k)
// and a(-k) where it needs them.
// Rows go to the workspace in column-blocked layout W[q][x / cb][y][x % cb],
// so that pass 2 reads each block of cb columns as one contiguous run.
// TIMED: some spectrum's multiplier involves omega (velocities, accelerations)
// or is M_h; the rest compile without those branches and their registers.
template <bool SHORT, bool MIRROR, bool TIMED>
__global__ __launch_bounds__(kRowThreads, kRowWaves) void synth_rows_kernel(SynthArgs a) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  double2* S = (double2*)lds_raw;  // [2][nx] packed spectrum of rows j, jm
  double2* tw = S + padded(2 * a.nx);   // [128] twiddles
  double2* nyq = tw + 128;         // [2] a(k), a(-k) at the Nyquist column
  load_twiddles(tw, a.twx);
  // Row pairs in XCD-contiguous order: rows j and j + 1 fill one 128-byte
  // line of the workspace between them, so their workgroups should run
  // behind the same L2.
  // Items are row pair major, snapshot minor: the snapshots of a row pair run
  // one after another on one XCD, so its coefficients and dispersion values
  // come from HBM once and from that XCD's L2 for the other snapshots.
  // With a time-paired spectrum (a.tpair: the plan's last, one field at
  // snapshot ta as component a and at ta + 1 as component b) an item covers
  // the snapshot pair (ta, ta + 1): the spectra of ta, the paired spectrum's
  // component a built into the FFT buffer and kept there, then H(a) at ta + 1,
  // the component b added to it, its transform, and the spectra of ta + 1.
  const int npairs = a.ny / 2 + 1;
  const int nst = a.tpair ? (a.n_times + 1) >> 1 : a.n_times;   // items per row pair
  const int item = osar::xcd_item(blockIdx.x, npairs * nst);
  const int si = item % nst;
  const int j = item / nst;
  const int ta = a.tpair ? 2 * si : si;
  const int ns = a.tpair ? min(2, a.n_times - ta) : 1;
  const int qp = a.tpair ? a.nq - 1 : -1;
  const int jm = (a.ny - j) & (a.ny - 1);
  const int nrow = (j == jm) ? 1 : 2;
  const int tid = threadIdx.x;
  const int nx = a.nx;
  const float kyj = a.ky[j], kyjm = a.ky[jm];
  double2 H[kRowPairs];    // H(a) at (j, i)
  for (int s = 0; s < ns; ++s) {
    const int ti = ta + s;
    const double t = a.times[ti];
    // Opaque per snapshot: the column addresses of the H(a) loop are
    // recomputed rather than held in registers across both snapshots.
    int th = tid;
    asm volatile("" : "+v"(th));
    // Every element of the register arrays is assigned on every path (columns
    // beyond nx are clamped and their results never stored), so the arrays stay
    // in registers.
#pragma unroll
    for (int u = 0; u < kRowPairs; ++u) {
      const int i = min(th + u * kRowThreads, nx - 1);
      const int im = (nx - i) & (nx - 1);
      double2 A0, A1;
      phased_pair<MIRROR>(a, j, i, jm, im, t, A0, A1);
      H[u] = make_double2(0.5 * (A0.x + A1.x), 0.5 * (A0.y - A1.y));
      if (i == nx / 2 && th + u * kRowThreads < nx) {   // only this thread reads them back
        nyq[0] = A0;
        nyq[1] = A1;
      }
      if (MIRROR && j == jm && th + u * kRowThreads < nx) S[pad(nx + i)] = A0;
    }
    if (MIRROR) __syncthreads();   // a self-paired row reads the coefficients of the mirror columns
    for (int qq = 0; qq < a.nq; ++qq) {
      // the time-paired spectrum comes last at ta and first at ta + 1
      const int qi = (s == 1) ? (qq == 0 ? qp : qq - 1) : qq;
      QPlan q = plan_at(a, qi);
      // mode 0: the spectrum written whole; 1: the paired spectrum's component a
      // (at ta) written; 2: its component b (at ta + 1) added
      const int mode = qi != qp ? 0 : (s == 0 ? 1 : 2);
      if (mode == 1) q.cb = kMonoNone;
      if (mode == 2) q.ca = kMonoNone;
      // Store the values at k (slot pk) and at the mirror -k (slot pm). In a
      // self-paired row (one row) the mirror slot is another thread's own slot,
      // which that thread fills itself: a second store there would race with a
      // value that may differ in the last bit (the two threads round the same
      // sums in different orders).
      auto put = [&](int pk, double2 vk, int pm, double2 vm) {
        if (mode == 2) {
          S[pad(pk)] = cadd2(S[pad(pk)], vk);
          if (nrow == 2) S[pad(pm)] = cadd2(S[pad(pm)], vm);
        } else {
          S[pad(pk)] = vk;
          if (nrow == 2) S[pad(pm)] = vm;
        }
      };
      const double ya = mono_y(q.ca, kyj), yb = mono_y(q.cb, kyj);
      const bool da = (q.ca & kMonoImag) != 0, db = (q.cb & kMonoImag) != 0;
      const bool need_ik = ((q.ca | q.cb) & 16) != 0;   // a multiplier divides by |k|
      const bool need_om = TIMED && ((q.ca | q.cb) & (3 * kMonoOmega)) != 0;
      const bool hmtf = TIMED && (q.ca & kMonoHmtf) != 0;   // M_h, alone in its spectrum
      const bool anti_a = TIMED && mono_anti(q.ca), anti_b = TIMED && mono_anti(q.cb);
      // the spectrum needs G(a) besides H(a)
      const bool need_g = hmtf || anti_a || anti_b;
      // Opaque per iteration: the index and address arithmetic of the build,
      // the FFT and the write-out is recomputed per spectrum instead of being
      // held in registers across all of them.
      int logx = a.logx, tq = tid;
      asm volatile("" : "+s"(logx), "+v"(tq));
      unsigned general = 0;   // column pairs that need the general form
      // the table values of every column pair, all loaded before the first is
      // used; re-read per spectrum at the opaque index, so the multipliers
      // derived from them are recomputed rather than held across spectra
      double kxv[kRowPairs], kxmv[kRowPairs], ikv[kRowPairs], omv[kRowPairs];
#pragma unroll
      for (int u = 0; u < kRowPairs; ++u) {
        const int i = min(tq + u * kRowThreads, nx - 1);
        const int im = (nx - i) & (nx - 1);
        kxv[u] = a.kx[i];
        kxmv[u] = a.kx[im];
        ikv[u] = need_ik ? a.ikt[(size_t)j * nx + i] : 0.;
        omv[u] = (need_om || hmtf) ? a.omega[(size_t)j * nx + i] : 0.;
      }
#pragma unroll
      for (int u = 0; u < kRowPairs; ++u) {
        const bool in = tq + u * kRowThreads < nx;
        const int i = min(tq + u * kRowThreads, nx - 1);
        const int im = (nx - i) & (nx - 1);
        const double kx = kxv[u], kxm = kxmv[u];
        if (kxm == -kx && (double)kyjm == -(double)kyj) {
          // the packed multiplier mA + i mB at -k is conj(mA(k)) + i conj(mB(k))
          const double2 h = H[u];
          // 1/|k| (0 off the good-k set) from its table: |-k| = |k| here
          const double ik = ikv[u];
          // omega(|k|) = omega(|-k|): from the dispersion table when a multiplier needs it
          const double om = omv[u];
          if (need_g) {
            // G(a) from a(k) and a(-k), recomputed; the thread of the mirror
            // column of a self-paired row computes -conj(G) to the bit
            double2 A0, A1;
            phased_pair<MIRROR>(a, j, i, jm, im, t, A0, A1);
            const double2 g = make_double2(0.5 * (A0.x - A1.x), 0.5 * (A0.y + A1.y));
            double2 ha, hb;
            if (hmtf) {
              // M_h(-k) = M_h(k): H(M_h a) = Re(M_h) H(a) + i Im(M_h) G(a)
              const double2 M = hmtf_mult((float)kx, kyj, om);
              ha = make_double2(M.x * h.x - M.y * g.y, M.x * h.y + M.y * g.x);
            } else {
              ha = cmul2(mono_c(da, mono_x(q.ca, kx, ik, om) * ya), anti_a ? g : h);
            }
            hb = cmul2(mono_c(db, mono_x(q.cb, kx, ik, om) * yb), anti_b ? g : h);
            if (in) {
              // ha + i hb at k, conj(ha) + i conj(hb) at -k
              put(i, make_double2(ha.x - hb.y, ha.y + hb.x), (nrow - 1) * nx + im,
                  make_double2(ha.x + hb.y, hb.x - ha.y));
            }
          } else {
            const double fa = mono_x(q.ca, kx, ik, om) * ya, fb = mono_x(q.cb, kx, ik, om) * yb;
            const double ar = da ? 0. : fa, ai = da ? fa : 0.;
            const double br = db ? 0. : fb, bi = db ? fb : 0.;
            if (in) {
              put(i, cmul2(make_double2(ar - bi, ai + br), h), (nrow - 1) * nx + im,
                  cmul2(make_double2(ar + bi, br - ai), make_double2(h.x, -h.y)));
            }
          }
        } else if (in) {
          general |= 1u << u;
        }
        // one column pair at a time: interleaving all of them would need more
        // than the 128 registers two workgroups per CU leave
        __builtin_amdgcn_sched_barrier(0);
      }
      // A Nyquist row or column is its own mirror: the general form, from a(k)
      // and a(-k) themselves (one column pair of each row pair, or the whole
      // Nyquist row). Kept out of the unrolled loop above.
#pragma unroll 1
      for (int u = 0; general != 0; ++u, general >>= 1) {
        if (!(general & 1u)) continue;
        const int i = tq + u * kRowThreads;
        const int im = (nx - i) & (nx - 1);
        const double kx = a.kx[i], kxm = a.kx[im], ky = kyj, kym = kyjm;
        double2 A0, A1;
        double k0;
        if (i == nx / 2 && j != jm) {
          A0 = nyq[0];
          A1 = nyq[1];
        } else if (MIRROR) {
          // self-paired row: a(-k) is this row's coefficient at column im
          A0 = S[pad(nx + i)];
          A1 = S[pad(nx + im)];
        } else {
          phased_pair<MIRROR>(a, j, i, jm, im, t, A0, A1);
        }
        (void)ky;
        (void)kym;
        const double ik = a.ikt[(size_t)j * nx + i];
        const double ikm = a.ikt[(size_t)jm * nx + im];
        const double om0 = (need_om || hmtf) ? a.omega[(size_t)j * nx + i] : 0.;
        const double omm = (need_om || hmtf) ? a.omega[(size_t)jm * nx + im] : 0.;
        const double yam = mono_y(q.ca, kym), ybm = mono_y(q.cb, kym);
        const double2 mak = hmtf ? hmtf_mult((float)kx, (float)ky, om0)
                                 : mono_c(da, mono_x(q.ca, kx, ik, om0) * ya);
        const double2 mbk = mono_c(db, mono_x(q.cb, kx, ik, om0) * yb);
        const double2 mam = hmtf ? hmtf_mult((float)kxm, (float)kym, omm)
                                 : mono_c(da, mono_x(q.ca, kxm, ikm, omm) * yam);
        const double2 mbm = mono_c(db, mono_x(q.cb, kxm, ikm, omm) * ybm);
        const double2 ha = herm(mak, A0, mam, A1);
        const double2 hb = herm(mbk, A0, mbm, A1);
        const double2 hc = herm(mam, A1, mak, A0);
        const double2 hd = herm(mbm, A1, mbk, A0);
        // ha + i hb at k, hc + i hd at -k
        put(i, make_double2(ha.x - hb.y, ha.y + hb.x), (nrow - 1) * nx + im,
            make_double2(hc.x - hd.y, hc.y + hd.x));
      }
      __syncthreads();
      // the paired spectrum's component a stays in the buffer for ta + 1
      if (mode == 1 && ns == 2) continue;
      lds_ifft_any<kRowThreads, 2 * kMaxN / kRowThreads, 3, SHORT>(S, nx, logx, nrow, tw);
      // column-blocked plane W[x / cbw][y][x % cbw] of this spectrum and snapshot
      // (the paired spectrum's at its first snapshot ta)
      unsigned char* wp =
          a.work + (size_t)(mode ? ta : ti) * (size_t)a.snap_bytes + (size_t)q.off;
      const int cbw = q.cbw, lcbw = q.lcbw;
      if (q.f32) {
        float2* w = (float2*)wp;
        for (int e = tq; e < nrow * nx; e += kRowThreads) {
          const int r = e >> logx, i = e & (nx - 1);
          const int row = r == 0 ? j : jm;
          const int xb = i >> lcbw, c = i & (cbw - 1);
          store_stream32(&w[((size_t)xb * a.ny + row) * cbw + c], S[pad(e)]);
        }
      } else {
        double2* w = (double2*)wp;
        for (int e = tq; e < nrow * nx; e += kRowThreads) {
          const int r = e >> logx, i = e & (nx - 1);
          const int row = r == 0 ? j : jm;
          const int xb = i >> lcbw, c = i & (cbw - 1);
          store_stream(&w[((size_t)xb * a.ny + row) * cbw + c], S[pad(e)]);
        }
      }
      __syncthreads();
    }
  }
}

// Pass 2: one workgroup per (block of cb columns, packed spectrum, snapshot),
// in XCD-contiguous order so that the blocks sharing a workspace run (cbw
// columns of a row) are read behind the same L2; FFT along y in LDS, then
// scaled by 1/(Nx Ny) and written as float32 fields, x-major (field[x][y]).
template <bool SHORT>
__global__ __launch_bounds__(kColThreads, kColWaves) void synth_cols_kernel(SynthArgs a) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const int cb = a.cb;
  double2* C = (double2*)lds_raw;  // [cb][ny]
  double2* tw = C + padded((size_t)a.ny * cb);  // [128] twiddles
  load_twiddles(tw, a.twy);
  // items: (block, spectrum, snapshot) of the spectra taken at every snapshot,
  // then (block, snapshot pair) of the time-paired spectrum (the plan's last)
  const int nxb = a.nx >> a.lcb;
  const int nq0 = a.nq - a.tpair;
  const int n0 = nxb * nq0 * a.n_times;
  const int item = osar::xcd_item(blockIdx.x, n0 + a.tpair * nxb * ((a.n_times + 1) >> 1));
  int xb, qi, ti;
  if (item < n0) {
    xb = item % nxb;
    qi = (item / nxb) % nq0;
    ti = item / (nxb * nq0);
  } else {
    xb = (item - n0) % nxb;
    qi = a.nq - 1;
    ti = 2 * ((item - n0) / nxb);
  }
  const int x0 = xb * cb;
  const int tid = threadIdx.x;
  const int ny = a.ny, nx = a.nx;
  const QPlan q = plan_at(a, qi);
  const int cbw = q.cbw;
  const size_t wofs = (size_t)(x0 >> q.lcbw) * ny * cbw + (x0 & (cbw - 1));
  const unsigned char* wp = a.work + (size_t)ti * (size_t)a.snap_bytes + (size_t)q.off;
  // every load of the block in flight before the first LDS store: the loads
  // are unconditional (indices past the block clamped to its last element)
  // and complex64 elements are widened after all of them are issued, so no
  // load waits for the one before it
  constexpr int kFill = kColElems / kColThreads;
  const int nblk = ny * cb;
  double2 fill[kFill];
  if (q.f32) {
    const float2* w = (const float2*)wp + wofs;
    nt_float2 v32[kFill];
#pragma unroll
    for (int k = 0; k < kFill; ++k) {
      const int e = min(tid + k * kColThreads, nblk - 1);
      v32[k] = __builtin_nontemporal_load(
          (const nt_float2*)&w[(size_t)(e >> a.lcb) * cbw + (e & (cb - 1))]);
    }
#pragma unroll
    for (int k = 0; k < kFill; ++k) fill[k] = make_double2((double)v32[k].x, (double)v32[k].y);
  } else {
    const double2* w = (const double2*)wp + wofs;
#pragma unroll
    for (int k = 0; k < kFill; ++k) {
      const int e = min(tid + k * kColThreads, nblk - 1);
      fill[k] = load_stream(&w[(size_t)(e >> a.lcb) * cbw + (e & (cb - 1))]);
    }
  }
#pragma unroll
  for (int k = 0; k < kFill; ++k) {
    const int e = tid + k * kColThreads;
    if (e < nblk) C[pad((e & (cb - 1)) * ny + (e >> a.lcb))] = fill[k];
  }
  __syncthreads();
  lds_ifft_any<kColThreads, kColElems / kColThreads, kColRadixLog, SHORT>(C, ny, a.logy, cb, tw);
  const double scale = 1.0 / ((double)nx * (double)ny);
  const double t = a.times[ti];
  const size_t plane = (size_t)nx * ny;
  if (q.keep) {
    // back into this block's own workspace positions (row y, column c), scaled
    double2* w = (double2*)wp + wofs;
    for (int e = tid; e < ny * cb; e += kColThreads) {
      const int c = e >> a.logy, y = e & (ny - 1);
      const double2 z = C[pad(e)];
      store_stream(&w[(size_t)y * cbw + c], make_double2(z.x * scale, z.y * scale));
    }
    return;
  }
  float* const outa = q.pa;
  float* const outb = q.pb;
  // a time-paired spectrum's component b is its field at the next snapshot
  const int tb = ti + q.tpair;
  const size_t ob = q.tpair ? plane : 0;
  const double t_b = tb < a.n_times ? a.times[tb] : t;
  for (int e = tid; e < ny * cb; e += kColThreads) {
    const int c = e >> a.logy, y = e & (ny - 1);  // consecutive threads walk y: contiguous stores
    const double2 z = C[pad(e)];
    const double re = z.x * scale, im = z.y * scale;
    const size_t o = (size_t)ti * plane + (size_t)(x0 + c) * ny + y;
    if (q.choppy) {
      __builtin_nontemporal_store((float)(re / (1. + im)), &outa[o]);
      continue;
    }
    if (q.outa >= 0) {
      double v = re;
      if (q.outa == 0) v = v + a.cur_x * t;   // Dx, Dy drift with the current
      if (q.outa == 1) v = v + a.cur_y * t;
      if (q.outa == 8) v = v + a.cur_x;       // Vx, Vy carry it
      if (q.outa == 9) v = v + a.cur_y;
      __builtin_nontemporal_store((float)v, &outa[o]);
    }
    if (q.outb >= 0 && tb < a.n_times) {
      double v = im;
      if (q.outb == 0) v = v + a.cur_x * t_b;
      if (q.outb == 1) v = v + a.cur_y * t_b;
      if (q.outb == 8) v = v + a.cur_x;
      if (q.outb == 9) v = v + a.cur_y;
      __builtin_nontemporal_store((float)v, &outb[o + ob]);
    }
  }
}
// Choppy curvatures (ocean.py:462-475) from the FP64 keep spectra of the
// column pass, per grid point and snapshot:
//   Diffx = a_x / (1 + a_xx), Diffy = a_y / (1 + a_yy),
//   Diffxx = ((1 + a_xx) Dxx - a_xxx a_x) / (1 + a_xx)^3,
//   Diffyy = ((1 + a_yy) Dyy - a_yyy a_y) / (1 + a_yy)^3,
//   Diffxy = ((1 + a_xx) Dxy - a_xxy a_x) / ((1 + a_xx)^2 (1 + a_yy)).
// Thread = one (x, y) with y fastest, so the float32 stores are contiguous.
__global__ __launch_bounds__(256) void synth_choppy_kernel(SynthArgs a) {
  const int64_t plane = (int64_t)a.nx * a.ny;
  const int64_t n = plane * a.n_times;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int ti = (int)(e / plane);
    const int64_t r = e - (int64_t)ti * plane;
    const int x = (int)(r >> a.logy), y = (int)(r & (a.ny - 1));
    double2 v[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      const QPlan q = plan_at(a, a.ch_q[k]);
      const double2* w = (const double2*)(a.work + (size_t)ti * (size_t)a.snap_bytes +
                                          (size_t)q.off);
      v[k] = load_stream(&w[((size_t)(x >> q.lcbw) * a.ny + y) * q.cbw + (x & (q.cbw - 1))]);
    }
    const double ax = v[0].x, axx = v[0].y, ay = v[1].x, ayy = v[1].y;
    const double ux = 1. + axx, uy = 1. + ayy;
    const size_t o = (size_t)e;
    if (a.ch_out[0]) __builtin_nontemporal_store((float)(ax / ux), &a.ch_out[0][o]);
    if (a.ch_out[1]) __builtin_nontemporal_store((float)(ay / uy), &a.ch_out[1][o]);
    if (a.ch_out[2])
      __builtin_nontemporal_store((float)((ux * v[2].x - v[2].y * ax) / (ux * ux * ux)),
                                  &a.ch_out[2][o]);
    if (a.ch_out[3])
      __builtin_nontemporal_store((float)((uy * v[3].x - v[3].y * ay) / (uy * uy * uy)),
                                  &a.ch_out[3][o]);
    if (a.ch_out[4])
      __builtin_nontemporal_store((float)((ux * v[4].x - v[4].y * ax) / (ux * ux * uy)),
                                  &a.ch_out[4][o]);
  }
}

int ilog2(int n) {
  int l = 0;
  while ((1 << l) < n) ++l;
  return (1 << l) == n ? l : -1;
}

// Monomial code of a multiplier (QPlan::ca / cb).
int mono_code(int m) {
  auto c = [](int px, int py, int pk, int imag, int neg, int po = 0) {
    return px | (py << 2) | (pk << 4) | (imag ? kMonoImag : 0) | (neg ? kMonoNeg : 0) |
           po * kMonoOmega;
  };
  switch (m) {
    case M_ONE: return c(0, 0, 0, 0, 0);
    case M_IKXN: return c(1, 0, 1, 1, 0);
    case M_IKYN: return c(0, 1, 1, 1, 0);
    case M_IKX: return c(1, 0, 0, 1, 0);
    case M_IKY: return c(0, 1, 0, 1, 0);
    case M_NKX2: return c(2, 0, 0, 0, 1);
    case M_NKY2: return c(0, 2, 0, 0, 1);
    case M_NKXKY: return c(1, 1, 0, 0, 1);
    case M_NKX2KI: return c(2, 0, 1, 0, 1);
    case M_NKY2KI: return c(0, 2, 1, 0, 1);
    case M_NIKX3KI: return c(3, 0, 1, 1, 1);
    case M_NIKY3KI: return c(0, 3, 1, 1, 1);
    case M_NIKX2KYKI: return c(2, 1, 1, 1, 1);
    case M_WKXN: return c(1, 0, 1, 0, 0, 1);
    case M_WKYN: return c(0, 1, 1, 0, 0, 1);
    case M_NIW: return c(0, 0, 0, 1, 1, 1);
    case M_NIW2KXN: return c(1, 0, 1, 1, 1, 2);
    case M_NIW2KYN: return c(0, 1, 1, 1, 1, 2);
    case M_NW2: return c(0, 0, 0, 0, 1, 2);
    case M_HMTF: return kMonoHmtf;
    default: return kMonoNone;
  }
}

// A spectrum of components a and b (multipliers ma, mb into fields outa, outb).
QPlan spectrum(int ma, int mb, int outa, int outb) {
  QPlan p{};
  p.ma = ma;
  p.mb = mb;
  p.outa = outa;
  p.outb = outb;
  return p;
}

// The spectra of a descriptor. ch_q (may be NULL): plan indices of the five
// keep spectra of the choppy curvatures, -1 when there are none.
int make_plan(const osar_surface_desc* d, QPlan* q_out, int* nq, int* ch_q = nullptr) {
  // Real outputs in pairing order with their multipliers. At most 10 spectra
  // result (14 monomial fields pair into 7 and M_h takes one; choppy slopes
  // take one each but then no curvatures; choppy curvatures take five, from
  // which the slopes come too, beside at most 9 monomial fields and M_h); q
  // has room for more.
  struct Item { int out, m; };
  QPlan q[2 * kMaxQ];
  static const int kMult[kNumFields] = {M_IKXN, M_IKYN, M_ONE, M_IKX, M_IKY, M_NKX2, M_NKY2,
                                        M_NKXKY, M_WKXN, M_WKYN, M_NIW, M_NIW2KXN, M_NIW2KYN,
                                        M_NW2, M_HMTF};
  // Dz, Dx, Dy, Diffx, Diffy, Diffxx, Diffyy, Diffxy, Vx, Vy, Vz, Ax, Ay, Az
  static const int kOrder[kNumFields - 1] = {2, 0, 1, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13};
  std::vector<Item> items;
  int n = 0;
  if (d->fields & ~((1 << kNumFields) - 1)) {
    osar_set_error("unknown field bits 0x%x", d->fields & ~((1 << kNumFields) - 1));
    return OSAR_EINVAL;
  }
  const bool curv = (d->fields & (OSAR_F_DIFFXX | OSAR_F_DIFFYY | OSAR_F_DIFFXY)) != 0;
  const bool choppy2 = d->choppy && curv;
  int chq[5] = {-1, -1, -1, -1, -1};
  if (choppy2) {
    // ocean.py:462-475: a_x, a_xx | a_y, a_yy | Dxx, a_xxx | Dyy, a_yyy | Dxy, a_xxy
    static const int kA[5] = {M_IKX, M_IKY, M_NKX2, M_NKY2, M_NKXKY};
    static const int kB[5] = {M_NKX2KI, M_NKY2KI, M_NIKX3KI, M_NIKY3KI, M_NIKX2KYKI};
    for (int k = 0; k < 5; ++k) {
      chq[k] = n;
      q[n] = spectrum(kA[k], kB[k], -1, -1);
      q[n++].keep = 1;
    }
  }
  for (int oi = 0; oi < kNumFields - 1; ++oi) {
    const int f = kOrder[oi];
    if (!(d->fields & (1 << f))) continue;
    if (choppy2 && f >= 3 && f <= 7) continue;   // from the keep spectra
    if (d->choppy && (f == 3 || f == 4)) {
      q[n] = spectrum(f == 3 ? M_IKX : M_IKY, f == 3 ? M_NKX2KI : M_NKY2KI, f, -1);
      q[n++].choppy = 1;
      continue;
    }
    items.push_back(Item{f, kMult[f]});
  }
  // slope_fp32: Diffx and Diffy share one complex64 spectrum, the other fields
  // pair up in complex128 spectra
  std::vector<Item> lo;
  if (d->slope_fp32 && !d->choppy) {
    std::vector<Item> keep;
    for (const Item& it : items) (it.out == 3 || it.out == 4 ? lo : keep).push_back(it);
    items.swap(keep);
  }
  // M_h is not a monomial: it is component a of its spectrum, whose component b
  // is the odd monomial field out, if any
  if (d->fields & OSAR_F_HMTF) {
    QPlan h = spectrum(M_HMTF, M_NONE, 14, -1);
    if (items.size() % 2) {
      h.mb = items.back().m;
      h.outb = items.back().out;
      items.pop_back();
    }
    q[n++] = h;
  }
  // A field left without a partner pairs with itself at the next snapshot
  // (time-paired spectrum, the plan's last): 2.5 instead of 3 transforms and
  // 8 B less workspace traffic per point and snapshot for five fields. Only a
  // monomial with m(-k) = conj(m(k)) and no omega, so that one H(a) per
  // snapshot serves both halves.
  Item lone{-1, M_NONE};
  if (items.size() % 2 && d->n_times >= 2) {
    const int c = mono_code(items.back().m);
    if (!(c & (3 * kMonoOmega | kMonoHmtf | kMonoNone)) && !mono_anti(c)) {
      lone = items.back();
      items.pop_back();
    }
  }
  for (size_t i = 0; i < items.size(); i += 2) {
    const bool two = i + 1 < items.size();
    q[n++] = spectrum(items[i].m, two ? items[i + 1].m : (int)M_NONE, items[i].out,
                      two ? items[i + 1].out : -1);
  }
  if (!lo.empty()) {
    const bool two = lo.size() > 1;
    q[n] = spectrum(lo[0].m, two ? lo[1].m : (int)M_NONE, lo[0].out, two ? lo[1].out : -1);
    q[n++].f32 = 1;
  }
  if (lone.out >= 0) {
    q[n] = spectrum(lone.m, lone.m, lone.out, lone.out);
    q[n++].tpair = 1;
  }
  if (n > kMaxQ) {
    osar_set_error("%d spectra exceed the %d a synthesis call takes", n, kMaxQ);
    return OSAR_EUNSUPPORTED;
  }
  for (int i = 0; i < n; ++i) {
    q[i].ca = mono_code(q[i].ma);
    q[i].cb = mono_code(q[i].mb);
    q_out[i] = q[i];
  }
  *nq = n;
  if (ch_q)
    for (int k = 0; k < 5; ++k) ch_q[k] = chq[k];
  return OSAR_OK;
}

// The row-pass variant: SHORT (padded addressing for nx >= 64), MIRROR, TIMED.
const void* rows_kernel(bool short_x, bool mirror, bool timed) {
  static const void* const k[8] = {
      (const void*)synth_rows_kernel<false, false, false>,
      (const void*)synth_rows_kernel<false, false, true>,
      (const void*)synth_rows_kernel<false, true, false>,
      (const void*)synth_rows_kernel<false, true, true>,
      (const void*)synth_rows_kernel<true, false, false>,
      (const void*)synth_rows_kernel<true, false, true>,
      (const void*)synth_rows_kernel<true, true, false>,
      (const void*)synth_rows_kernel<true, true, true>};
  return k[(short_x ? 4 : 0) + (mirror ? 2 : 0) + (timed ? 1 : 0)];
}

// Column block (columns per column-pass workgroup) of an ny-point transform.
int col_block(int nx, int ny) {
  int cb = kColElems / ny;
  if (cb > 8) cb = 8;
  if (cb > nx) cb = nx;
  return cb;
}

// Workspace layout: per snapshot, one plane per spectrum, column-blocked with
// runs of 128 bytes per row (8 complex128 or 16 complex64 columns, at least the
// column block, at most nx). Returns the bytes per snapshot.
long long plan_layout(int nx, int ny, QPlan* q, int nq) {
  const int cb = col_block(nx, ny);
  long long off = 0;
  for (int i = 0; i < nq; ++i) {
    int w = q[i].f32 ? 2 * kWsBlock : kWsBlock;
    if (w < cb) w = cb;
    if (w > nx) w = nx;
    q[i].cbw = w;
    q[i].lcbw = ilog2(w);
    q[i].off = off;
    off += (long long)nx * ny * (q[i].f32 ? 8 : 16);
  }
  return off;
}

// Two-level twiddle tables of exp(+2 pi i k / N) for every length N = 16 ..
// 8192 (entries 0..63 hold k = 0..63, entries 64..191 hold k = 64 m),
// computed on the host in long double and copied to each device once, when
// the first context of that device is created (osar_synth_device_setup).
constexpr int kTwLens = 10;   // log2 N = 4 .. 13
__device__ double2 g_twiddles[kTwLens][kTwEntries];

// Scratch of osar_surface_synth_noalloc: its synthesis workspace, one per
// device for the life of the process (every call on a device shares it).
static __device__ __align__(256) unsigned char scratch[OSAR_SYNTH_SCRATCH_BYTES];

struct DeviceTables {
  bool ready = false;
  double2* twiddles = nullptr;     // device address of g_twiddles
  unsigned char* scratch = nullptr;  // device address of scratch
};
constexpr int kMaxDevices = 64;
DeviceTables g_dev[kMaxDevices];
std::mutex g_dev_mu;

int twiddles(osar_ctx* ctx, int n, const double2** out) {
  const int l = ilog2(n);
  OSAR_REQUIRE(ctx->device >= 0 && ctx->device < kMaxDevices && l >= 4 && l < 4 + kTwLens &&
                   g_dev[ctx->device].ready,
               OSAR_EUNSUPPORTED, "twiddle table for device %d, n %d", ctx->device, n);
  *out = g_dev[ctx->device].twiddles + (size_t)(l - 4) * kTwEntries;
  return OSAR_OK;
}

}  // namespace

// Device address of the twiddle table of length n on a set-up device.
int osar_twiddles(osar_ctx* ctx, int n, const double2** out) { return twiddles(ctx, n, out); }

// Once per device, with that device current: the device addresses of the
// twiddle tables and of the no-allocation scratch, and the twiddle values.
int osar_synth_device_setup(int device) {
  OSAR_REQUIRE(device >= 0 && device < kMaxDevices, OSAR_EUNSUPPORTED,
               "device %d beyond the %d the library tracks", device, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_dev_mu);
  DeviceTables& t = g_dev[device];
  if (t.ready) return OSAR_OK;
  void* tw = nullptr;
  void* sc = nullptr;
  OSAR_HIP(hipGetSymbolAddress(&tw, HIP_SYMBOL(g_twiddles)));
  OSAR_HIP(hipGetSymbolAddress(&sc, HIP_SYMBOL(scratch)));
  std::vector<double2> h((size_t)kTwLens * kTwEntries);
  const long double pi = 3.14159265358979323846264338327950288L;
  for (int l = 4; l < 4 + kTwLens; ++l) {
    const long n = 1L << l;
    for (int e = 0; e < kTwEntries; ++e) {
      const long k = e < 64 ? e : 64L * (e - 64);
      const long double ang = 2.0L * pi * (long double)(k % n) / (long double)n;
      h[(size_t)(l - 4) * kTwEntries + e] = make_double2((double)cosl(ang), (double)sinl(ang));
    }
  }
  OSAR_HIP(hipMemcpy(tw, h.data(), sizeof(double2) * h.size(), hipMemcpyHostToDevice));
  t.twiddles = (double2*)tw;
  t.scratch = (unsigned char*)sc;
  t.ready = true;
  return OSAR_OK;
}

extern "C" int64_t osar_surface_workspace_bytes(const osar_surface_desc* d) {
  if (!d) return -1;
  if (ilog2(d->nx) < 0 || ilog2(d->ny) < 0) return -1;
  QPlan q[kMaxQ];
  int nq = 0;
  if (make_plan(d, q, &nq) != OSAR_OK) return -1;
  // packed spectra [n_times][nq planes], then the omega and 1/|k| tables [ny][nx]
  return (int64_t)plan_layout(d->nx, d->ny, q, nq) * d->n_times +
         2 * (int64_t)d->nx * d->ny * (int64_t)sizeof(double);
}

namespace {

// The synthesis of descriptor d with the workspace `work` (at least
// osar_surface_workspace_bytes(d) bytes), or with the context's workspace,
// grown on demand, when work is NULL.
int surface_synth(osar_ctx* ctx, const osar_surface_desc* d, void* stream, unsigned char* work,
                  size_t work_bytes) {
  OSAR_PTR(ctx);
  OSAR_PTR(d);
  const int lx = ilog2(d->nx), ly = ilog2(d->ny);
  OSAR_REQUIRE(lx >= 4 && ly >= 4 && d->nx <= kMaxN && d->ny <= kMaxN, OSAR_EUNSUPPORTED,
               "surface %dx%d: sizes must be powers of two in [16, %d]", d->nx, d->ny, kMaxN);
  OSAR_REQUIRE(d->n_times >= 1, OSAR_EINVAL, "n_times must be >= 1");
  OSAR_PTR(d->kx);
  OSAR_PTR(d->ky);
  OSAR_PTR(d->coefs);
  OSAR_PTR(d->times);
  SynthArgs a{};
  OSAR_REQUIRE(make_plan(d, a.q, &a.nq, a.ch_q) == OSAR_OK, OSAR_EUNSUPPORTED, "%s",
               osar_last_error());
  if (a.nq == 0) return OSAR_OK;
  {
    static const int kChOut[5] = {3, 4, 5, 6, 7};   // Diffx, Diffy, Diffxx, Diffyy, Diffxy
    for (int k = 0; k < 5; ++k)
      a.ch_out[k] = a.ch_q[0] >= 0 && (d->fields & (1 << kChOut[k])) ? d->out[kChOut[k]] : nullptr;
  }
  for (int f = 0; f < kNumFields; ++f) {
    if (d->fields & (1 << f)) {
      OSAR_REQUIRE(d->out[f] != nullptr, OSAR_EINVAL, "output field %d requested but NULL", f);
    }
  }
  for (int i = 0; i < a.nq; ++i) {
    a.q[i].pa = a.q[i].outa >= 0 ? d->out[a.q[i].outa] : nullptr;
    a.q[i].pb = a.q[i].outb >= 0 ? d->out[a.q[i].outb] : nullptr;
  }
  OSAR_HIP(hipSetDevice(ctx->device));
  const size_t need = (size_t)osar_surface_workspace_bytes(d);
  if (work != nullptr) {
    OSAR_REQUIRE(need <= work_bytes, OSAR_EINVAL,
                 "surface %dx%d with %d snapshots needs %zu B of workspace; the "
                 "no-allocation scratch holds %zu B (split the snapshots over calls)",
                 d->nx, d->ny, d->n_times, need, work_bytes);
  } else {
    if (need > ctx->work_bytes) {
      if (ctx->work) OSAR_HIP(hipFree(ctx->work));
      ctx->work = nullptr;
      ctx->work_bytes = 0;
      OSAR_HIP(hipMalloc(&ctx->work, need));
      ctx->work_bytes = need;
    }
    work = (unsigned char*)ctx->work;
  }
  a.nx = d->nx;
  a.ny = d->ny;
  a.logx = lx;
  a.logy = ly;
  a.depth = d->depth;
  a.cur_x = d->current_x;
  a.cur_y = d->current_y;
  a.thr = (d->kx_res < d->ky_res ? d->kx_res : d->ky_res) / 2.0;
  a.kx = d->kx;
  a.ky = d->ky;
  a.coefs = d->coefs;
  a.times = d->times;
  a.snap_bytes = plan_layout(d->nx, d->ny, a.q, a.nq);
  a.work = work;
  a.omega = (double*)(a.work + (size_t)a.snap_bytes * d->n_times);
  a.ikt = a.omega + (size_t)d->nx * d->ny;
  int rc = twiddles(ctx, d->nx, &a.twx);
  if (rc) return rc;
  rc = twiddles(ctx, d->ny, &a.twy);
  if (rc) return rc;
  const int cb = col_block(d->nx, d->ny);
  a.cb = cb;
  a.lcb = ilog2(cb);
  hipStream_t s = (hipStream_t)stream;
  const size_t lds_rows = (padded(2 * (size_t)d->nx) + 128 + 2) * sizeof(double2);
  const size_t lds_cols = (padded((size_t)cb * d->ny) + 128) * sizeof(double2);
  // short padded addressing for transforms of 64 points or more
  const bool short_x = d->nx >= 64, short_y = d->ny >= 64;
  const bool mirror = d->mirror_grid != 0;
  bool timed = false;
  for (int i = 0; i < a.nq; ++i)
    timed = timed || (((a.q[i].ca | a.q[i].cb) & (3 * kMonoOmega | kMonoHmtf)) != 0);
  const void* rows_k = rows_kernel(short_x, mirror, timed);
  const void* cols_k = short_y ? (const void*)synth_cols_kernel<true>
                               : (const void*)synth_cols_kernel<false>;
  {
    const int rc1 = osar_reserve_lds(rows_k, lds_rows);
    if (rc1) return rc1;
    const int rc2 = osar_reserve_lds(cols_k, lds_cols);
    if (rc2) return rc2;
  }
  a.n_times = d->n_times;
  a.tpair = a.q[a.nq - 1].tpair;
  {
    const int64_t n = (int64_t)d->nx * d->ny;
    const int blocks = (int)((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096);
    hipLaunchKernelGGL(omega_kernel, dim3(blocks), dim3(256), 0, s, a);
    OSAR_LAUNCHED("omega_kernel");
  }
  // a time-paired plan: row items per snapshot pair, the paired spectrum's
  // column items per snapshot pair
  const int n_pair = (d->n_times + 1) / 2;
  const dim3 grid_rows((d->ny / 2 + 1) * (a.tpair ? n_pair : d->n_times));
  const dim3 grid_cols((d->nx / cb) * ((a.nq - a.tpair) * d->n_times + a.tpair * n_pair));
  {
    void* args[] = {&a};
    OSAR_HIP(hipLaunchKernel(rows_k, grid_rows, dim3(kRowThreads), args, lds_rows, s));
  }
  OSAR_LAUNCHED("synth_rows_kernel");
  if (short_y)
    hipLaunchKernelGGL(synth_cols_kernel<true>, grid_cols, dim3(kColThreads), lds_cols, s, a);
  else
    hipLaunchKernelGGL(synth_cols_kernel<false>, grid_cols, dim3(kColThreads), lds_cols, s, a);
  OSAR_LAUNCHED("synth_cols_kernel");
  if (a.ch_q[0] >= 0) {
    const int64_t n = (int64_t)d->nx * d->ny * d->n_times;
    const int blocks = (int)((n + 255) / 256 < 8192 ? (n + 255) / 256 : 8192);
    hipLaunchKernelGGL(synth_choppy_kernel, dim3(blocks), dim3(256), 0, s, a);
    OSAR_LAUNCHED("synth_choppy_kernel");
  }
  return OSAR_OK;
}

}  // namespace

extern "C" int osar_surface_synth(osar_ctx* ctx, const osar_surface_desc* d, void* stream) {
  return surface_synth(ctx, d, stream, nullptr, 0);
}

extern "C" int osar_surface_synth_noalloc(osar_ctx* ctx, const osar_surface_desc* d,
                                          void* stream) {
  OSAR_PTR(ctx);
  OSAR_REQUIRE(ctx->device >= 0 && ctx->device < kMaxDevices && g_dev[ctx->device].ready,
               OSAR_EINVAL, "context of device %d has no scratch", ctx->device);
  return surface_synth(ctx, d, stream, g_dev[ctx->device].scratch, OSAR_SYNTH_SCRATCH_BYTES);
}
