/*
 * oceansar_hip.h -- C ABI of liboceansar_hip.so, the MI355X (gfx950) kernels
 * behind the OCEANSAR raw-data generation path.
 *
 * The reference path is pure Python (numpy / numba); its "interfaces" are
 * Python call sites. Each entry point below replaces one of them:
 *
 *   osar_surface_synth   OceanSurface.t setter
 *                        (oceansar/surfaces/ocean.py:418-526)
 *   osar_facet_bragg     RCSRomeiser97.rcs -> rcs_numba -> _rcs_numba, plus the
 *                        sqrt(RCS) and NRCS sums of the driver
 *                        (oceansar/nrcs/romeiser97.py:95-107,251-393;
 *                         oceansar/radarsim/sar_raw_nompi.py:432-467)
 *   osar_echo_accumulate the per-pulse body of sar_raw from the surface fields
 *                        to the range lines: geometry, Bragg scene, beam,
 *                        channel phases and chan_profile_numba /
 *                        profile_integrator
 *                        (oceansar/radarsim/sar_raw_nompi.py:349-535;
 *                         oceansar/radarsim/range_profile.py:22-49,83-116)
 *   osar_echo_deposit    chan_profile_numba alone (range_profile.py:83-116):
 *                        deposit given complex facet values at given ranges
 *   osar_sar_focus       the processing loop of sar_focus for one channel
 *                        (oceansar/radarsim/sar_processor.py:131-225), the
 *                        consumer of the raw data
 *   osar_insar_coregister  the azimuth sub-sample shift and flat-earth
 *                        removal of insar_process
 *                        (oceansar/radarsim/insar_processor.py:165-185)
 *   osar_insar_multilook the multilooked intensities, interferograms and
 *                        coherences of insar_process (:192-214), and the
 *                        smoothing of ati_process (ati_processor.py:127-134)
 *
 * Conventions
 *   - Every pointer is a caller-owned device buffer; the library never frees
 *     caller memory. Workspace is owned by the context.
 *   - Return 0 on success, a negative OSAR_E* code on error; the message of
 *     the last error on the calling thread is osar_last_error().
 *   - Calls are asynchronous on the given HIP stream (a hipStream_t passed as
 *     void*, NULL = default stream).
 *   - One context per host thread or stream; contexts share no mutable state.
 */
#ifndef OCEANSAR_HIP_H
#define OCEANSAR_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OSAR_OK 0
#define OSAR_EINVAL (-1)        /* bad argument */
#define OSAR_ENOMEM (-2)        /* out of device memory */
#define OSAR_ELAUNCH (-3)       /* kernel launch / runtime failure */
#define OSAR_EUNSUPPORTED (-4)  /* configuration not supported */

#define OSAR_MAX_CH 8

typedef struct osar_ctx osar_ctx;

/* float2 / complex64 as laid out in memory: {re, im} */
typedef struct osar_c32 { float re, im; } osar_c32;

/* ---- context ---------------------------------------------------------- */
int osar_ctx_create(int device, osar_ctx** out);
int osar_ctx_destroy(osar_ctx* ctx);
const char* osar_last_error(void);
/* Library version: major*10000 + minor*100 + patch. */
int osar_version(void);
/* sizeof of the descriptor structs (0 echo, 1 bragg, 2 surface, 3 focus, 4 init,
   5 coregistration, 6 multilook), so
   that bindings can check their layout; -1 for an unknown id. */
int64_t osar_struct_size(int which);

/* ---- echo accumulation ------------------------------------------------ */
/* Flags of osar_echo_desc.flags */
#define OSAR_ECHO_BRAGG 1      /* add the two Bragg components */
#define OSAR_ECHO_XTI 2        /* apply the cross-track phase (any b_xti != 0) */
#define OSAR_ECHO_SPECULAR 4   /* add the Kirchhoff stationary-phase specular field
                                  (RCSKA.field 'spa', oceansar/nrcs/ka.py:105-284,
                                  used at sar_raw_nompi.py:401-428) */

typedef struct osar_echo_desc {
  /* sizes */
  int32_t nx, ny;           /* facets per row, rows */
  int32_t n_pulses_total;   /* pulses (rows) of the raw cube */
  int32_t pulse0, npulses;  /* window of pulses processed by this call */
  int32_t nbins;            /* bins per range line (rg_samp) */
  int32_t n_ch, n_pol;      /* channels (<= OSAR_MAX_CH), polarisations (1 or 2) */
  int32_t K;                /* surface snapshot cadence in pulses */
  int32_t snap0;            /* snapshot index stored in slot 0 of dx/dy/dz */
  int32_t n_snap_slots;     /* slots in dx/dy/dz */
  int32_t rcs0;             /* RCS refresh index stored in slot 0 of amp */
  int32_t n_rcs_slots;      /* slots in amp */
  int32_t flags;            /* OSAR_ECHO_* */
  /* launch shape (0 = library default) */
  int32_t pulses_per_block; /* 2, 4, 8 or 16 */
  int32_t rows_per_block;
  int32_t window_bins;      /* LDS range-line window per (pol, ch, pulse) */
  int32_t pol;              /* 0 = vv, 1 = hh, 2 = DP (line 0 hh, line 1 vv) */
  int32_t cols_per_block;   /* columns per workgroup (0 = all); below nx, sr_col must
                               be non-decreasing (the window starts at the chunk's
                               first column) */
  int32_t scat_len;         /* entries of scat_m, scat_p (N + n_pulses_total; the
                               larger if they differ), at most 2^29 (OSAR_ECHO_BRAGG) */
  /* FP64 scalars */
  double t_step, t_span, v_ground, sr0, gr0, k0;
  double min_sr, binsize;   /* bin position x = (sr - min_sr) / binsize */
  double ltx_over_l, lrx_over_l;  /* antenna lengths over wavelength */
  double sr_col_min, sr_col_max;  /* range of sr_col */
  double resid_max;         /* bound on |-cos_inc Dz + sin_inc Dx + sin_az Dy| */
  double b_ati[OSAR_MAX_CH], b_xti[OSAR_MAX_CH];
  /* per-column FP64 tables [nx] */
  const double* sr_col;     /* slant range minus its minimum */
  const double* sin_inc;
  const double* cos_inc;
  const double* dinc_col;   /* inc - inc_angle */
  const double* cap_rate;   /* 2 k0 t_step c_b; Bragg phase is cap_rate * (p + 1) */
  /* per-row FP64 table [ny] */
  const double* y_row;
  /* surface snapshots, float32 [n_snap_slots][nx][ny] (x-major) */
  const float* dx;
  const float* dy;
  const float* dz;
  /* Bragg amplitudes sqrt(RCS), float32 [n_rcs_slots][n_pol][2][nx][ny];
     component 0 pairs with scat_m, component 1 with scat_p */
  const float* amp;
  const int32_t* rcs_slot;  /* [n_pulses_total] RCS refresh index in force:
                               non-decreasing, at most +1 from one pulse to the next */
  /* scatterer series, extended: s_ext[i] = s[i mod N], i < N + n_pulses_total */
  const osar_c32* scat_m;
  const osar_c32* scat_p;
  const int32_t* ind0_m;    /* [nx][ny], 0 <= ind0 < N */
  const int32_t* ind0_p;
  const float* sinc_tab;    /* [161] range response table */
  /* output: accumulated, complex64 [n_pol][n_ch][n_pulses_total][nbins] */
  osar_c32* raw;
  /* specular (OSAR_ECHO_SPECULAR): slope and curvature snapshots, float32
     [n_snap_slots][nx][ny] like dx/dy/dz; incidence per column [nx]; facet size */
  const float* sx;          /* Diffx */
  const float* sy;          /* Diffy */
  const float* sxx;         /* Diffxx */
  const float* syy;         /* Diffyy */
  const float* sxy;         /* Diffxy */
  const double* inc_col;
  double facet_dx, facet_dy;
  /* out: accumulated sum of |E_spec|^2 per pulse [n_pulses_total], of the last
     polarisation (vv when present), as the driver's NRCS_avg does */
  double* spec_energy;
  /* osar_echo_accumulate_fp64 only: FP64 amplitudes [n_rcs_slots][n_pol][2][nx][ny]
     (osar_facet_bragg with fp64 = 1), complex128 scatterer series (interleaved
     re, im, extended like scat_m / scat_p) and the complex128 raw cube
     [n_pol][n_ch][n_pulses_total][nbins] */
  const double* amp64;
  const double* scat64_m;
  const double* scat64_p;
  double* raw64;
} osar_echo_desc;

int osar_echo_accumulate(osar_ctx* ctx, const osar_echo_desc* d, void* stream);

/* Persistent batched accumulation: the launches of n descriptors (scenes or
   chunks of any sizes, 1 to 4 channels each; one polarisation mode and one
   pulses_per_block, 4 or 8, for all; no specular or cross-track terms) in ONE
   persistent launch. No counterpart in the reference (which runs one scene per
   process); it replaces n osar_echo_accumulate calls with identical results
   up to the float32 order of the raw-cube additions. The grid holds as many
   workgroups as fit the device at once (or `workgroups`); each takes work
   items (P pulses x R rows x a column chunk of one descriptor) from a queue
   per XCD until the queue is empty, so small and large scenes share the
   device. The plan (descriptors, item tables, zeroed counters) is written by
   the call into the caller's pinned host buffer and copied to the caller's
   device buffer on the stream: both must stay untouched until the stream
   has run the launch (for a captured graph: as long as the graph is
   replayed); concurrent batches need their own plan buffers. */
#define OSAR_ECHO_BATCH_MAX 4096
typedef struct osar_echo_batch_desc {
  int32_t n;                    /* descriptors, 1 .. OSAR_ECHO_BATCH_MAX */
  int32_t workgroups;           /* grid size, 0: as many as fit the device at once */
  const osar_echo_desc* descs;  /* host array [n], checked as osar_echo_accumulate checks */
  void* plan_host;              /* pinned host buffer of osar_echo_batch_plan_bytes(n) */
  void* plan_dev;               /* device buffer of the same size */
} osar_echo_batch_desc;

/* Plan bytes of an n-descriptor batch, -1 for n out of range. */
int64_t osar_echo_batch_plan_bytes(int32_t n);
int osar_echo_accumulate_batch(osar_ctx* ctx, const osar_echo_batch_desc* b, void* stream);

/* The same accumulation with every value FP64 (the inputs the reference keeps
   in float32 aside: surface fields and the range-response table), into raw64:
   the precision reference of osar_echo_accumulate, at the same call site
   (sar_raw_nompi.py:349-535). Launch-shape fields are ignored. The FP64 lines
   of a pulse (n_pol x n_ch x nbins complex128) must fit the LDS. */
int osar_echo_accumulate_fp64(osar_ctx* ctx, const osar_echo_desc* d, void* stream);

/* Zero-fill n complex64 values (a kernel, so it may be captured in a graph). */
int osar_fill_zero_c32(osar_ctx* ctx, osar_c32* p, int64_t n, void* stream);

/* chan_profile_numba: out[d - 4 + j] += tab[sb + 20 j] * scene[i] for every
   facet i, with d = floor((srg - min_sr)/binsize) and
   sb = rint(20 (1 - frac)); taps outside [0, nbins) are dropped.
   srg float64 [n], scene complex64 [n], out complex64 [nbins]. */
int osar_echo_deposit(osar_ctx* ctx, const double* srg, const osar_c32* scene,
                      int64_t n, double binsize, double min_sr,
                      const float* sinc_tab, int32_t n_sinc_samples,
                      int32_t sinc_ovs, osar_c32* out, int32_t nbins,
                      void* stream);

/* ---- Bragg facet kernel ----------------------------------------------- */
typedef struct osar_bragg_desc {
  int32_t nx, ny;
  int32_t pol;              /* 0 = vv, 1 = hh, 2 = DP (hh then vv) */
  int32_t spread;           /* 0 = elfouhaily, 1 = cos2s */
  int32_t n_spec;           /* spectrum table length */
  int32_t pulse;            /* refresh pulse (sets the azimuth of every row) when
                               n_refresh == 0 */
  double k0, d, wind_dir, wind_U, wind_fetch;
  double dx, dy;            /* facet size (RCS = NRCS * dx * dy) */
  double t_step, t_span, v_ground, gr0;
  int32_t spec_uniform;     /* 1: spec_k[i] == spec_k0 + i * spec_dk exactly (product
                               rounded first, as numpy.arange fills), so the search
                               computes grid points instead of loading them */
  int32_t fp64;             /* 0: the facet model in float32, amplitudes to amp;
                               1: every value FP64, FP64 amplitudes to amp64 (the
                               reference-precision path) */
  double spec_k0, spec_dk;  /* spec_k[0] and the grid step (spec_uniform) */
  double spec_inv_dk;       /* 1 / mean step of spec_k: starting guess of the search */
  const double* cos_inc;    /* [nx] cos and sin of the incidence angle */
  const double* sin_inc;
  const double* y_row;      /* [ny] */
  const double* spec_k;     /* [n_spec] ascending */
  const double* spec_S;     /* [n_spec] */
  const double* spec_slope; /* [n_spec] (spec_S[i] - spec_S[i-1]) / (spec_k[i] -
                               spec_k[i-1]) for i >= 1, or NULL (computed per
                               facet; the same value) */
  const float* diffx;       /* [nx][ny] (x-major) */
  const float* diffy;
  float* amp;               /* out: [n_pol][2][nx][ny] sqrt(RCS) per refresh */
  double* rcs_sum;          /* out: [n_refresh][n_pol] sum of RCS over facets and
                               components (accumulated; zero it first); may be NULL */
  /* Several refreshes in one launch (n_refresh > 0): refresh i uses pulse
     refresh_pulse[i] and the slopes at diffx/diffy + (refresh_snap[i] - snap0) *
     diff_stride,
     and writes amp + i * amp_stride and rcs_sum + i * n_pol. n_refresh == 0 is
     one refresh at `pulse` with the slopes at diffx/diffy. */
  int32_t n_refresh;
  int32_t snap0;            /* snapshot index held in slot 0 of diffx/diffy */
  const int32_t* refresh_pulse;  /* device [n_refresh] */
  const int32_t* refresh_snap;   /* device [n_refresh] */
  int64_t diff_stride;      /* elements between snapshot slots of diffx/diffy */
  int64_t amp_stride;       /* elements between amplitude slots */
  double* amp64;            /* out (fp64 = 1): as amp, float64 */
  const float* hmtf;        /* NULL, or the hydrodynamic MTF field of the slopes'
                               snapshot (same layout and diff_stride as diffx): every
                               RCS is multiplied by 1 + max(hmtf, -1), computed in
                               float32 (use_hmtf, sar_raw_nompi.py:448-456) */
} osar_bragg_desc;

int osar_facet_bragg(osar_ctx* ctx, const osar_bragg_desc* d, void* stream);
int osar_fill_zero_f64(osar_ctx* ctx, double* p, int64_t n, void* stream);

/* ---- surface synthesis ------------------------------------------------ */
/* Output field bits of osar_surface_desc.fields */
#define OSAR_F_DX 1
#define OSAR_F_DY 2
#define OSAR_F_DZ 4
#define OSAR_F_DIFFX 8
#define OSAR_F_DIFFY 16
#define OSAR_F_DIFFXX 32
#define OSAR_F_DIFFYY 64
#define OSAR_F_DIFFXY 128
#define OSAR_F_VX 256           /* orbital velocities (ocean.py:484-494) */
#define OSAR_F_VY 512
#define OSAR_F_VZ 1024
#define OSAR_F_AX 2048          /* accelerations (ocean.py:496-506) */
#define OSAR_F_AY 4096
#define OSAR_F_AZ 8192
#define OSAR_F_HMTF 16384       /* hydrodynamic modulation (ocean.py:508-526) */
#define OSAR_NUM_FIELDS 15

typedef struct osar_surface_desc {
  int32_t nx, ny;           /* powers of two, 16 .. 2048 */
  int32_t fields;           /* OSAR_F_* mask of outputs to write */
  int32_t choppy;           /* 1: choppy slope / curvature corrections */
  int32_t n_times;          /* snapshots computed by this call */
  int32_t mirror_grid;      /* 1: kx, ky are a 2 pi fftfreq grid (kx[0] = 0 and
                               kx[nx - i] = -kx[i] for i != nx / 2, the same for
                               ky): selects the faster row pass; 0 gives the
                               same results for any tables */
  int32_t slope_fp32;       /* 1: Diffx and Diffy share one spectrum whose row-pass
                               result is stored as complex64 (every other field, and
                               both slopes when choppy, round-trip in complex128):
                               5 fields take 40 instead of 48 workspace bytes per grid
                               point; the slopes then carry the float32 rounding of
                               the intermediate (about 1e-7 of their RMS) */
  int32_t reserved;
  double depth;             /* water depth, <= 0 for deep water */
  double current_x, current_y;
  double kx_res, ky_res;    /* float32 grid steps (good-k threshold) */
  const float* kx;          /* [nx] float32 wave numbers (2 pi fftfreq) */
  const float* ky;          /* [ny] */
  const osar_c32* coefs;    /* [ny][nx] complex64 wave coefficients */
  const double* times;      /* [n_times] snapshot times */
  /* outputs: float32 [n_times][nx][ny] (x-major: field[x][y]) for every
     requested field, NULL otherwise */
  float* out[OSAR_NUM_FIELDS]; /* indexed by bit position: dx, dy, dz, diffx, diffy,
                               diffxx, diffyy, diffxy, vx, vy, vz, ax, ay, az, hmtf */
} osar_surface_desc;

/* Workspace bytes osar_surface_synth needs for this descriptor. */
int64_t osar_surface_workspace_bytes(const osar_surface_desc* d);
int osar_surface_synth(osar_ctx* ctx, const osar_surface_desc* d, void* stream);

/* osar_surface_synth without any device allocation. No counterpart in the
   reference: for callers that cannot allocate device memory (code under HIP
   graph capture, tight serving loops, harnesses that forbid hipMalloc). The
   same kernels and results as osar_surface_synth, with the synthesis workspace
   taken from a scratch array of OSAR_SYNTH_SCRATCH_BYTES that the library holds
   on every device it runs on until the process ends, instead of the context's
   workspace. A descriptor whose workspace (osar_surface_workspace_bytes) exceeds
   it is rejected with OSAR_EINVAL (split its snapshots over several calls). All
   calls on one device share the scratch: calls that overlap on one device race
   (serialise them on one stream). */
#define OSAR_SYNTH_SCRATCH_BYTES (256ll << 20)
int osar_surface_synth_noalloc(osar_ctx* ctx, const osar_surface_desc* d, void* stream);

/* ---- surface initialisation on the device ---------------------------- */
/* OceanSurface.init's spectrum and coefficients (oceansar/surfaces/ocean.py:
   131-216) on the device, so that a served job's surface never leaves HBM.
   Per grid point (row j, column i; row-major [ny][nx]): |k|, the good-k mask
   |k| > min(kx_res, ky_res) / 2 and 1/|k| in float32 as the reference's
   tables; theta = angle(exp(i (atan2(ky, kx) - wind_dir_eff))); the Elfouhaily
   spectrum S(k) (oceansar/spec/elfouhaily.py:43-99) and the spreading G(k,
   theta) (oceansar/spread/elfouhaily.py:23-56 or cos2s.py:13-18) in FP64,
   each rounded to float32; dirspec = kinv S G and coefs = nx ny sqrt(2 dirspec
   kx_res ky_res) cg in float32 as the reference's expressions, plus the
   directional swell term (oceansar/swell_spec/dir_swell_spec.py:24-39,
   ocean.py:206-216) when swell is set. The random draws replace the
   reference's unseeded numpy RNG with a counter-based generator:
   Philox4x32-10 (Salmon et al., SC'11) with key (seed & 0xffffffff, seed >>
   32) and counter (e & 0xffffffff, e >> 32, 0, 0) at point e = j nx + i gives
   x0..x3; cg = sqrt(-2 ln u0) exp(2 pi i u1) / sqrt(2) with u = (x + 0.5)
   2^-32 from (x0, x1) for the wind sea and (x2, x3) for the swell (FP64,
   rounded to complex64). Scalars that depend only on the wind and the swell
   parameters are the caller's, computed as the reference computes them. */
typedef struct osar_init_desc {
  int32_t nx, ny;           /* grid, 1 .. 2^30 points */
  int32_t spread;           /* 0 Elfouhaily, 1 cos^2 */
  int32_t swell;            /* 1: add the directional swell coefficients */
  uint64_t seed;            /* Philox key */
  float kx_res, ky_res;     /* float32 grid steps */
  float wind_dir_eff;       /* float32 effective wind direction [rad] */
  float reserved;
  /* Elfouhaily spectrum: k_p, c_p, Omega = U / c_p, alpha_p, sigma, gamma,
     alpha_m, c_m, k_m (elfouhaily.py:43-99) */
  double k_p, c_p, omega_w, alpha_p, sigma, gamma, alpha_m, c_m, k_m;
  /* Elfouhaily spreading: a_0, a_p, a_m, cK_p, cK_m (spread/elfouhaily.py:26-52) */
  double a_0, a_p, a_m, cK_p, cK_m;
  /* directional swell: direction [rad], freq_r [Hz], sigf [Hz], sigs [rad],
     amplitude (Hs / 4)^2 / sigf / sqrt(2 pi) */
  double swell_dir, freq_r, sigf, sigs, swell_amp;
  const float* kx;          /* [nx] float32 wave numbers */
  const float* ky;          /* [ny] */
  osar_c32* coefs;          /* out: [ny][nx] complex64 wave coefficients */
  float* dirspec;           /* out (NULL: not written): [ny][nx] wind-sea kinv S G */
  double* rmss_sums;        /* out (NULL: not written): [2] sums of dirspec kx^2
                               kx_res ky_res and dirspec ky^2 kx_res ky_res */
} osar_init_desc;

int osar_surface_init(osar_ctx* ctx, const osar_init_desc* d, void* stream);

/* ---- SAR focusing ----------------------------------------------------- */
/* One channel of sar_focus (oceansar/radarsim/sar_processor.py:131-225):
   the raw planes zero-padded to n_az x n_rg, FFT along range then azimuth,
   times exp(4 pi i rcmc[a] fr[r] / c), inverse FFT along range, times h[a],
   inverse FFT along azimuth, lines [az_lo, az_hi) and range samples [0, n_rg0)
   kept. FP64 (complex128) throughout, like the reference's numpy arrays. The
   1-D tables (fr = fftfreq(n_rg, 1 / rg_sampling), rcmc = sr0 / sqrt(1 - (fa
   l0 / 2 v)^2) - sr0, h = exp(i ph_ac) weighting / beam) are the host's, as
   the reference computes them. */
typedef struct osar_focus_desc {
  int32_t n_pol;            /* polarisation planes of the channel */
  int32_t n_az0, n_rg0;     /* raw lines and samples per line */
  int32_t n_az, n_rg;       /* padded sizes: powers of two in [16, 8192] */
  int32_t az_lo, az_hi;     /* kept lines of the image, 0 <= az_lo < az_hi <= n_az0 */
  int32_t reserved;
  int64_t raw_pol_stride;   /* complex64 elements between polarisation planes of raw */
  const osar_c32* raw;      /* [n_pol][n_az0][n_rg0] (planes raw_pol_stride apart) */
  const double* fr;         /* [n_rg] range frequencies */
  const double* rcmc;       /* [n_az] range migration in the Doppler domain */
  const double* h;          /* [n_az] complex128 azimuth filter, interleaved re, im */
  double* work;             /* [n_pol][n_az][n_rg] complex128 workspace (caller's) */
  double* slc;              /* out: [n_pol][az_hi - az_lo][n_rg0] complex128 */
} osar_focus_desc;

int osar_sar_focus(osar_ctx* ctx, const osar_focus_desc* d, void* stream);

/* ---- InSAR L1b processing ---------------------------------------------- */
/* Sub-sample azimuth shift and flattening of one channel's SLC planes
   (oceansar/radarsim/insar_processor.py:165-185): per range sample r and
   plane, y = IFFT(FFT(x) S) along azimuth (lines of any length n_az), where
   S is the reference's shift phasor. The host gives the spectrum kernel =
   FFT_M(g) / M of the wrapped circular kernel g (g[j] = h[j], g[M - j] =
   h[n_az - j], h = IFFT(S)), M = n_fft a power of two >= 2 n_az - 1 and
   <= 16384; the product is one linear convolution per column in LDS. The
   shifted values are rounded to complex64 (the reference stores them into the
   complex64 array it read from the SLC file) and, when flat is given,
   multiplied by flat[r] = exp(i k0 b_xti (inc(r) - inc0)), in FP64. */
typedef struct osar_coreg_desc {
  int32_t n_pol;            /* planes of the channel */
  int32_t n_az, n_rg;       /* SLC lines and range samples */
  int32_t n_fft;            /* M */
  int64_t in_pol_stride;    /* complex64 elements between input planes */
  const osar_c32* in;       /* [n_pol][n_az][n_rg] SLC as read from the file */
  const double* kernel;     /* [M] complex128, interleaved re, im */
  const double* flat;       /* [n_rg] complex128 flat-earth phasor, or NULL */
  double* out;              /* out: [n_pol][n_az][n_rg] complex128 */
} osar_coreg_desc;

int osar_insar_coregister(osar_ctx* ctx, const osar_coreg_desc* d, void* stream);

/* Multilook of the coregistered images (insar_processor.py:192-214 with
   oceansar/utils/misc.py:108-197): for complex images the intensities
   |s_c|^2 and the interferograms s_c2 conj(s_c1) of every pair c1 < c2 (the
   reference's loop order), each smoothed along range then azimuth, cropped to
   [az_lo, az_hi) x [rg_lo, rg_hi), and the coherence t / sqrt(I_c1 I_c2) as
   magnitude and phase; real images (is_complex = 0) are smoothed as they are
   into `intensity`. A flat window (rg_flat / az_flat = 1) is the reference's
   edge-normalised moving average over offsets -(L/2) .. L - L/2 - 1, summed
   in that order in FP64; otherwise the L normalised taps (numpy's window /
   its sum) with zeros beyond the edges, as fftconvolve 'same'. */
#define OSAR_ML_MAX_IMG 16
typedef struct osar_multilook_desc {
  int32_t n_img;            /* images (channels x polarisations), 1 .. 16 */
  int32_t n_az, n_rg;       /* image lines and samples */
  int32_t is_complex;       /* 1: complex128 images, 0: float64 images */
  int32_t rg_len, az_len;   /* window lengths */
  int32_t rg_flat, az_flat; /* 1: flat window, 0: taps */
  int32_t az_lo, az_hi, rg_lo, rg_hi;   /* crop of the outputs */
  int64_t img_stride;       /* elements between images */
  const void* img;          /* [n_img][n_az][n_rg] */
  const double* rg_taps;    /* [rg_len] when !rg_flat */
  const double* az_taps;    /* [az_len] when !az_flat */
  double* work;             /* osar_multilook_work_bytes() of device memory */
  double* intensity;        /* out: [n_img][az_hi - az_lo][rg_hi - rg_lo] */
  double* coherence;        /* out: [pairs][..] |coherence| (complex images) */
  double* phase;            /* out: [pairs][..] its phase in (-pi, pi] */
} osar_multilook_desc;

int64_t osar_multilook_work_bytes(const osar_multilook_desc* d);
int osar_insar_multilook(osar_ctx* ctx, const osar_multilook_desc* d, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* OCEANSAR_HIP_H */
