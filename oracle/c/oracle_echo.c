/*
 * C restatement of the sar_raw per-pulse body, FP64 throughout.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle): used by tests/ to check the GPU
 * path at sizes the numpy oracle is too slow for, and by bench.py as the
 * timed CPU baseline (a stand-in for the reference's numba nopython loop).
 *
 * Per pulse p (one OpenMP thread per pulse, so output lines are private):
 *   fields      w_l S[m] + w_h S[m+1] in float32, w_h = (p mod K)/K,
 *               w_l = 1 - w_h (both rounded to float32)    balancer.py:274-282
 *   geometry    az, sin_az, sr_surface, wave_dinc            sar_raw_nompi.py:352-369
 *   Bragg scene sum_pm amp * exp(i phase_pm) * s_pm          sar_raw_nompi.py:470-480
 *   beam        sinc_bp(L_tx) * sinc_bp(L_rx)                sar_raw_nompi.py:485-486
 *   channels    exp(-i k0 b_ati sin_az) exp(-i k0 b_xti dinc) sar_raw_nompi.py:488-495
 *   integrator  profile_integrator                           range_profile.py:22-49,106-108
 * in the numpy oracle's operation order; build with -ffp-contract=off.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define MAXCH 8

typedef struct oracle_desc {
  int32_t nx, ny, K, snap0, n_snap, p0, p1, nbins, n_ch, n_pol, rcs0, n_rcs, bragg, n_sinc, ovs,
      pad;
  double t_step, t_span, v_ground, sr0, gr0, k0, min_sr, binsize, ltx_l, lrx_l, inc_angle;
  double b_ati[MAXCH], b_xti[MAXCH];
  const double *sr, *sin_inc, *cos_inc, *inc, *cap_rate, *y;
  const float *dx, *dy, *dz;         /* [n_snap][ny][nx] */
  const double* amp;                 /* [n_rcs][n_pol][2][ny][nx] = sqrt(RCS) */
  const int32_t* rcs_slot;           /* [n_pulses] */
  const double *sm, *sp;             /* series, interleaved re/im, lengths nm, np */
  int64_t nm, np_;
  const int64_t *ind0_m, *ind0_p;    /* [ny][nx] */
  const float* sinc;                 /* [n_sinc*ovs+1] */
  double* out;                       /* [n_pol][n_ch][p1-p0][nbins][2] */
} oracle_desc;

static double np_sinc(double x) {
  double y = M_PI * (x == 0.0 ? 1.0e-20 : x);
  return sin(y) / y;
}

static double lerp(const float* f, size_t o0, size_t plane, int rk, float wl, float wh) {
  const float a = f[o0];
  if (rk == 0) return (double)a;
  const float b = f[o0 + plane];
  const float x = wl * a;
  const float y = wh * b;
  return (double)(x + y);
}

int oracle_echo(const oracle_desc* d) {
  const int npulse = d->p1 - d->p0;
  const size_t plane = (size_t)d->nx * d->ny;
  const double two_k0 = 2 * d->k0;
#pragma omp parallel for schedule(dynamic, 1)
  for (int ip = 0; ip < npulse; ++ip) {
    const int p = d->p0 + ip;
    const int m = p / d->K, rk = p % d->K;
    const float wh = (float)((double)rk / (double)d->K);
    const float wl = (float)(1.0 - (double)wh);
    const size_t so = (size_t)(m - d->snap0) * plane;
    const int rs = d->bragg ? d->rcs_slot[p] - d->rcs0 : 0;
    const double t_now = p * d->t_step;
    const double az_now = (t_now - d->t_span / 2.) * d->v_ground;
    for (int y = 0; y < d->ny; ++y) {
      const double az = d->y[y] - az_now;
      const double sin_az = az / d->sr0;
      const double mig = az / 2 * sin_az;
      const double beam = np_sinc(d->ltx_l * sin_az) * np_sinc(d->lrx_l * sin_az);
      double cre[MAXCH], cim[MAXCH];
      for (int ch = 0; ch < d->n_ch; ++ch) {
        const double ph = -(d->k0 * d->b_ati[ch]) * sin_az;
        cre[ch] = cos(ph);
        cim[ch] = sin(ph);
      }
      for (int x = 0; x < d->nx; ++x) {
        const size_t idx = (size_t)y * d->nx + x;
        const double Dx = lerp(d->dx, so + idx, plane, rk, wl, wh);
        const double Dy = lerp(d->dy, so + idx, plane, rk, wl, wh);
        const double Dz = lerp(d->dz, so + idx, plane, rk, wl, wh);
        double srs = d->sr[x] - d->cos_inc[x] * Dz;
        srs = srs + mig;
        srs = srs + Dx * d->sin_inc[x];
        srs = srs + Dy * sin_az;
        const double dinc = (Dz * d->sin_inc[x] + Dx * d->sin_inc[x]) / d->sr0;
        double sre[2] = {0., 0.}, sim[2] = {0., 0.};
        if (d->bragg) {
          const double surf = -two_k0 * srs;
          const double cap = d->cap_rate[x] * (double)(p + 1);
          const double ph0 = surf - cap, ph1 = surf + cap;
          const double e0r = cos(ph0), e0i = sin(ph0), e1r = cos(ph1), e1i = sin(ph1);
          const int64_t im = (d->ind0_m[idx] + p) % d->nm;
          const int64_t ipp = (d->ind0_p[idx] + p) % d->np_;
          const double smr = d->sm[2 * im], smi = d->sm[2 * im + 1];
          const double spr = d->sp[2 * ipp], spi = d->sp[2 * ipp + 1];
          for (int pol = 0; pol < d->n_pol; ++pol) {
            const size_t ao = ((size_t)(rs * d->n_pol + pol) * 2) * plane + idx;
            const double a0 = d->amp[ao], a1 = d->amp[ao + plane];
            const double t0r = a0 * e0r, t0i = a0 * e0i;   /* amp * exp(i ph) */
            const double t1r = a1 * e1r, t1i = a1 * e1i;
            const double u0r = t0r * smr - t0i * smi, u0i = t0r * smi + t0i * smr;
            const double u1r = t1r * spr - t1i * spi, u1i = t1r * spi + t1i * spr;
            sre[pol] = u0r + u1r;
            sim[pol] = u0i + u1i;
          }
        }
        /* bin position (range_profile.py:106-108) */
        const double xb = (srs - d->min_sr) / d->binsize;
        const double fl = floor(xb);
        const int db = (int)fl;
        const int sb = (int)nearbyint(d->ovs * (1 - (xb - fl)));
        const int b0 = db - d->n_sinc / 2;
        for (int ch = 0; ch < d->n_ch; ++ch) {
          double xr = 1., xi = 0.;
          if (d->b_xti[ch] != 0.) {
            const double tot = (d->inc[x] - d->inc_angle) + dinc;
            const double ph = -(d->k0 * d->b_xti[ch]) * tot;
            xr = cos(ph);
            xi = sin(ph);
          }
          for (int pol = 0; pol < d->n_pol; ++pol) {
            /* (scene * beam) * exp(i ati) * exp(i xti) */
            double vr = sre[pol] * beam, vi = sim[pol] * beam;
            double tr = vr * cre[ch] - vi * cim[ch], ti = vr * cim[ch] + vi * cre[ch];
            vr = tr * xr - ti * xi;
            vi = tr * xi + ti * xr;
            double* line = d->out + (((size_t)pol * d->n_ch + ch) * npulse + ip) * d->nbins * 2;
            for (int j = 0; j < d->n_sinc; ++j) {
              const int b = b0 + j;
              if (b < 0 || b >= d->nbins) continue;
              const double t = (double)d->sinc[sb + j * d->ovs];
              line[2 * b] += t * vr;
              line[2 * b + 1] += t * vi;
            }
          }
        }
      }
    }
  }
  return 0;
}

int64_t oracle_desc_size(void) { return (int64_t)sizeof(oracle_desc); }

/* OpenMP threads of the pulse loop (the environment is read only once). */
int oracle_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
}
