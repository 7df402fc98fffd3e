// Surface initialisation on the device (osar_surface_init): the wave spectrum
// and the random coefficients of OceanSurface.init (oceansar/surfaces/
// ocean.py:131-216), one thread per grid point.
//
// The float32 structure of the reference's tables is kept where it decides
// bits: |k| = sqrt(kx^2 + ky^2), the good-k mask and 1/|k| with every
// operation rounded to float32, dirspec = (kinv * S) * G and the amplitude
// nx ny sqrt(((2 dirspec) kx_res) ky_res) as float32 products. The spectrum
// S(k), the spreading G and the swell spectrum are evaluated in FP64 and
// rounded to float32 (numpy evaluates them in float32 with its own
// transcendental routines: a few float32 ulp apart).
// The draws come from Philox4x32-10 (J. Salmon, M. Moraes, R. Dror, D. Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11): ten rounds of two
// 32x32->64 multiplies, with the Weyl key schedule between rounds.
#include "common.h"

namespace {

constexpr double kG = 9.80665;          // scipy.constants.g
constexpr double kRho = 1000., kSurfT = 0.072;
constexpr double kPi = 3.14159265358979323846;

struct u32x4 {
  uint32_t x0, x1, x2, x3;
};

__device__ __forceinline__ u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
  constexpr uint32_t kM0 = 0xD2511F53u, kM1 = 0xCD9E8D57u;
  constexpr uint32_t kW0 = 0x9E3779B9u, kW1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t hi0 = __umulhi(kM0, c.x0), lo0 = kM0 * c.x0;
    const uint32_t hi1 = __umulhi(kM1, c.x2), lo1 = kM1 * c.x2;
    c = u32x4{hi1 ^ c.x1 ^ k0, lo1, hi0 ^ c.x3 ^ k1, lo0};
    k0 += kW0;
    k1 += kW1;
  }
  return c;
}

// (N(0,1) + i N(0,1)) / sqrt(2) from two 32-bit words (Box-Muller, FP64)
__device__ __forceinline__ osar_c32 complex_gaussian(uint32_t a, uint32_t b) {
  constexpr double kScale = 1. / 4294967296.;
  const double u0 = ((double)a + 0.5) * kScale, u1 = ((double)b + 0.5) * kScale;
  const double r = sqrt(-2. * log(u0)) * 0.70710678118654752440;
  double s, c;
  sincospi(2. * u1, &s, &c);
  return osar_c32{(float)(r * c), (float)(r * s)};
}

// angle(exp(i x)) in FP64: x wrapped to (-pi, pi]
__device__ __forceinline__ double wrap_angle(double x) { return atan2(sin(x), cos(x)); }

// np.angle(np.exp(1j * x)) of a float32 x as numpy evaluates it: exp of a
// complex64 (cos and sin of x rounded to float32), then the float32 atan2 of
// those, each taken as the correctly rounded float32 of the FP64 value
__device__ __forceinline__ float angle_exp32(float x) {
  const float c = (float)cos((double)x), s = (float)sin((double)x);
  return (float)atan2((double)s, (double)c);
}

// Elfouhaily omnidirectional spectrum (B_l + B_h) / k^3 (elfouhaily.py:43-99)
__device__ double elfouhaily_spec(const osar_init_desc& d, double k) {
  const double c = sqrt(kG / k + kSurfT / kRho * k);
  const double q = sqrt(k / d.k_p) - 1.;
  const double Gamma = exp(-(q * q) / (2. * d.sigma * d.sigma));
  const double J_p = pow(d.gamma, Gamma);
  const double r = d.k_p / k;
  const double L_pm = exp(-5. / 4. * (r * r));
  const double F_p = L_pm * J_p * exp(-d.omega_w / sqrt(10.) * q);
  const double B_l = 0.5 * d.alpha_p * d.c_p / c * F_p;
  const double m = k / d.k_m - 1.;
  const double F_m = L_pm * exp(-0.25 * (m * m));
  const double B_h = 0.5 * d.alpha_m * d.c_m / c * F_m;
  return (B_l + B_h) / (k * k * k);
}

// Spreading G(k, theta): Elfouhaily (spread/elfouhaily.py:23-56) or cos^2
// (spread/cos2s.py:13-18), zero outside the upwind half plane
__device__ double spreading(const osar_init_desc& d, double k, double th) {
  if (d.spread == 1) {
    if (th < -kPi / 2. || th > kPi / 2.) return 0.;
    const double c = cos(th);
    return 1. / kPi * (c * c);
  }
  if (!(th > -kPi / 2. && th < kPi / 2.)) return 0.;
  const double cK = sqrt(kG / k + kSurfT / kRho * k);
  const double delta =
      tanh(d.a_0 + d.a_p * pow(cK / d.cK_p, 2.5) + d.a_m * pow(d.cK_m / cK, 2.5));
  return (1. + delta * cos(2. * th)) / kPi;
}

// Directional swell spectrum (dir_swell_spec.py:24-39)
__device__ double swell_spec(const osar_init_desc& d, double k, double th) {
  const double f_k = 1. / 2. / kPi * sqrt(kG * k);
  const double fac_f_k = 1. / 4. / kPi * sqrt(kG / k);
  const double df = f_k - d.freq_r;
  const double efs = (d.swell_amp * exp(-(df * df) / (2. * d.sigf * d.sigf)) + 1e-5) * fac_f_k;
  const double ang = wrap_angle(th - d.swell_dir);
  const double dirss = exp(-(ang * ang) / (2. * d.sigs * d.sigs));
  const double factor = sqrt(2. * kPi) * d.sigs;
  return efs * dirss / factor;
}

__global__ __launch_bounds__(256) void surface_init_kernel(osar_init_desc d) {
  const int64_t n = (int64_t)d.nx * d.ny;
  const float thr = fminf(d.kx_res, d.ky_res) / 2.f;
  const float nn = (float)n;    // nx ny as float32 (the reference's int * float32 array)
  const uint32_t key0 = (uint32_t)d.seed, key1 = (uint32_t)(d.seed >> 32);
  double sx = 0., sy = 0.;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int j = (int)(e / d.nx), i = (int)(e - (int64_t)j * d.nx);
    const float kx = d.kx[i], ky = d.ky[j];
    const float k32 = __fsqrt_rn(__fadd_rn(__fmul_rn(kx, kx), __fmul_rn(ky, ky)));
    const bool good = k32 > thr;
    const float kinv = good ? __fdiv_rn(1.f, k32) : 0.f;
    const u32x4 r = philox4x32_10(u32x4{(uint32_t)e, (uint32_t)(e >> 32), 0u, 0u}, key0, key1);
    float ds = 0.f;
    osar_c32 coef{0.f, 0.f};
    const double atn = atan2((double)ky, (double)kx);
    if (good) {
      const double k = (double)k32;
      // theta = angle(exp(i (atan2(ky, kx) - wind_dir_eff))) (float32 operands)
      const float th32 = angle_exp32(__fsub_rn((float)atn, d.wind_dir_eff));
      const float S = (float)elfouhaily_spec(d, k);
      const float G = (float)spreading(d, k, (double)th32);
      ds = __fmul_rn(__fmul_rn(kinv, S), G);
    }
    {
      const float amp = __fmul_rn(nn, __fsqrt_rn(__fmul_rn(__fmul_rn(__fmul_rn(2.f, ds), d.kx_res),
                                                           d.ky_res)));
      const osar_c32 cg = complex_gaussian(r.x0, r.x1);
      coef = osar_c32{__fmul_rn(amp, cg.re), __fmul_rn(amp, cg.im)};
    }
    if (d.swell) {
      float ss = 0.f;
      if (good) {
        const float ths = angle_exp32((float)atn);
        ss = __fmul_rn(kinv, (float)swell_spec(d, (double)k32, (double)ths));
      }
      const float amp = __fmul_rn(nn, __fsqrt_rn(__fmul_rn(__fmul_rn(__fmul_rn(2.f, ss), d.kx_res),
                                                           d.ky_res)));
      const osar_c32 cg = complex_gaussian(r.x2, r.x3);
      coef = osar_c32{__fadd_rn(coef.re, __fmul_rn(amp, cg.re)),
                      __fadd_rn(coef.im, __fmul_rn(amp, cg.im))};
    }
    d.coefs[e] = coef;
    if (d.dirspec) d.dirspec[e] = ds;
    // rmss sums: dirspec kx^2 kx_res ky_res as float32 products (ocean.py:181-182)
    sx += (double)__fmul_rn(__fmul_rn(__fmul_rn(ds, __fmul_rn(kx, kx)), d.kx_res), d.ky_res);
    sy += (double)__fmul_rn(__fmul_rn(__fmul_rn(ds, __fmul_rn(ky, ky)), d.kx_res), d.ky_res);
  }
  if (!d.rmss_sums) return;
  __shared__ double red[2][4];
  for (int off = 32; off > 0; off >>= 1) {
    sx += __shfl_xor(sx, off);
    sy += __shfl_xor(sy, off);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = sx;
    red[1][threadIdx.x >> 6] = sy;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(&d.rmss_sums[0], red[0][0] + red[0][1] + red[0][2] + red[0][3]);
    atomicAdd(&d.rmss_sums[1], red[1][0] + red[1][1] + red[1][2] + red[1][3]);
  }
}

__global__ void zero2_kernel(double* p) {
  if (threadIdx.x < 2) p[threadIdx.x] = 0.;
}

}  // namespace

extern "C" int osar_surface_init(osar_ctx* ctx, const osar_init_desc* d, void* stream) {
  OSAR_PTR(ctx);
  OSAR_PTR(d);
  OSAR_PTR(d->kx);
  OSAR_PTR(d->ky);
  OSAR_PTR(d->coefs);
  OSAR_REQUIRE(d->nx >= 1 && d->ny >= 1 && (int64_t)d->nx * d->ny <= (1ll << 30), OSAR_EINVAL,
               "surface %dx%d: 1 .. 2^30 grid points", d->nx, d->ny);
  OSAR_REQUIRE(d->spread == 0 || d->spread == 1, OSAR_EUNSUPPORTED, "spreading model %d",
               d->spread);
  OSAR_REQUIRE(d->kx_res > 0.f && d->ky_res > 0.f, OSAR_EINVAL, "grid steps must be positive");
  OSAR_HIP(hipSetDevice(ctx->device));
  hipStream_t s = (hipStream_t)stream;
  if (d->rmss_sums) {
    hipLaunchKernelGGL(zero2_kernel, dim3(1), dim3(64), 0, s, d->rmss_sums);
    OSAR_LAUNCHED("zero2_kernel");
  }
  const int64_t n = (int64_t)d->nx * d->ny;
  const int64_t want = (n + 255) / 256;
  const int blocks = (int)(want < 16384 ? want : 16384);
  hipLaunchKernelGGL(surface_init_kernel, dim3(blocks), dim3(256), 0, s, *d);
  OSAR_LAUNCHED("surface_init_kernel");
  return OSAR_OK;
}
