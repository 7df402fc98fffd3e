// Shared helpers of liboceansar_hip.so (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string>

#include "oceansar_hip.h"

struct osar_ctx {
  int device;
  int n_cu;
  size_t max_lds;             // bytes of LDS one workgroup may allocate
  void* work;                 // synthesis workspace (grows on demand)
  size_t work_bytes;
};

// Raise a kernel's dynamic LDS limit to `bytes` on the current device if it is
// not already that high. The attribute is only set when the need grows, so
// that steady-state launches (and launches under stream capture, after a
// first eager run) make no other runtime calls.
int osar_reserve_lds(const void* kernel, size_t bytes);

// Per-device set-up of the synthesis (twiddle tables, no-allocation scratch),
// once per device with that device current; called by osar_ctx_create.
int osar_synth_device_setup(int device);

// Device address of the two-level twiddle table of an n-point transform (n a
// power of two in 16 .. 8192; see fft_lds.h) of the context's device.
int osar_twiddles(osar_ctx* ctx, int n, const double2** out);

// Thread-local last error message (osar_last_error).
void osar_set_error(const char* fmt, ...);

#define OSAR_REQUIRE(cond, code, ...)      \
  do {                                     \
    if (!(cond)) {                         \
      osar_set_error(__VA_ARGS__);         \
      return (code);                       \
    }                                      \
  } while (0)

#define OSAR_HIP(call)                                                   \
  do {                                                                   \
    hipError_t e_ = (call);                                              \
    if (e_ != hipSuccess) {                                              \
      osar_set_error("%s failed: %s", #call, hipGetErrorString(e_));     \
      return e_ == hipErrorOutOfMemory ? OSAR_ENOMEM : OSAR_ELAUNCH;     \
    }                                                                    \
  } while (0)

// Check the launch that was just issued.
#define OSAR_LAUNCHED(name)                                                      \
  do {                                                                           \
    hipError_t e_ = hipGetLastError();                                           \
    if (e_ != hipSuccess) {                                                      \
      osar_set_error("launch of %s failed: %s", name, hipGetErrorString(e_));    \
      return OSAR_ELAUNCH;                                                       \
    }                                                                            \
  } while (0)

// Pointer must be set (the ABI takes raw device pointers).
#define OSAR_PTR(p) OSAR_REQUIRE((p) != nullptr, OSAR_EINVAL, "null pointer: %s", #p)

namespace osar {

constexpr double kTwoPi = 6.283185307179586476925286766559;
constexpr double kInvTwoPi = 0.15915494309189533576888376337251;

// exp(i * phase) for an FP64 phase of any size: the phase is reduced to
// revolutions in FP64 (exact subtraction of the nearest integer), and only the
// reduced fraction in [-1/2, 1/2] goes through the hardware sin/cos, which take
// their argument in revolutions (v_sin_f32 / v_cos_f32; max abs error 1.3e-7
// over [-1/2, 1/2], measured against FP64 sin/cos on gfx950).
__device__ __forceinline__ float2 cis_fp64(double phase) {
  const double r = phase * kInvTwoPi;
  const float f = (float)(r - rint(r));
  return make_float2(__builtin_amdgcn_cosf(f), __builtin_amdgcn_sinf(f));
}

// exp(2 pi i r) for an FP64 phase r already in revolutions: its fractional part
// (v_fract_f64, exact) is rounded to float32 and goes through the hardware
// sin/cos, which take revolutions and reduce their argument themselves; half
// the FP64 work of cis_fp64 when the caller can form the phase in revolutions.
__device__ __forceinline__ float2 cis_rev(double r) {
  const float f = (float)__builtin_amdgcn_fract(r);
  return make_float2(__builtin_amdgcn_cosf(f), __builtin_amdgcn_sinf(f));
}

// a / b for b > 0 given inv = 1 / b correctly rounded: one residual correction
// of a * inv gives the correctly rounded quotient (Markstein), i.e. the same
// bits as IEEE division, at three FP64 operations.
__device__ __forceinline__ double div_by(double a, double b, double inv) {
  const double q0 = a * inv;
  const double r = fma(-q0, b, a);
  return fma(r, inv, q0);
}

// Workgroups are dispatched to the 8 XCDs round-robin by linear block index,
// and each XCD has its own L2. xcd_item maps block h of G to a work item so
// that XCD x runs the contiguous item range [x G / 8, (x + 1) G / 8) in order:
// neighbouring items run at the same time behind the same L2.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_item(int h, int G) {
  const int q = G / kXcds, r = G % kXcds;
  const int x = h % kXcds, s = h / kXcds;
  return x * q + min(x, r) + s;
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

}  // namespace osar
