// Context, error reporting and fill kernels of liboceansar_hip.so.
#include <stdarg.h>
#include <stdlib.h>

#include <map>
#include <mutex>
#include <utility>

#include "common.h"

static thread_local char g_last_error[1024] = "";

void osar_set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
}

extern "C" const char* osar_last_error(void) { return g_last_error; }

extern "C" int osar_version(void) { return 100; }

int osar_reserve_lds(const void* kernel, size_t bytes) {
  static std::mutex mu;
  static std::map<std::pair<int, const void*>, size_t> done;
  int dev = 0;
  OSAR_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  size_t& have = done[{dev, kernel}];
  if (bytes > have) {
    OSAR_HIP(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    have = bytes;
  }
  return OSAR_OK;
}

extern "C" int64_t osar_struct_size(int which) {
  switch (which) {
    case 0: return (int64_t)sizeof(osar_echo_desc);
    case 1: return (int64_t)sizeof(osar_bragg_desc);
    case 2: return (int64_t)sizeof(osar_surface_desc);
    case 3: return (int64_t)sizeof(osar_focus_desc);
    case 4: return (int64_t)sizeof(osar_init_desc);
    case 5: return (int64_t)sizeof(osar_coreg_desc);
    case 6: return (int64_t)sizeof(osar_multilook_desc);
    default: return -1;
  }
}

extern "C" int osar_ctx_create(int device, osar_ctx** out) {
  OSAR_PTR(out);
  int n = 0;
  OSAR_HIP(hipGetDeviceCount(&n));
  OSAR_REQUIRE(device >= 0 && device < n, OSAR_EINVAL, "device %d out of range (%d devices)",
               device, n);
  OSAR_HIP(hipSetDevice(device));
  {
    const int rc = osar_synth_device_setup(device);
    if (rc) return rc;
  }
  hipDeviceProp_t prop;
  OSAR_HIP(hipGetDeviceProperties(&prop, device));
  osar_ctx* c = (osar_ctx*)calloc(1, sizeof(osar_ctx));
  OSAR_REQUIRE(c != nullptr, OSAR_ENOMEM, "host allocation of the context failed");
  c->device = device;
  c->n_cu = prop.multiProcessorCount;
  c->max_lds = prop.sharedMemPerBlock;
  if (prop.maxSharedMemoryPerMultiProcessor > c->max_lds)
    c->max_lds = prop.maxSharedMemoryPerMultiProcessor;
  if (c->max_lds > 160 * 1024) c->max_lds = 160 * 1024;
  *out = c;
  return OSAR_OK;
}

extern "C" int osar_ctx_destroy(osar_ctx* ctx) {
  if (!ctx) return OSAR_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->work) (void)hipFree(ctx->work);
  free(ctx);
  return OSAR_OK;
}

__global__ void fill_zero_u32(uint32_t* p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = 0u;
}

static int fill_zero_words(osar_ctx* ctx, void* p, int64_t words, hipStream_t s) {
  if (words == 0) return OSAR_OK;
  int64_t blocks = (words + 255) / 256;
  if (blocks > 8 * 1024) blocks = 8 * 1024;
  hipLaunchKernelGGL(fill_zero_u32, dim3((unsigned)blocks), dim3(256), 0, s,
                     (uint32_t*)p, words);
  OSAR_LAUNCHED("fill_zero_u32");
  return OSAR_OK;
}

extern "C" int osar_fill_zero_c32(osar_ctx* ctx, osar_c32* p, int64_t n, void* stream) {
  OSAR_PTR(ctx);
  OSAR_PTR(p);
  OSAR_REQUIRE(n >= 0, OSAR_EINVAL, "negative length");
  return fill_zero_words(ctx, p, 2 * n, (hipStream_t)stream);
}

extern "C" int osar_fill_zero_f64(osar_ctx* ctx, double* p, int64_t n, void* stream) {
  OSAR_PTR(ctx);
  OSAR_PTR(p);
  OSAR_REQUIRE(n >= 0, OSAR_EINVAL, "negative length");
  return fill_zero_words(ctx, p, 2 * n, (hipStream_t)stream);
}
