// Surface synthesis (placeholder until the FFT kernels land).
#include "common.h"

extern "C" int64_t osar_surface_workspace_bytes(const osar_surface_desc* d) {
  (void)d;
  return 0;
}

extern "C" int osar_surface_synth(osar_ctx* ctx, const osar_surface_desc* d, void* stream) {
  (void)ctx;
  (void)d;
  (void)stream;
  osar_set_error("osar_surface_synth: not implemented in this build");
  return OSAR_EUNSUPPORTED;
}
