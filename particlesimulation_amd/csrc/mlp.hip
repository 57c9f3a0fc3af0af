// Learned influence-function correction (filled in by the MLP kernels).
#include "pmgpu_internal.h"

namespace pmgpu {

float* mlp_correction_grid(pmgpu_ctx* c) {
  (void)c;
  fail(PMGPU_ERR_UNSUPPORTED, "learned correction not available");
}

}  // namespace pmgpu

extern "C" {

int64_t pmgpu_mlp_num_params(void) { return 4 * 64 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 64 + 1; }

int pmgpu_set_greens_correction(pmgpu_ctx* c, const float* weights) {
  if (!c) return PMGPU_ERR_INVALID_ARGUMENT;
  if (!weights) {
    c->mlp_weights.clear();
    return PMGPU_OK;
  }
  c->mlp_weights.assign(weights, weights + pmgpu_mlp_num_params());
  return PMGPU_OK;
}

}  // extern "C"
