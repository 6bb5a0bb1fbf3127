// FP8 e4m3 block GEMM with per-block power-of-two scales (placeholder until
// the gfx950 f8f6f4 MFMA path lands): reports BBTS_E_UNSUPPORTED.
#include "common.h"

namespace bbts_hip {
int launch_gemm_fp8(bbts_hip_ctx *ctx, const gemm_args &g, int refine, hipStream_t s) {
  (void)ctx; (void)g; (void)refine; (void)s;
  set_error("FP8 GEMM is not built yet");
  return BBTS_E_UNSUPPORTED;
}
}  // namespace bbts_hip
