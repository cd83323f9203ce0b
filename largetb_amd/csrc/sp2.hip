// Dense SP2 purification (placeholder until the MFMA GEMM lands).
#include <hip/hip_runtime.h>

#include "ctx.hpp"

extern "C" int ltb_sp2_dense(ltb_ctx* c, double Ne, int max_iter, int fp64_tail, double* diagP,
                             double* pop, ltb_sp2_result* res) {
  if (!c) return LTB_ERR_ARG;
  return set_err(c, LTB_ERR_STATE, "ltb_sp2_dense is not built yet");
}
