// Mixed-precision (int8 MFMA) genotype-stream kernels -- placeholder until
// the kernels land; the FP64 path is complete without them.
#include "internal.h"

namespace slfmm {
struct EpiArgs;
void launch_func_i8(slfmm_ctx*, const slfmm_geno*, const double*, uint64_t, const EpiArgs&, bool) {
  throw Error(SLFMM_E_ARG, "SLFMM_PREC_I8 association kernel is not available in this build");
}
void rsvd_project_i8(slfmm_ctx*, const uint8_t*, uint64_t, uint64_t, uint64_t, const double*, uint64_t,
                     const int32_t*, const int32_t*, const double*, const double*, double*, uint64_t) {
  throw Error(SLFMM_E_ARG, "SLFMM_PREC_I8 RSVD kernel is not available in this build");
}
void rsvd_accumulate_i8(slfmm_ctx*, const uint8_t*, uint64_t, uint64_t, uint64_t, const double*, uint64_t,
                        uint64_t, const double*, const double*, double*, uint64_t) {
  throw Error(SLFMM_E_ARG, "SLFMM_PREC_I8 RSVD kernel is not available in this build");
}
}  // namespace slfmm
