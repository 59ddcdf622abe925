// Specialised register-resident passes for the hot sizes.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/pfft.h"
#include "fast_pass.hpp"

namespace pfft {

void fast_select(const Plan& p, PassPlan& pp) {
  (void)p;
  (void)pp;
}

int fast_prepare(PassPlan& pp) {
  (void)pp;
  return PFFT_OK;
}

int fast_launch(const PassPlan& pp, const PassArgs& a, hipStream_t stream) {
  (void)a;
  (void)stream;
  return fail(PFFT_ERR_UNSUPPORTED, "no specialised kernel variant " + std::to_string(pp.variant));
}

}  // namespace pfft
