// Specialised register-resident passes for the hot sizes (declarations).
#pragma once

#include "plan.hpp"

namespace pfft {

// Picks a specialised kernel for the pass when one exists (sets pp.variant,
// threads, grid, lds_bytes); leaves the generic pass otherwise.
void fast_select(const Plan& p, PassPlan& pp);
int fast_prepare(PassPlan& pp);
int fast_launch(const PassPlan& pp, const PassArgs& a, hipStream_t stream);
const char* fast_name(int variant);  // kernel name of a variant ("generic" for 0)

}  // namespace pfft
