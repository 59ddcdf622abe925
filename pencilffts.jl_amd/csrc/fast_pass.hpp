// Specialised register-resident passes for the hot sizes (declarations).
#pragma once

#include "plan.hpp"

namespace pfft {

// Variant id bit: the pass walks its tiles in descending order.
constexpr int kVariantReverse = 1 << 21;
// Variant id bit: the pass stores its output with the nontemporal hint.
constexpr int kVariantStream = 1 << 22;

// Specialised kernels able to run a pass, as variant ids (first = default).
// op: variants that carry a spectral operator (pfft_execute_op)
std::vector<int> fast_candidates(const Plan& p, const PassPlan& pp, bool op = false);
// Switches a pass to a specialised variant (tile shape, grid, LDS).
void fast_apply(PassPlan& pp, int variant);
int fast_prepare(const PassPlan& pp);
int fast_launch(const PassPlan& pp, const PassArgs& a, hipStream_t stream);
const char* fast_name(int variant);  // kernel name of a variant ("generic" for 0)
// Digest of this build's kernel table (names of every entry, in order): tune
// data names kernels by table index, so it is valid for one table only.
std::string fast_build_signature();

}  // namespace pfft
