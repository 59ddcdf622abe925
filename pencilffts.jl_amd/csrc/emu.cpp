// Host emulator of the generic pass, for the CPU test-suite only.
//
// Runs the exact phase functions of generic_pass.hpp that the GPU kernel runs
// (same pass descriptors, same twiddle tables rounded to the plan precision),
// with the workgroup's threads executed one after another between barriers.
// It is built into its own library (libpencilffts_emu.so) that only tests
// load; the product library has no host execution path.
#include <string.h>

#include <string>
#include <vector>

#include "../../include/pfft.h"
#include "generic_pass.hpp"
#include "plan.hpp"

using namespace pfft;

struct pfft_emu_plan {
  Plan p;
  int32_t tile_order = 0, tile_rev = 0;   // tile traversal (tile_split)
};

template <typename R>
static void emu_pass(const PassPlan& pp, const PassArgs& a) {
  std::vector<char> lds((size_t)pp.lds_bytes);  // exact size: overruns show under ASan
  const int nthr = pp.threads;
  for (int64_t tile = 0; tile < a.ntiles; ++tile) {
    TileCtx c;
    tile_init(a, tile, c);
    for (int t = 0; t < nthr; ++t) phase_load<R>(a, c, t, nthr, lds.data());
    for (int t = 0; t < nthr; ++t) phase_pre<R>(a, c, t, nthr, lds.data());
    for (int s = 0; s < fft_steps(a); ++s)
      for (int t = 0; t < nthr; ++t) phase_fft_step<R>(a, c, s, t, nthr, lds.data());
    for (int t = 0; t < nthr; ++t) phase_post<R>(a, c, t, nthr, lds.data());
  }
}

extern "C" {

__attribute__((visibility("default"))) int pfft_emu_plan_create(
    pfft_emu_plan** out, int32_t ndims, const int64_t* dims, const int32_t* kinds,
    int32_t real_type, int32_t permute, int32_t n_extra, const int64_t* extra, uint32_t flags) {
  *out = nullptr;
  pfft_emu_plan* h = new pfft_emu_plan();
  int r = plan_build(h->p, ndims, dims, kinds, real_type, permute, n_extra, extra,
                     flags | PFFT_FLAG_GENERIC_ONLY);
  if (r) {
    delete h;
    return r;
  }
  *out = h;
  return PFFT_OK;
}

__attribute__((visibility("default"))) int pfft_emu_plan_destroy(pfft_emu_plan* h) {
  delete h;
  return PFFT_OK;
}

__attribute__((visibility("default"))) const char* pfft_emu_last_error(void) {
  return last_error();
}

__attribute__((visibility("default"))) int pfft_emu_nbytes(const pfft_emu_plan* h,
                                                          size_t* in_b, size_t* out_b) {
  *in_b = h->p.in_bytes;
  *out_b = h->p.out_bytes;
  return PFFT_OK;
}

// Host buffers in the same layouts as pfft_execute's device buffers; op as
// in pfft_execute_op (null: none).
__attribute__((visibility("default"))) int pfft_emu_execute_op(pfft_emu_plan* h, int32_t dir,
                                                              const pfft_spectral_op* op,
                                                              const void* in, void* out) {
  Plan& p = h->p;
  int op_pass = -1, op_side = 0;
  if (op && op->op) {
    const int r0 = op_setup(p, dir, *op, &op_pass, &op_side);
    if (r0) return r0;
  }
  if (dir < PFFT_FORWARD || dir > PFFT_ADJOINT_BACKWARD)
    return fail(PFFT_ERR_ARGUMENT, "unknown direction");
  if ((dir == PFFT_ADJOINT_FORWARD || dir == PFFT_ADJOINT_BACKWARD) && !p.adjoint_ok)
    return fail(PFFT_ERR_UNSUPPORTED, "no adjoint for R2R plans");
  const bool fwd_shape = dir == PFFT_FORWARD || dir == PFFT_ADJOINT_BACKWARD;
  if (p.inplace && in != out)
    return fail(PFFT_ERR_ARGUMENT, "input and output arrays for in-place plan must be the same");
  // the two work buffers as separate exact-size allocations
  std::vector<char> work0(p.work0_bytes), work1(p.work_bytes - p.work0_bytes);
  const std::vector<PassPlan>& passes = fwd_shape ? p.fwd : p.bwd;
  void* bufs[3] = {nullptr, work0.data(), work1.data()};
  const int np = (int)passes.size();
  for (int i = 0; i < np; ++i) {
    const PassPlan& pp = passes[i];
    PassArgs a = pp.args;
    a.in = pp.src == 0 ? in : bufs[pp.src];
    a.out = pp.dst == 0 ? out : bufs[pp.dst];
    bind_tables(p, pp, a, true);
    if (dir == PFFT_BACKWARD_SCALED && i == np - 1) {
      a.do_scale = 1;
      a.scale = 1.0 / p.scale;
    }
    a.adjoint = (dir == PFFT_ADJOINT_FORWARD || dir == PFFT_ADJOINT_BACKWARD) ? 1 : 0;
    a.tile_order = h->tile_order;
    a.tile_rev = h->tile_rev;
    if (i == op_pass) op_fill(p, dir, *op, op_side, a);
    if (p.real_type == PFFT_F32) emu_pass<float>(pp, a);
    else emu_pass<double>(pp, a);
  }
  return PFFT_OK;
}

// Tile traversal of every pass (the specialised kernels' tuned tile order and
// direction), to check on the host that each order covers every line once.
__attribute__((visibility("default"))) int pfft_emu_set_traversal(pfft_emu_plan* h,
                                                                 int32_t order, int32_t rev) {
  h->tile_order = order ? 1 : 0;
  h->tile_rev = rev ? 1 : 0;
  return PFFT_OK;
}

// Convolution length of the Bluestein form of each forward pass (0: direct
// stages); returns the number of passes.
__attribute__((visibility("default"))) int pfft_emu_bluestein(pfft_emu_plan* h, int64_t* out,
                                                             int32_t n) {
  const int np = (int)h->p.fwd.size();
  for (int i = 0; i < np && i < n; ++i) out[i] = h->p.fwd[i].args.bs_M;
  return np;
}

__attribute__((visibility("default"))) int pfft_emu_execute(pfft_emu_plan* h, int32_t dir,
                                                           const void* in, void* out) {
  return pfft_emu_execute_op(h, dir, nullptr, in, out);
}

}  // extern "C"
