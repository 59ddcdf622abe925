// extern "C" entry points declared in include/pfft.h.  Every function catches
// everything: nothing is thrown across the ABI.
#include <string.h>

#include <new>
#include <string>

#include "../../include/pfft.h"
#include "plan.hpp"

using namespace pfft;

struct pfft_plan {
  Plan p;
};

#define PFFT_GUARD(body)                                              \
  try {                                                               \
    body                                                              \
  } catch (const std::bad_alloc&) {                                   \
    return fail(PFFT_ERR_ALLOC, "host allocation failed");            \
  } catch (...) {                                                     \
    return fail(PFFT_ERR_ARGUMENT, "unexpected exception");           \
  }

#define NEED(ptr)                                                      \
  if (!(ptr)) return fail(PFFT_ERR_ARGUMENT, "null pointer: " #ptr);

extern "C" {

const char* pfft_version(void) { return "pencilffts-hip 0.1.0 (PencilFFTs.jl v0.15.3 API)"; }
const char* pfft_last_error(void) { return last_error(); }

int pfft_transform_check(int32_t kind) {
  std::string why;
  if (!kind_valid(kind, &why)) return fail(PFFT_ERR_ARGUMENT, why);
  return PFFT_OK;
}
int pfft_transform_length_output(int32_t kind, int64_t n_in, int64_t* n_out) {
  NEED(n_out);
  PFFT_GUARD(return length_output(kind, n_in, n_out);)
}
int pfft_transform_binv(int32_t kind, int64_t d, int32_t* kind_inv) {
  NEED(kind_inv);
  PFFT_GUARD(return binv(kind, d, kind_inv);)
}
int pfft_transform_scale_factor(int32_t kind, int64_t n_in, int64_t* scale) {
  NEED(scale);
  PFFT_GUARD(return scale_factor_1d(kind, n_in, scale);)
}
int pfft_transform_eltype_input(int32_t kind, int32_t real_type, int32_t* eltype) {
  NEED(eltype);
  PFFT_GUARD(return eltype_input(kind, real_type, eltype);)
}
int pfft_transform_eltype_output(int32_t kind, int32_t eltype_in, int32_t* eltype) {
  NEED(eltype);
  PFFT_GUARD(return eltype_output(kind, eltype_in, eltype);)
}
// Transforms.jl:139-147: all in-place -> 1, all out-of-place -> 0, mixed -> -1
int pfft_transforms_is_inplace(int32_t n, const int32_t* kinds, int32_t* result) {
  NEED(kinds);
  NEED(result);
  if (n < 1) return fail(PFFT_ERR_ARGUMENT, "empty transform list");
  for (int i = 0; i < n; ++i) {
    std::string why;
    if (!kind_valid(kinds[i], &why)) return fail(PFFT_ERR_ARGUMENT, why);
  }
  const bool a = kind_inplace(kinds[0]);
  for (int i = 1; i < n; ++i) {
    if (kind_inplace(kinds[i]) != a) {
      *result = -1;
      return PFFT_OK;
    }
  }
  *result = a ? 1 : 0;
  return PFFT_OK;
}
// Transforms.jl:285-291 (same transform along every dim) and r2c.jl:114-124
// (RFFT -> (RFFT, FFT, ...); BRFFT -> (BFFT, ..., BFFT, BRFFT)).
int pfft_transform_expand_dims(int32_t kind, int32_t n, int32_t* out) {
  NEED(out);
  std::string why;
  if (!kind_valid(kind, &why)) return fail(PFFT_ERR_ARGUMENT, why);
  if (n < 0) return fail(PFFT_ERR_ARGUMENT, "negative dimension count");
  const int32_t ip = kind & PFFT_INPLACE;
  const int32_t b = kind_base(kind);
  for (int i = 0; i < n; ++i) {
    int32_t k = kind;
    if (b == PFFT_RFFT && i > 0) k = PFFT_FFT | ip;
    if ((b == PFFT_BRFFT_EVEN || b == PFFT_BRFFT_ODD) && i < n - 1) k = PFFT_BFFT | ip;
    out[i] = k;
  }
  return PFFT_OK;
}
int pfft_transform_name(int32_t kind, char* buf, size_t buflen) {
  NEED(buf);
  std::string why;
  if (!kind_valid(kind, &why)) return fail(PFFT_ERR_ARGUMENT, why);
  const std::string s = kind_name(kind);
  if (buflen < s.size() + 1) return fail(PFFT_ERR_ARGUMENT, "buffer too small");
  memcpy(buf, s.c_str(), s.size() + 1);
  return PFFT_OK;
}

int pfft_plan_create(pfft_plan** out, int32_t ndims, const int64_t* dims_logical,
                     const int32_t* kinds, int32_t real_type, int32_t permute_dims,
                     int32_t n_extra, const int64_t* extra_dims, uint32_t flags) {
  NEED(out);
  NEED(dims_logical);
  NEED(kinds);
  if (n_extra > 0 && !extra_dims) return fail(PFFT_ERR_ARGUMENT, "null extra_dims");
  *out = nullptr;
  PFFT_GUARD({
    pfft_plan* h = new pfft_plan();
    int r = plan_build(h->p, ndims, dims_logical, kinds, real_type, permute_dims, n_extra,
                       extra_dims, flags);
    if (r) {
      delete h;
      return r;
    }
    *out = h;
    return PFFT_OK;
  })
}

int pfft_plan_destroy(pfft_plan* plan) {
  if (!plan) return PFFT_OK;
  PFFT_GUARD({
    device_release(plan->p);
    delete plan;
    return PFFT_OK;
  })
}

int pfft_plan_prepare(pfft_plan* plan, void* hip_stream) {
  NEED(plan);
  (void)hip_stream;
  PFFT_GUARD(return device_prepare(plan->p);)
}

int pfft_plan_ndims(const pfft_plan* plan, int32_t* ndims, int32_t* n_extra) {
  NEED(plan);
  if (ndims) *ndims = plan->p.N;
  if (n_extra) *n_extra = plan->p.Ne;
  return PFFT_OK;
}

// which: 0 = input, 1 = output (logical global dims, plus extra dims)
int pfft_plan_size_global(const pfft_plan* plan, int32_t which, int64_t* dims) {
  NEED(plan);
  NEED(dims);
  const Plan& p = plan->p;
  for (int i = 0; i < p.N; ++i) dims[i] = which ? p.size_out[i] : p.size_in[i];
  for (int i = 0; i < p.Ne; ++i) dims[p.N + i] = p.extra[i];
  return PFFT_OK;
}

// Memory-order dims of the input (which = 0) / output (which = 1) array.
int pfft_plan_mem_dims(const pfft_plan* plan, int32_t which, int64_t* dims) {
  NEED(plan);
  NEED(dims);
  const Plan& p = plan->p;
  const Stage& s = which ? p.stages.back() : p.stages.front();
  const int64_t* ld = which ? p.size_out : p.size_in;
  for (int i = 0; i < p.N; ++i) dims[i] = ld[s.perm[i]];
  for (int i = 0; i < p.Ne; ++i) dims[p.N + i] = p.extra[i];
  return PFFT_OK;
}

// Memory dims of the parent array behind the input (which = 0) or output
// (1) view: the memory dims, except that the real input of an in-place RFFT!
// plan lives in a parent padded to 2 (n/2 + 1) along memory axis 0
// (multiarrays_r2c.jl:75-86).
int pfft_plan_parent_dims(const pfft_plan* plan, int32_t which, int64_t* dims) {
  int r = pfft_plan_mem_dims(plan, which, dims);
  if (r) return r;
  if (which == 0 && plan->p.pad_in0) dims[0] = plan->p.pad_in0;
  return PFFT_OK;
}

// 1-based permutation (memory axis i holds logical dim perm[i]).
int pfft_plan_permutation(const pfft_plan* plan, int32_t which, int32_t* perm) {
  NEED(plan);
  NEED(perm);
  const Plan& p = plan->p;
  const Stage& s = which ? p.stages.back() : p.stages.front();
  for (int i = 0; i < p.N; ++i) perm[i] = s.perm[i] + 1;
  return PFFT_OK;
}

int pfft_plan_eltypes(const pfft_plan* plan, int32_t* eltype_in, int32_t* eltype_out) {
  NEED(plan);
  if (eltype_in) *eltype_in = plan->p.elt_in;
  if (eltype_out) *eltype_out = plan->p.elt_out;
  return PFFT_OK;
}

int pfft_plan_is_inplace(const pfft_plan* plan, int32_t* inplace) {
  NEED(plan);
  NEED(inplace);
  *inplace = plan->p.inplace ? 1 : 0;
  return PFFT_OK;
}

int pfft_plan_scale_factor(const pfft_plan* plan, double* scale) {
  NEED(plan);
  NEED(scale);
  *scale = plan->p.scale;
  return PFFT_OK;
}

int pfft_plan_kinds(const pfft_plan* plan, int32_t* kinds_fwd, int32_t* kinds_bwd) {
  NEED(plan);
  for (int i = 0; i < plan->p.N; ++i) {
    if (kinds_fwd) kinds_fwd[i] = plan->p.stages[i].kind_fwd;
    if (kinds_bwd) kinds_bwd[i] = plan->p.stages[i].kind_bwd;
  }
  return PFFT_OK;
}

int pfft_plan_nbytes(const pfft_plan* plan, size_t* in_bytes, size_t* out_bytes,
                     size_t* work_bytes) {
  NEED(plan);
  if (in_bytes) *in_bytes = plan->p.in_bytes;
  if (out_bytes) *out_bytes = plan->p.out_bytes;
  if (work_bytes) *work_bytes = plan->p.work_bytes;
  return PFFT_OK;
}

int pfft_plan_stage(const pfft_plan* plan, int32_t stage, int32_t* perm_in, int64_t* dims_in,
                    int32_t* perm_out, int64_t* dims_out, int32_t* mem_axis) {
  NEED(plan);
  const Plan& p = plan->p;
  if (stage < 0 || stage >= p.N) return fail(PFFT_ERR_ARGUMENT, "stage out of range");
  const Stage& s = p.stages[stage];
  for (int i = 0; i < p.N; ++i) {
    if (perm_in) perm_in[i] = s.perm[i] + 1;
    if (perm_out) perm_out[i] = s.perm[i] + 1;  // Pi and Po share the permutation
    if (dims_in) dims_in[i] = s.dims_before[i];
    if (dims_out) dims_out[i] = s.dims_after[i];
  }
  if (mem_axis) *mem_axis = s.mem_axis + 1;
  return PFFT_OK;
}

int pfft_plan_describe(const pfft_plan* plan, char* buf, size_t buflen) {
  NEED(plan);
  NEED(buf);
  PFFT_GUARD({
    const std::string s = plan_describe(plan->p);
    if (buflen < s.size() + 1) return fail(PFFT_ERR_ARGUMENT, "buffer too small");
    memcpy(buf, s.c_str(), s.size() + 1);
    return PFFT_OK;
  })
}

int pfft_execute(pfft_plan* plan, int32_t direction, const void* in, void* out, void* work,
                 size_t work_bytes, void* hip_stream) {
  NEED(plan);
  PFFT_GUARD(return device_execute(plan->p, direction, in, out, work, work_bytes,
                                   (hipStream_t)hip_stream);)
}

int pfft_execute_noalloc(pfft_plan* plan, int32_t direction, const void* in, void* out,
                         void* hip_stream) {
  NEED(plan);
  PFFT_GUARD(return device_execute_noalloc(plan->p, direction, in, out,
                                           (hipStream_t)hip_stream);)
}

int pfft_execute_op(pfft_plan* plan, int32_t direction, const pfft_spectral_op* op,
                    const void* in, void* out, void* work, size_t work_bytes, void* hip_stream) {
  NEED(plan);
  NEED(op);
  PFFT_GUARD(return device_execute(plan->p, direction, in, out, work, work_bytes,
                                   (hipStream_t)hip_stream, op->op ? op : nullptr);)
}

int pfft_execute_batch(int32_t n_items, const pfft_batch_item* items, void* hip_stream) {
  if (n_items < 0) return fail(PFFT_ERR_ARGUMENT, "negative item count");
  if (n_items == 0) return PFFT_OK;
  NEED(items);
  PFFT_GUARD({
    std::vector<Plan*> plans((size_t)n_items);
    for (int32_t i = 0; i < n_items; ++i) {
      if (!items[i].plan) return fail(PFFT_ERR_ARGUMENT, "null plan in batch item");
      plans[i] = &items[i].plan->p;
    }
    return device_execute_batch(n_items, plans.data(), items, (hipStream_t)hip_stream);
  })
}

int pfft_project_divergence_free(pfft_plan* plan, void* uh_x, void* uh_y, void* uh_z,
                                 const double* scale, void* hip_stream) {
  NEED(plan);
  void* const uh[3] = {uh_x, uh_y, uh_z};
  PFFT_GUARD(return device_project(plan->p, uh, scale, (hipStream_t)hip_stream);)
}

int pfft_execute_many(pfft_plan* plan, int32_t direction, int32_t n_arrays,
                      const void* const* ins, void* const* outs, void* work, size_t work_bytes,
                      void* hip_stream) {
  NEED(plan);
  NEED(ins);
  NEED(outs);
  if (n_arrays < 0) return fail(PFFT_ERR_ARGUMENT, "negative array count");
  PFFT_GUARD({
    for (int i = 0; i < n_arrays; ++i) {
      int r = device_execute(plan->p, direction, ins[i], outs[i], work, work_bytes,
                             (hipStream_t)hip_stream);
      if (r) return r;
    }
    return PFFT_OK;
  })
}

int pfft_plan_kernels(pfft_plan* plan, int32_t direction, char* buf, size_t buflen) {
  NEED(plan);
  NEED(buf);
  PFFT_GUARD({
    std::string s;
    int r = device_pass_names(plan->p, direction, &s);
    if (r) return r;
    if (buflen < s.size() + 1) return fail(PFFT_ERR_ARGUMENT, "buffer too small");
    memcpy(buf, s.c_str(), s.size() + 1);
    return PFFT_OK;
  })
}

int pfft_tune_export(char* buf, size_t buflen, size_t* needed) {
  PFFT_GUARD({
    std::string s;
    int r = tune_export(&s);
    if (r) return r;
    if (needed) *needed = s.size() + 1;
    if (!buf) return PFFT_OK;
    if (buflen < s.size() + 1) return fail(PFFT_ERR_ARGUMENT, "buffer too small");
    memcpy(buf, s.c_str(), s.size() + 1);
    return PFFT_OK;
  })
}

int pfft_tune_import(const char* text) {
  NEED(text);
  PFFT_GUARD(return tune_import(text);)
}

int pfft_tune_forget(void) {
  PFFT_GUARD({
    tune_forget();
    return PFFT_OK;
  })
}

int pfft_plan_set_timing(pfft_plan* plan, int32_t enable) {
  NEED(plan);
  PFFT_GUARD(return device_set_timing(plan->p, enable != 0);)
}

int pfft_plan_stage_times(pfft_plan* plan, int32_t n, float* ms) {
  NEED(plan);
  NEED(ms);
  PFFT_GUARD(return device_stage_times(plan->p, n, ms);)
}

}  // extern "C"
