// Device side of the transform pipeline: kernels, device tables and the
// execution layer (the reference's _apply_plans_out_of_place!,
// src/operations.jl:227-257, with the transposes fused into the passes and
// ldiv!'s normalisation, operations.jl:50,64-66, fused into the last pass).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <sstream>
#include <string>

#include "../../include/pfft.h"
#include "fast_pass.hpp"
#include "generic_pass.hpp"
#include "plan.hpp"

namespace pfft {

// ---- generic LDS pass ---------------------------------------------------------
// (NT: the workgroup size it is compiled for; the launch uses exactly that.
// PRIME: with the radix-7/11/13 butterflies, whose registers would cost
// every other pass occupancy; used for lines that have such a stage)
template <typename R, int NT, bool PRIME>
__global__ __launch_bounds__(NT) void k_generic(PassArgs a) {
  extern __shared__ __align__(16) char lds[];
  for (int64_t tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
    TileCtx c;
    tile_init(a, tile, c);
    phase_load<R>(a, c, threadIdx.x, blockDim.x, lds);
    __syncthreads();
    if (!a.load_w0) {
      phase_pre<R>(a, c, threadIdx.x, blockDim.x, lds);
      __syncthreads();
    }
    const int nsteps = fft_steps(a);
    for (int s = 0; s < nsteps; ++s) {
      phase_fft_step<R, PRIME>(a, c, s, threadIdx.x, blockDim.x, lds);
      __syncthreads();
    }
    phase_post<R>(a, c, threadIdx.x, blockDim.x, lds);
    __syncthreads();
  }
}

// lines with a radix-7, 11 or 13 stage (the butterfly instance)
static bool prime_stages(const PassArgs& a) {
  if (a.bs_M) return false;
  for (int i = 0; i < a.nstages; ++i)
    if (a.radix[i] == 7 || a.radix[i] == 11 || a.radix[i] == 13) return true;
  return false;
}

// the generic kernel instance of a pass (element type, workgroup size,
// prime butterflies)
template <typename R>
static const void* generic_fn_t(int threads, bool prime) {
  if (prime) return (const void*)k_generic<R, kBlockThreadsHeavy, true>;
  return threads == kBlockThreadsHeavy ? (const void*)k_generic<R, kBlockThreadsHeavy, false>
                                       : (const void*)k_generic<R, kBlockThreads, false>;
}
static const void* generic_fn(int real_type, const PassPlan& pp) {
  const bool prime = prime_stages(pp.args) && pp.threads == kBlockThreadsHeavy;
  return real_type == PFFT_F32 ? generic_fn_t<float>(pp.threads, prime)
                               : generic_fn_t<double>(pp.threads, prime);
}

template <typename R>
static void launch_generic_t(const PassPlan& pp, const PassArgs& a, hipStream_t st) {
  const dim3 g((unsigned)pp.grid), b(pp.threads);
  const size_t lds = (size_t)pp.lds_bytes;
  if (pp.threads == kBlockThreadsHeavy && prime_stages(a))
    hipLaunchKernelGGL((k_generic<R, kBlockThreadsHeavy, true>), g, b, lds, st, a);
  else if (pp.threads == kBlockThreadsHeavy)
    hipLaunchKernelGGL((k_generic<R, kBlockThreadsHeavy, false>), g, b, lds, st, a);
  else
    hipLaunchKernelGGL((k_generic<R, kBlockThreads, false>), g, b, lds, st, a);
}
static void launch_generic(const PassPlan& pp, int real_type, const PassArgs& a,
                           hipStream_t st) {
  if (real_type == PFFT_F32) launch_generic_t<float>(pp, a, st);
  else launch_generic_t<double>(pp, a, st);
}

#define HIP_TRY(expr)                                                             \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess)                                                         \
      return fail(PFFT_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct DeviceState {
  int device = -1;
  void* work = nullptr;
  size_t work_size = 0;
  bool timing = false;
  hipEvent_t ev[2 * kMaxDims + 2] = {};
  int n_ev = 0;
};

static int set_lds_limit(const void* fn, int64_t bytes) {
  if (bytes > 64 * 1024)
    HIP_TRY(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  return PFFT_OK;
}

// ---- autotuner ------------------------------------------------------------------
// A pass can usually run on several specialised variants (tile shape, lines
// per wave).  The first use of a pass shape times every candidate on scratch
// buffers of the pass's size and keeps the fastest; the choice is cached per
// process by the pass signature and its place in the plan.  Within one line
// length the variants differ in how work is spread, not in the arithmetic,
// but their results agree to rounding only: 1024-point lines' two-wavefront
// variants (W = 2) run the 8.8.4.4 radix schedule instead of 16.16.4, and
// separate template instances may contract a multiply-add differently.  A
// plan pinned by PFFT_FLAG_NO_AUTOTUNE (or by imported tune data) is
// bit-reproducible.
static std::mutex g_tune_mu;
static std::map<std::string, int> g_tune_cache;

// The key names the pass signature and the candidate set, so plans whose
// flags admit different kernels (precision variants) never share a choice.
static std::string pass_key(const Plan& p, const PassPlan& pp, const std::vector<int>& cands) {
  const PassArgs& a = pp.args;
  std::ostringstream o;
  for (int v : cands) o << v << ',';
  o << p.real_type << ':' << a.kind << ':' << a.M << ':' << a.n << ':' << a.L_in << ':' << a.L_out
    << ':' << a.sk_in << ':' << a.sk_out << ':' << a.in_cplx << a.out_cplx << ':' << a.D;
  for (int d = 0; d < a.D; ++d) o << ':' << a.ldim[d] << ',' << a.lsin[d] << ',' << a.lsout[d];
  return o.str();
}

// Signature of a whole plan for the tune cache: every pass of both
// directions, in-place or not, and whether the in-chain phases run.
static std::string plan_signature(const Plan& p, bool chain) {
  std::string s = std::string(p.inplace ? "ip" : "oop") + (chain ? ",chain" : ",pass");
  const std::vector<int> none;
  for (const std::vector<PassPlan>* list : {&p.fwd, &p.bwd}) {
    s += "[";
    for (const PassPlan& pp : *list) s += pass_key(p, pp, none) + ";";
    s += "]";
  }
  return s;
}

// op_side > 0 times an operator-carrying variant with a derivative operator
// on that side (the arithmetic of the operator included).  With a producer
// (the pass before this one in the chain, already tuned), each timed launch
// follows a launch of the producer writing this pass's input, so the input
// meets the caches as it does in the pipeline.
static constexpr int kTrials = 5;
static int time_variant(Plan& p, const PassPlan& base, int variant, void* in, void* out,
                        const PassPlan* prod, void* prod_in, hipStream_t st,
                        const hipEvent_t* e0, const hipEvent_t* e1, float* ms,
                        int op_side = 0) {
  PassPlan pp = base;
  fast_apply(pp, variant);
  int r = fast_prepare(pp);
  if (r) return r;
  PassArgs a = pp.args;
  if (op_side) {
    const pfft_spectral_op op = {PFFT_OP_DERIVATIVE, 0, 1.0, 0};
    op_fill(p, op_side == 2 ? PFFT_FORWARD : PFFT_BACKWARD, op, op_side, a);
  }
  a.in = in;
  a.out = out;
  bind_tables(p, pp, a);
  PassArgs pa;
  if (prod) {
    pa = prod->args;
    pa.in = prod_in;
    pa.out = in;
    bind_tables(p, *prod, pa);
  }
  auto produce = [&]() { return prod ? fast_launch(*prod, pa, st) : PFFT_OK; };
  r = produce();
  if (!r) r = fast_launch(pp, a, st);  // warm-up
  if (r) return r;
  // the median of a few launches: a typical launch, robust to the occasional
  // slow one (a minimum would favour erratic variants); fewer on big passes.
  // The launches are queued back to back and waited for once, so the device
  // stays busy as in the pipeline.
  const int trials = (double)pp.in_elems * eltype_bytes(pp.elt_in) > 4e9 ? 3 : kTrials;
  float t[kTrials];
  for (int i = 0; i < trials; ++i) {
    r = produce();
    if (r) return r;
    HIP_TRY(hipEventRecord(e0[i], st));
    r = fast_launch(pp, a, st);
    if (r) return r;
    HIP_TRY(hipEventRecord(e1[i], st));
  }
  HIP_TRY(hipEventSynchronize(e1[trials - 1]));
  for (int i = 0; i < trials; ++i) HIP_TRY(hipEventElapsedTime(&t[i], e0[i], e1[i]));
  HIP_TRY(hipGetLastError());
  std::sort(t, t + trials);
  *ms = t[trials / 2];
  return PFFT_OK;
}

// A job times the candidates of one pass: a plain pass (op_side 0, the
// choice is applied to it) or the operator-carrying form of the last forward /
// first backward pass (op_side 2 / 1, the choice is kept aside in variant_op
// for pfft_execute_op).
struct TuneJob {
  PassPlan* pp;
  PassPlan base;
  std::vector<int> cands;
  std::string key;
  int op_side;
  const PassPlan* prod;   // the pass writing this one's input (nullptr: none)
  int list, index;        // position in the plan: list 0 forward, 1 backward
  std::vector<float> t1;  // phase-1 time of each candidate
};

// One pass of the chain on the given buffers (the timing loops of the tuner).
static int launch_pass(const Plan& p, const PassPlan& pp, const void* src, void* dst,
                       bool scale, hipStream_t st) {
  PassArgs a = pp.args;
  a.in = src;
  a.out = dst;
  bind_tables(p, pp, a);
  if (scale) {
    a.do_scale = 1;
    a.scale = 1.0 / p.scale;
  }
  if (pp.variant == 0) {
    launch_generic(pp, p.real_type, a, st);
    return PFFT_OK;
  }
  return fast_launch(pp, a, st);
}

static int prepare_pass(const Plan& p, const PassPlan& pp) {
  if (pp.variant != 0) return fast_prepare(pp);
  return set_lds_limit(generic_fn(p.real_type, pp), pp.lds_bytes);
}

// Second tuning phase: every plain pass's leading candidates again, judged by
// the time of whole mul! + ldiv! pairs on buffers of the plan's sizes (the
// other passes at their current choice).  The caches, clocks and neighbouring
// kernels are those of the pipeline, and a choice that speeds its own pass
// but slows the next one loses.  Out-of-place plans; skipped when the buffers
// do not fit.
static int refine_in_chain(Plan& p, std::vector<TuneJob>& jobs, bool log, bool* ran) {
  *ran = true;
  if (p.inplace) return PFFT_OK;
  bool any = false;
  for (const TuneJob& j : jobs) any = any || j.op_side == 0;
  if (!any) return PFFT_OK;
  *ran = false;
  void *u = nullptr, *v = nullptr, *w = nullptr;
  hipStream_t st = nullptr;
  const int kPairs = 5;
  hipEvent_t e0[kPairs] = {}, e1[kPairs] = {};
  auto cleanup = [&]() {
    if (st) (void)hipStreamSynchronize(st);
    for (void* b : {u, v, w})
      if (b) (void)hipFree(b);
    for (int i = 0; i < kPairs; ++i) {
      if (e0[i]) (void)hipEventDestroy(e0[i]);
      if (e1[i]) (void)hipEventDestroy(e1[i]);
    }
    if (st) (void)hipStreamDestroy(st);
  };
  // plan-sized scratch only while it leaves most of the device free
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess ||
      2 * (p.in_bytes + p.out_bytes + p.work_bytes) > free_b) {
    (void)hipGetLastError();
    return PFFT_OK;
  }
  bool ok = hipMalloc(&u, p.in_bytes) == hipSuccess && hipMalloc(&v, p.out_bytes) == hipSuccess &&
            (!p.work_bytes || hipMalloc(&w, p.work_bytes) == hipSuccess) &&
            hipStreamCreate(&st) == hipSuccess;
  for (int i = 0; ok && i < kPairs; ++i)
    ok = hipEventCreate(&e0[i]) == hipSuccess && hipEventCreate(&e1[i]) == hipSuccess;
  if (!ok) {   // no room: keep the first phase's choices
    cleanup();
    (void)hipGetLastError();
    return PFFT_OK;
  }
  *ran = true;
  (void)hipMemsetD32Async((hipDeviceptr_t)u, 0x3f800000u, p.in_bytes / 4, st);
  void* bufs[3] = {nullptr, w, w ? (char*)w + p.work0_bytes : nullptr};
  // runs `pairs` mul! + ldiv! pairs, recording events around pair i
  auto run_pairs = [&](int pairs) -> int {
    for (int i = 0; i < pairs; ++i) {
      if (i < kPairs) HIP_TRY(hipEventRecord(e0[i], st));
      for (int d = 0; d < 2; ++d) {
        const std::vector<PassPlan>& passes = d == 0 ? p.fwd : p.bwd;
        void* in = d == 0 ? u : v;
        void* out = d == 0 ? v : u;
        const int np = (int)passes.size();
        for (int k = 0; k < np; ++k) {
          const PassPlan& pp = passes[k];
          const void* src = pp.src == 0 ? in : bufs[pp.src];
          void* dst = pp.dst == 0 ? out : bufs[pp.dst];
          const int r = launch_pass(p, pp, src, dst, d == 1 && k == np - 1, st);
          if (r) return r;
        }
      }
      if (i < kPairs) HIP_TRY(hipEventRecord(e1[i], st));
    }
    HIP_TRY(hipGetLastError());
    return PFFT_OK;
  };
  int rc = PFFT_OK;
  for (auto* list : {&p.fwd, &p.bwd})
    for (PassPlan& pp : *list) {
      rc = prepare_pass(p, pp);
      if (rc) break;
    }
  const double pair_bytes = 2.0 * ((double)p.in_bytes + (double)p.out_bytes);
  const int rounds = pair_bytes > 8e9 ? 1 : 2;
  // a candidate replaces the incumbent only when the pairs run faster by
  // kMargin, twice (see below)
  const float kMargin = 0.992f;
  for (TuneJob& j : jobs) {
    if (rc) break;
    if (j.op_side) continue;
    const int lst = j.list, idx = j.index;
    // the phase-1 leaders, each walking its tiles both ways and storing with
    // and without the streaming hint: both decide what the pass leaves in
    // and finds in the last-level cache, which only the chain shows
    std::vector<size_t> order(j.cands.size());
    for (size_t c = 0; c < order.size(); ++c) order[c] = c;
    if (j.t1.size() == j.cands.size())
      std::stable_sort(order.begin(), order.end(),
                       [&](size_t x, size_t y) { return j.t1[x] < j.t1[y]; });
    std::vector<int> cands;
    for (size_t c = 0; c < order.size() && c < 3; ++c)
      for (int bits : {0, kVariantReverse, kVariantStream, kVariantReverse | kVariantStream})
        cands.push_back(j.cands[order[c]] | bits);
    const int vop = j.pp->variant_op;
    auto apply = [&](int var) {
      *j.pp = j.base;
      fast_apply(*j.pp, var);
      j.pp->variant_op = vop;
      return fast_prepare(*j.pp);
    };
    // median pair time with candidate c in place, folded into t[c] as a minimum
    std::vector<float> t(cands.size(), 1e30f);
    auto time_cand = [&](size_t c) -> int {
      int r = apply(cands[c]);
      if (!r) r = run_pairs(1);   // warm-up
      if (!r) r = run_pairs(kPairs);
      if (r) return r;
      if (hipEventSynchronize(e1[kPairs - 1]) != hipSuccess)
        return fail(PFFT_ERR_DEVICE, "autotune: event synchronisation failed");
      float ts[kPairs];
      for (int i = 0; i < kPairs; ++i)
        if (hipEventElapsedTime(&ts[i], e0[i], e1[i]) != hipSuccess) ts[i] = 1e30f;
      std::sort(ts, ts + kPairs);
      t[c] = std::min(t[c], ts[kPairs / 2]);
      return PFFT_OK;
    };
    for (int round = 0; round < rounds && !rc; ++round)
      for (size_t c = 0; c < cands.size() && !rc; ++c) rc = time_cand(c);
    if (rc) break;
    // cands[0] is the first phase's choice: another candidate replaces it
    // only when the pairs run faster by kMargin, twice (the median pair time
    // of one trial scatters by a few tenths of a percent; a streaming store
    // that speeds up the next pass gains about 1 % of the pair)
    size_t best = 0;
    for (size_t c = 1; c < cands.size(); ++c)
      if (t[c] < t[best]) best = c;
    if (t[best] > kMargin * t[0]) best = 0;
    if (best != 0) {
      // the fastest of a dozen noisy timings is biased low: the challenger is
      // timed again from scratch, interleaved with the incumbent, and must win
      // by the same margin
      const float first0 = t[0], firstb = t[best];
      t[0] = t[best] = 1e30f;
      for (int round = 0; round < 4 && !rc; ++round) {
        rc = time_cand(best);
        if (!rc) rc = time_cand(0);
      }
      if (rc) break;
      if (log)
        fprintf(stderr, "pfft autotune (in chain): %s %.4f -> %.4f ms vs %.4f -> %.4f ms\n",
                fast_name(cands[best]), firstb, t[best], first0, t[0]);
      if (t[best] > kMargin * t[0]) best = 0;
    }
    if (log) {
      fprintf(stderr, "pfft autotune (in chain, ms per pair): %s pass %d\n",
              lst == 0 ? "forward" : "backward", idx + 1);
      for (size_t c = 0; c < cands.size(); ++c)
        fprintf(stderr, "  %9.4f ms %s%s\n", t[c], fast_name(cands[c]),
                c == best ? "  <- chosen" : "");
    }
    rc = apply(cands[best]);
    if (rc) break;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[j.key] = cands[best];
  }
  // The streaming bits of all tuned passes together: a streaming store speeds
  // up the pass after it, so the bits interact along the chain, and the pass
  // by pass search above settles in whichever local optimum its visiting
  // order meets first.  With the kernels and tile directions fixed, every
  // assignment of the bits is timed (2^P pairs runs, P <= 6).
  std::vector<TuneJob*> sj;
  for (TuneJob& j : jobs)
    if (!j.op_side) sj.push_back(&j);
  const int P = (int)sj.size();
  if (!rc && pair_bytes <= 8e9 && P >= 2 && P <= 6) {
    std::vector<int> basev(P);
    int cur = 0;
    for (int i = 0; i < P; ++i) {
      basev[i] = sj[i]->pp->variant & ~kVariantStream;
      if (sj[i]->pp->variant & kVariantStream) cur |= 1 << i;
    }
    auto apply_mask = [&](int m) -> int {
      for (int i = 0; i < P; ++i) {
        TuneJob& j = *sj[i];
        const int vop = j.pp->variant_op;
        *j.pp = j.base;
        fast_apply(*j.pp, basev[i] | (((m >> i) & 1) ? kVariantStream : 0));
        j.pp->variant_op = vop;
        const int r = fast_prepare(*j.pp);
        if (r) return r;
      }
      return PFFT_OK;
    };
    std::vector<float> tm((size_t)1 << P, 1e30f);
    auto time_mask = [&](int m) -> int {
      int r = apply_mask(m);
      if (!r) r = run_pairs(1);
      if (!r) r = run_pairs(kPairs);
      if (r) return r;
      if (hipEventSynchronize(e1[kPairs - 1]) != hipSuccess)
        return fail(PFFT_ERR_DEVICE, "autotune: event synchronisation failed");
      float ts[kPairs];
      for (int i = 0; i < kPairs; ++i)
        if (hipEventElapsedTime(&ts[i], e0[i], e1[i]) != hipSuccess) ts[i] = 1e30f;
      std::sort(ts, ts + kPairs);
      tm[m] = std::min(tm[m], ts[kPairs / 2]);
      return PFFT_OK;
    };
    for (int m = 0; m < (1 << P) && !rc; ++m) rc = time_mask(m);
    int best = cur;
    for (int m = 0; m < (1 << P); ++m)
      if (tm[m] < tm[best]) best = m;
    if (!rc && best != cur && tm[best] < kMargin * tm[cur]) {
      const float firstb = tm[best], first0 = tm[cur];
      tm[best] = tm[cur] = 1e30f;
      for (int round = 0; round < 4 && !rc; ++round) {
        rc = time_mask(best);
        if (!rc) rc = time_mask(cur);
      }
      if (log)
        fprintf(stderr,
                "pfft autotune (streaming bits): mask %d %.4f -> %.4f ms vs mask %d %.4f -> "
                "%.4f ms\n",
                best, firstb, tm[best], cur, first0, tm[cur]);
      if (tm[best] > kMargin * tm[cur]) best = cur;
    } else {
      best = cur;
    }
    if (!rc) rc = apply_mask(best);
    if (!rc) {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      for (int i = 0; i < P; ++i) g_tune_cache[sj[i]->key] = sj[i]->pp->variant;
    }
  }
  cleanup();
  return rc;
}

static int autotune(Plan& p) {
  using Job = TuneJob;
  std::vector<Job> jobs;
  size_t in_max = 0, out_max = 0, prod_max = 0;
  const uint32_t forced = (p.flags >> 16) & 0xffu;
  // PFFT_TUNE_CHAIN=0 keeps the first phase's choices
  const char* chain_env = getenv("PFFT_TUNE_CHAIN");
  const bool chain = !(chain_env && chain_env[0] == '0') && !p.inplace;
  const std::string psig = plan_signature(p, chain);
  auto settle = [&](PassPlan& pp, int op_side, int v) {
    if (op_side) pp.variant_op = v;
    else fast_apply(pp, v);
  };
  auto consider = [&](PassPlan& pp, const PassPlan& base, int op_side, const PassPlan* prod,
                      int list, int index) {
    std::vector<int> c = fast_candidates(p, base, op_side != 0);
    if (c.empty()) return;
    const int rev = ((p.flags & PFFT_FLAG_REVERSE_TILES) ? kVariantReverse : 0) |
                    ((p.flags & PFFT_FLAG_STREAM_STORES) ? kVariantStream : 0);
    if (forced) {
      settle(pp, op_side, c[(forced - 1) % c.size()] | rev);
      return;
    }
    // untuned plans take the first candidate whatever a tuned plan of the
    // same shape chose before (the pinned, reproducible choice)
    if (c.size() == 1 || (p.flags & PFFT_FLAG_NO_AUTOTUNE)) {
      settle(pp, op_side, c[0] | rev);
      return;
    }
    // The key holds the pass signature, its position in the chain and the
    // signature of the whole plan: the in-chain phases choose with the
    // pass's neighbours in place, so the same pass in another chain (or in a
    // chain tuned without them) may choose differently.
    const std::string key = pass_key(p, base, c) + (op_side ? ":op" + std::to_string(op_side) : "") +
                            "@" + std::to_string(list) + "." + std::to_string(index) + "|" + psig;
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      // a cached choice (imported ones included) must be one of this pass's
      // candidates, up to the direction and streaming bits
      const int base_v = it->second & ~(kVariantReverse | kVariantStream);
      if (std::find(c.begin(), c.end(), base_v) != c.end()) {
        settle(pp, op_side, it->second);
        return;
      }
    }
    // estimate plans take what earlier tuning (or imported tune data) chose
    // for this pass, else the first candidate, and time nothing
    if (p.flags & PFFT_FLAG_ESTIMATE) {
      settle(pp, op_side, c[0]);
      return;
    }
    jobs.push_back({&pp, base, c, key, op_side, prod, list, index, {}});
    in_max = std::max(in_max, (size_t)pp.in_elems * eltype_bytes(pp.elt_in));
    out_max = std::max(out_max, (size_t)pp.out_elems * eltype_bytes(pp.elt_out));
    if (prod) {
      prod_max = std::max(prod_max, (size_t)prod->in_elems * eltype_bytes(prod->elt_in));
      in_max = std::max(in_max, (size_t)prod->out_elems * eltype_bytes(prod->elt_out));
    }
  };
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (int l = 0; l < 2; ++l) {
      std::vector<PassPlan>& list = l == 0 ? p.fwd : p.bwd;
      for (size_t i = 0; i < list.size(); ++i)
        consider(list[i], list[i], 0, i > 0 ? &list[i - 1] : nullptr, l, (int)i);
    }
    // operator carriers (when the plan admits spectral operators)
    const pfft_spectral_op probe = {PFFT_OP_DERIVATIVE, 0, 1.0, 0};
    for (int dir : {PFFT_FORWARD, PFFT_BACKWARD}) {
      int idx = -1, side = 0;
      if (op_setup(p, dir, probe, &idx, &side) != PFFT_OK) {
        clear_error();
        continue;
      }
      std::vector<PassPlan>& list = dir == PFFT_FORWARD ? p.fwd : p.bwd;
      PassPlan& pp = list[idx];
      PassPlan q = pp;
      q.args = q.gen_args;
      q.threads = q.gen_threads;
      q.lds_bytes = q.gen_lds;
      q.grid = q.gen_grid;
      q.variant = 0;
      // a pass without fast variants runs (and carries the operator) on the
      // generic kernel; the others are considered whether or not the plain
      // pass was settled from the cache above, so that a fresh and a warm
      // tune cache tune the same set
      if (fast_candidates(p, q, false).empty()) continue;
      consider(pp, q, side, idx > 0 ? &list[idx - 1] : nullptr, dir == PFFT_FORWARD ? 0 : 1,
               idx);
    }
  }
  if (jobs.empty()) return PFFT_OK;
  void *in = nullptr, *out = nullptr, *pin = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0[kTrials] = {}, e1[kTrials] = {};
  int rc = PFFT_OK;
  auto cleanup = [&]() {
    if (st) (void)hipStreamSynchronize(st);
    if (in) (void)hipFree(in);
    if (out) (void)hipFree(out);
    if (pin) (void)hipFree(pin);
    for (int i = 0; i < kTrials; ++i) {
      if (e0[i]) (void)hipEventDestroy(e0[i]);
      if (e1[i]) (void)hipEventDestroy(e1[i]);
    }
    if (st) (void)hipStreamDestroy(st);
  };
  if (hipMalloc(&in, in_max) != hipSuccess || hipMalloc(&out, out_max) != hipSuccess ||
      (prod_max && hipMalloc(&pin, prod_max) != hipSuccess) ||
      hipStreamCreate(&st) != hipSuccess || [&]() {
        for (int i = 0; i < kTrials; ++i)
          if (hipEventCreate(&e0[i]) != hipSuccess || hipEventCreate(&e1[i]) != hipSuccess)
            return true;
        return false;
      }()) {
    cleanup();
    // no room to tune: fall back to the default variants
    for (Job& j : jobs) settle(*j.pp, j.op_side, j.cands[0]);
    (void)hipGetLastError();
    return PFFT_OK;
  }
  // a nonzero input pattern (1.0f words): the kernels' time does not depend
  // on the data, but the memory system is exercised as with real fields
  (void)hipMemsetD32Async((hipDeviceptr_t)in, 0x3f800000u, in_max / 4, st);
  if (pin) (void)hipMemsetD32Async((hipDeviceptr_t)pin, 0x3f800000u, prod_max / 4, st);
  // PFFT_TUNE_LOG=1 in the environment prints every timing and choice
  const char* log_env = getenv("PFFT_TUNE_LOG");
  const bool tune_log = log_env && log_env[0] == '1';
  auto tune = [&](Job& j, const PassPlan* prod) -> int {
    // the producer, as tuned (a generic producer is left out: its LDS limit
    // is set after tuning)
    if (prod && prod->variant == 0) prod = nullptr;
    if (prod) {
      const int r = fast_prepare(*prod);
      if (r) return r;
    }
    // two rounds over all candidates, keeping each one's better median: the
    // first round also brings the clocks up, so early candidates are not
    // judged cold
    std::vector<float> t(j.cands.size(), 1e30f);
    for (int round = 0; round < 2; ++round)
      for (size_t i = 0; i < j.cands.size(); ++i) {
        float ms = 0;
        const int r = time_variant(p, j.base, j.cands[i], in, out, prod, pin, st, e0, e1, &ms,
                                   j.op_side);
        if (r) return r;
        t[i] = ms < t[i] ? ms : t[i];
      }
    int best = j.cands[0];
    float best_ms = 1e30f;
    for (size_t i = 0; i < j.cands.size(); ++i)
      if (t[i] < best_ms) { best_ms = t[i]; best = j.cands[i]; }
    j.t1 = t;
    if (tune_log) {
      fprintf(stderr, "pfft autotune: pass %s%s\n", j.key.c_str(), prod ? " (after producer)" : "");
      for (size_t i = 0; i < j.cands.size(); ++i)
        fprintf(stderr, "  %9.4f ms %s%s\n", t[i], fast_name(j.cands[i]),
                j.cands[i] == best ? "  <- chosen" : "");
    }
    settle(*j.pp, j.op_side, best);
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[j.key] = best;
    return PFFT_OK;
  };
  for (Job& j : jobs) {
    rc = tune(j, j.prod);
    if (rc) break;
  }
  cleanup();
  if (!rc && chain) {
    bool ran = false;
    rc = refine_in_chain(p, jobs, tune_log, &ran);
    if (!ran) {
      // the in-chain phases did not fit in device memory: these phase-1
      // choices are not what a chain-tuned plan of this signature would
      // take, so they are not kept for later plans
      std::lock_guard<std::mutex> lk(g_tune_mu);
      for (const Job& j : jobs) g_tune_cache.erase(j.key);
    }
  }
  return rc;
}

// ---- tune data export / import ------------------------------------------------
// The tune cache as text: a header naming the format, the kernel table of
// this build and the device, then one "variant<TAB>key" line per choice.
// Variant ids index this build's kernel table, so data from another build
// or another kind of device is refused.
static std::string device_tag() {
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
    (void)hipGetLastError();
    return "none";
  }
  return std::string(prop.gcnArchName) + "/" + std::to_string(prop.multiProcessorCount);
}

static std::string tune_header() {
  return "pfft-tune 1\nbuild " + fast_build_signature() + "\ndevice " + device_tag() + "\n";
}

int tune_export(std::string* out) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  *out = tune_header();
  for (const auto& kv : g_tune_cache) *out += std::to_string(kv.second) + "\t" + kv.first + "\n";
  return PFFT_OK;
}

int tune_import(const char* text) {
  const std::string s(text);
  const std::string head = tune_header();
  if (s.compare(0, head.size(), head) != 0)
    return fail(PFFT_ERR_ARGUMENT,
                "tune data of another format, build or device (expected a header of\n" + head +
                    ")");
  std::map<std::string, int> got;
  size_t pos = head.size();
  while (pos < s.size()) {
    size_t nl = s.find('\n', pos);
    if (nl == std::string::npos) nl = s.size();
    const std::string line = s.substr(pos, nl - pos);
    pos = nl + 1;
    if (line.empty()) continue;
    const size_t tab = line.find('\t');
    char* end = nullptr;
    const long v = tab == std::string::npos ? 0 : strtol(line.c_str(), &end, 10);
    if (tab == std::string::npos || end != line.c_str() + tab || v < 0 || tab + 1 >= line.size())
      return fail(PFFT_ERR_ARGUMENT, "malformed tune data line: " + line);
    got[line.substr(tab + 1)] = (int)v;
  }
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (const auto& kv : got) g_tune_cache[kv.first] = kv.second;
  return PFFT_OK;
}

void tune_forget() {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune_cache.clear();
}

int device_prepare(Plan& p) {
  if (p.dev) return PFFT_OK;
  DeviceState* d = new DeviceState();
  // on failure the tables go too, so a later prepare starts clean
  auto abandon = [&](int code) {
    delete d;
    for (Table& t : p.tables) {
      if (t.dev) (void)hipFree(t.dev);
      t.dev = nullptr;
    }
    return code;
  };
  if (hipGetDevice(&d->device) != hipSuccess)
    return abandon(fail(PFFT_ERR_DEVICE, "no HIP device available"));
  for (Table& t : p.tables) {
    const size_t bytes = t.host.size();
    hipError_t e = hipMalloc(&t.dev, bytes);
    if (e != hipSuccess) {
      t.dev = nullptr;
      return abandon(
          fail(PFFT_ERR_ALLOC, std::string("hipMalloc(tables): ") + hipGetErrorString(e)));
    }
    e = hipMemcpy(t.dev, t.host.data(), bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess)
      return abandon(
          fail(PFFT_ERR_DEVICE, std::string("hipMemcpy(tables): ") + hipGetErrorString(e)));
  }
  int r = autotune(p);
  if (r) return abandon(r);
  for (auto* list : {&p.fwd, &p.bwd}) {
    for (PassPlan& pp : *list) {
      if (pp.variant == 0) {
        r = set_lds_limit(generic_fn(p.real_type, pp), pp.lds_bytes);
      } else {
        r = fast_prepare(pp);
      }
      if (r) return abandon(r);
    }
  }
  p.dev = d;
  return PFFT_OK;
}

// ---- batches of plans (pfft_execute_batch) -----------------------------------
// Per device: four streams (default priority) and the fork / join events,
// created on the device's first batch and kept to the end of the process.
struct BatchStreams {
  hipStream_t s[4] = {};
  hipEvent_t fork = nullptr;
  hipEvent_t join[4] = {};
};
static std::mutex g_batch_mu;
static std::map<int, BatchStreams> g_batch;

int device_execute_batch(int n, Plan* const* plans, const pfft_batch_item* items,
                         hipStream_t stream) {
  // the plans in order of first appearance: plan k runs on stream k mod 4
  std::vector<const Plan*> order;
  std::vector<int> lane((size_t)n);
  for (int i = 0; i < n; ++i) {
    size_t k = 0;
    while (k < order.size() && order[k] != plans[i]) ++k;
    if (k == order.size()) order.push_back(plans[i]);
    lane[i] = (int)(k % 4);
  }
  auto run = [&](int i, hipStream_t st) {
    const pfft_batch_item& it = items[i];
    const pfft_spectral_op* op = it.op && it.op->op ? it.op : nullptr;
    return device_execute(*plans[i], it.direction, it.in, it.out, it.work, it.work_bytes, st, op);
  };
  if (order.size() == 1) {   // one plan: its items in order on the caller's stream
    for (int i = 0; i < n; ++i) {
      const int r = run(i, stream);
      if (r) return r;
    }
    return PFFT_OK;
  }
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_batch_mu);
  BatchStreams& b = g_batch[dev];
  if (!b.fork) {
    for (int k = 0; k < 4; ++k) {
      HIP_TRY(hipStreamCreateWithFlags(&b.s[k], hipStreamNonBlocking));
      HIP_TRY(hipEventCreateWithFlags(&b.join[k], hipEventDisableTiming));
    }
    HIP_TRY(hipEventCreateWithFlags(&b.fork, hipEventDisableTiming));
  }
  const int used = (int)std::min<size_t>(order.size(), 4);
  HIP_TRY(hipEventRecord(b.fork, stream));
  for (int k = 0; k < used; ++k) HIP_TRY(hipStreamWaitEvent(b.s[k], b.fork, 0));
  int rc = PFFT_OK;
  for (int i = 0; i < n && !rc; ++i) rc = run(i, b.s[lane[i]]);
  // join every forked stream back, also after a failed item
  for (int k = 0; k < used; ++k) {
    HIP_TRY(hipEventRecord(b.join[k], b.s[k]));
    HIP_TRY(hipStreamWaitEvent(stream, b.join[k], 0));
  }
  return rc;
}

int device_pass_names(Plan& p, int dir, std::string* out) {
  int r = device_prepare(p);
  if (r) return r;
  const std::vector<PassPlan>& passes = dir == PFFT_FORWARD ? p.fwd : p.bwd;
  out->clear();
  for (size_t i = 0; i < passes.size(); ++i) {
    if (i) *out += ";";
    *out += passes[i].variant == 0
                ? std::string(p.real_type == PFFT_F32 ? "k_generic<f32>" : "k_generic<f64>")
                : std::string(fast_name(passes[i].variant));
    if (passes[i].variant == 0 && passes[i].args.bs_M)
      *out += "/bluestein" + std::to_string(passes[i].args.bs_M);
  }
  return PFFT_OK;
}

void device_release(Plan& p) {
  for (Table& t : p.tables) {
    if (t.dev) (void)hipFree(t.dev);
    t.dev = nullptr;
  }
  if (p.dev) {
    if (p.dev->work) (void)hipFree(p.dev->work);
    for (int i = 0; i < p.dev->n_ev; ++i) (void)hipEventDestroy(p.dev->ev[i]);
    delete p.dev;
    p.dev = nullptr;
  }
}

// ---- divergence-free projection (pfft_project_divergence_free) -----------------
struct ProjArgs {
  int64_t n[3];       // output memory dims of the transformed axes
  int64_t rest;       // product of the extra dims
  int32_t ld[3];      // logical dim of each memory axis
  int32_t mode[3];    // per logical dim: 1 rfftfreq, 0 fftfreq
  int64_t kn[3];      // per logical dim: spectrum length
  double sc[3];       // per logical dim: 2 pi / L
};

template <typename R>
__device__ __forceinline__ R proj_k(const ProjArgs& a, int d, int64_t c) {
  const int64_t n = a.kn[d];
  const int64_t k = a.mode[d] ? c : (c < (n + 1) / 2 ? c : c - n);
  return (R)((double)k * a.sc[d]);
}

// One thread per (memory axis 0, 1) element; the other axes in a loop.
// The projection reads all three components before writing any.
template <typename R>
__global__ __launch_bounds__(256) void k_project(cpx<R>* __restrict__ ux, cpx<R>* __restrict__ uy,
                                                 cpx<R>* __restrict__ uz, ProjArgs a) {
  const int64_t m0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m0 >= a.n[0]) return;
  const int64_t n01 = a.n[0] * a.n[1];
  const int64_t outer = a.n[2] * a.rest;
  for (int64_t m1 = blockIdx.y; m1 < a.n[1]; m1 += gridDim.y) {
    R k[3];
    k[a.ld[0]] = proj_k<R>(a, a.ld[0], m0);
    k[a.ld[1]] = proj_k<R>(a, a.ld[1], m1);
    for (int64_t r = blockIdx.z; r < outer; r += gridDim.z) {
      k[a.ld[2]] = proj_k<R>(a, a.ld[2], r % a.n[2]);
      const R k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
      if (k2 == R(0)) continue;
      const int64_t e = m0 + a.n[0] * m1 + n01 * r;
      const cpx<R> x = ux[e], y = uy[e], z = uz[e];
      // (k . u) / |k|^2, complex
      const R inv = R(1) / k2;
      const cpx<R> s = mk((k[0] * x.re + k[1] * y.re + k[2] * z.re) * inv,
                          (k[0] * x.im + k[1] * y.im + k[2] * z.im) * inv);
      ux[e] = mk(x.re - k[0] * s.re, x.im - k[0] * s.im);
      uy[e] = mk(y.re - k[1] * s.re, y.im - k[1] * s.im);
      uz[e] = mk(z.re - k[2] * s.re, z.im - k[2] * s.im);
    }
  }
}

int device_project(Plan& p, void* const* uh, const double* scale, hipStream_t stream) {
  if (p.N != 3)
    return fail(PFFT_ERR_ARGUMENT, "the divergence-free projection needs a 3D plan");
  for (int d = 0; d < 3; ++d) {
    const int32_t b = kind_base(p.kinds[d]);
    if (b != PFFT_FFT && b != PFFT_BFFT && b != PFFT_RFFT)
      return fail(PFFT_ERR_UNSUPPORTED,
                  "the divergence-free projection needs wavenumbers: FFT/BFFT/RFFT dimensions");
    if (!uh[d]) return fail(PFFT_ERR_ARGUMENT, "null component array");
  }
  for (int i = 0; i < 3; ++i)
    for (int j = i + 1; j < 3; ++j)
      if (uh[i] == uh[j]) return fail(PFFT_ERR_ARGUMENT, "the three components must be distinct");
  ProjArgs a;
  const Stage& s = p.stages.back();
  for (int i = 0; i < 3; ++i) {
    a.ld[i] = s.perm[i];
    a.n[i] = p.size_out[s.perm[i]];
  }
  a.rest = p.E;
  for (int d = 0; d < 3; ++d) {
    a.mode[d] = kind_base(p.kinds[d]) == PFFT_RFFT ? 1 : 0;
    a.kn[d] = p.size_out[d];
    a.sc[d] = scale ? scale[d] : 1.0;
  }
  const dim3 block(256);
  const dim3 grid((unsigned)((a.n[0] + 255) / 256), (unsigned)std::min<int64_t>(a.n[1], 65535),
                  (unsigned)std::min<int64_t>(a.n[2] * a.rest, 65535));
  if (p.real_type == PFFT_F32)
    hipLaunchKernelGGL(k_project<float>, grid, block, 0, stream, (cpx<float>*)uh[0],
                       (cpx<float>*)uh[1], (cpx<float>*)uh[2], a);
  else
    hipLaunchKernelGGL(k_project<double>, grid, block, 0, stream, (cpx<double>*)uh[0],
                       (cpx<double>*)uh[1], (cpx<double>*)uh[2], a);
  HIP_TRY(hipGetLastError());
  return PFFT_OK;
}

int device_set_timing(Plan& p, bool on) {
  int r = device_prepare(p);
  if (r) return r;
  DeviceState* d = p.dev;
  if (on && d->n_ev == 0) {
    d->n_ev = (int)std::max(p.fwd.size(), p.bwd.size()) + 1;
    for (int i = 0; i < d->n_ev; ++i) HIP_TRY(hipEventCreate(&d->ev[i]));
  }
  d->timing = on;
  return PFFT_OK;
}

int device_stage_times(Plan& p, int n, float* ms) {
  if (!p.dev || !p.dev->timing) return fail(PFFT_ERR_ARGUMENT, "timing is not enabled on this plan");
  DeviceState* d = p.dev;
  const int np = d->n_ev - 1;
  if (n < np) return fail(PFFT_ERR_ARGUMENT, "buffer too small for the stage times");
  HIP_TRY(hipEventSynchronize(d->ev[np]));
  for (int i = 0; i < np; ++i) HIP_TRY(hipEventElapsedTime(&ms[i], d->ev[i], d->ev[i + 1]));
  return PFFT_OK;
}

// ---- execution without allocation (pfft_execute_noalloc) -------------------------
// The intermediate pencils of a plan (the workspace of pfft_execute) live in
// this array instead of caller or plan memory: each GPU the library runs on
// holds one copy of kNoAllocScratch bytes until the process ends.  Host code
// never passes `scratch` itself to a kernel or copy -- on the host that name
// stands for a host-side shadow, not a device address -- but the device
// address hipGetSymbolAddress gives for the current device, looked up once
// per device and kept in g_scratch_ptr.  Calls on one GPU share the array:
// two executes that overlap in time on one device (other streams, other
// threads) race on it.  Every pass writes its whole output before the next
// reads it, so nothing in it needs clearing between calls.
constexpr size_t kNoAllocScratch = size_t(256) << 20;
static __device__ __align__(256) char scratch[kNoAllocScratch];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

static int scratch_for_device(void** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    return fail(PFFT_ERR_DEVICE, "device index " + std::to_string(dev) + " out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(PFFT_ERR_DEVICE,
                  std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return PFFT_OK;
}

int device_execute_noalloc(Plan& p, int dir, const void* in, void* out, hipStream_t stream) {
  if (p.work_bytes > kNoAllocScratch)
    return fail(PFFT_ERR_ARGUMENT,
                "pfft_execute_noalloc: the plan needs " + std::to_string(p.work_bytes) +
                    " bytes of intermediate storage, above the static scratch of " +
                    std::to_string(kNoAllocScratch) + " bytes; use pfft_execute");
  if (!p.dev)
    return fail(PFFT_ERR_ARGUMENT,
                "pfft_execute_noalloc: the plan is not prepared (call pfft_plan_prepare "
                "first: preparing uploads tables and tunes kernels, which allocates)");
  void* ptr = nullptr;
  const int r = scratch_for_device(&ptr);
  if (r) return r;
  return device_execute(p, dir, in, out, ptr, kNoAllocScratch, stream);
}

static bool overlaps(const void* a, size_t na, const void* b, size_t nb) {
  const char* x = (const char*)a;
  const char* y = (const char*)b;
  return x < y + nb && y < x + na;
}

int device_execute(Plan& p, int dir, const void* in, void* out, void* work, size_t work_bytes,
                   hipStream_t stream, const pfft_spectral_op* op) {
  if (dir < PFFT_FORWARD || dir > PFFT_ADJOINT_BACKWARD)
    return fail(PFFT_ERR_ARGUMENT, "unknown direction " + std::to_string(dir));
  if (!in || !out) return fail(PFFT_ERR_ARGUMENT, "null array pointer");
  const bool adjoint = dir == PFFT_ADJOINT_FORWARD || dir == PFFT_ADJOINT_BACKWARD;
  if (adjoint && !p.adjoint_ok)
    return fail(PFFT_ERR_UNSUPPORTED,
                "adjoints are provided for FFT/BFFT/RFFT/BRFFT/NoTransform plans only");
  // data flows input-shaped -> output-shaped for mul! and the adjoint of bmul!
  const bool fwd_shape = dir == PFFT_FORWARD || dir == PFFT_ADJOINT_BACKWARD;
  const size_t nin = fwd_shape ? p.in_bytes : p.out_bytes;
  const size_t nout = fwd_shape ? p.out_bytes : p.in_bytes;
  if (p.inplace) {
    // operations.jl:100-105: one ManyPencilArray is both input and output
    if (in != out)
      return fail(PFFT_ERR_ARGUMENT, "input and output arrays for in-place plan must be the same");
  } else if (overlaps(in, nin, out, nout)) {
    return fail(PFFT_ERR_ARGUMENT, "out-of-place plan applied to aliased data");
  }
  int op_pass = -1, op_side = 0;
  if (op) {
    const int r0 = op_setup(p, dir, *op, &op_pass, &op_side);
    if (r0) return r0;
  }
  int r = device_prepare(p);
  if (r) return r;
  DeviceState* d = p.dev;
  char* wk = (char*)work;
  if (p.work_bytes) {
    if (!wk) {
      if (d->work_size < p.work_bytes) {
        if (d->work) HIP_TRY(hipFree(d->work));
        d->work = nullptr;
        hipError_t e = hipMalloc(&d->work, p.work_bytes);
        if (e != hipSuccess)
          return fail(PFFT_ERR_ALLOC, std::string("hipMalloc(work): ") + hipGetErrorString(e));
        d->work_size = p.work_bytes;
      }
      wk = (char*)d->work;
    } else if (work_bytes < p.work_bytes) {
      return fail(PFFT_ERR_ARGUMENT, "workspace too small: need " + std::to_string(p.work_bytes) +
                                         " bytes, got " + std::to_string(work_bytes));
    } else if (overlaps(wk, p.work_bytes, in, nin) || overlaps(wk, p.work_bytes, out, nout)) {
      return fail(PFFT_ERR_ARGUMENT, "workspace overlaps the input or output array");
    }
  }
  // the adjoint of mul! runs the backward passes, the adjoint of bmul! the
  // forward ones, each with the Hermitian weights of its r2c/c2r stage
  const std::vector<PassPlan>& passes = fwd_shape ? p.fwd : p.bwd;
  void* bufs[3] = {nullptr, wk, wk ? wk + p.work0_bytes : nullptr};
  const int np = (int)passes.size();
  if (d->timing) HIP_TRY(hipEventRecord(d->ev[0], stream));
  for (int i = 0; i < np; ++i) {
    PassPlan opp;   // the operator-carrying form of this pass
    if (i == op_pass) {
      opp = passes[i];
      opp.args = opp.gen_args;
      opp.threads = opp.gen_threads;
      opp.lds_bytes = opp.gen_lds;
      opp.grid = opp.gen_grid;
      opp.variant = 0;
      if (passes[i].variant != 0) {
        int v = passes[i].variant_op;
        if (v < 0) {
          const std::vector<int> c = fast_candidates(p, opp, true);
          v = c.empty() ? 0 : c[0];
        }
        if (v > 0) {
          fast_apply(opp, v);
          r = fast_prepare(opp);
          if (r) return r;
        }
      }
      op_fill(p, dir, *op, op_side, opp.args);
    }
    const PassPlan& pp = i == op_pass ? opp : passes[i];
    PassArgs a = pp.args;
    a.in = pp.src == 0 ? in : bufs[pp.src];
    a.out = pp.dst == 0 ? out : bufs[pp.dst];
    bind_tables(p, pp, a);
    if (dir == PFFT_BACKWARD_SCALED && i == np - 1) {
      a.do_scale = 1;
      a.scale = 1.0 / p.scale;
    }
    a.adjoint = adjoint ? 1 : 0;
    if (pp.variant == 0) {
      launch_generic(pp, p.real_type, a, stream);
    } else {
      r = fast_launch(pp, a, stream);
      if (r) return r;
    }
    HIP_TRY(hipGetLastError());
    if (d->timing) HIP_TRY(hipEventRecord(d->ev[i + 1], stream));
  }
  return PFFT_OK;
}

}  // namespace pfft
