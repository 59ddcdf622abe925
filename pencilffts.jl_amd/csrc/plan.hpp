// Host-side plan of the transform pipeline (the PLAN layer of PencilFFTs:
// src/plans.jl, src/global_params.jl, src/Transforms/*.jl), and the pass
// descriptors the execution layer hands to the kernels.
#pragma once

#include <stdint.h>

#include <complex>
#include <string>
#include <vector>

#include "../../include/pfft.h"
#include "common.hpp"

namespace pfft {

// ---- errors ----------------------------------------------------------------
struct Error {
  int code;
  std::string msg;
};
void set_error(int code, const std::string& msg);
int fail(int code, const std::string& msg);  // set_error + return code
const char* last_error();
void clear_error();

// ---- 1D transform metadata (src/Transforms) ---------------------------------
bool kind_valid(int32_t kind, std::string* why);
bool kind_inplace(int32_t kind);
int32_t kind_base(int32_t kind);  // without the in-place bit
std::string kind_name(int32_t kind);       // Base.show of the transform
std::string kind_type_name(int32_t kind);  // fully qualified type name
std::string eltype_name(int32_t eltype);
int length_output(int32_t kind, int64_t n_in, int64_t* n_out);
int binv(int32_t kind, int64_t d, int32_t* inv);
int scale_factor_1d(int32_t kind, int64_t n_in, int64_t* scale);
int eltype_input(int32_t kind, int32_t real_type, int32_t* out);
int eltype_output(int32_t kind, int32_t eltype_in, int32_t* out);

// ---- pass (one fused kernel launch over the whole array) --------------------
struct PassPlan {
  PassArgs args;            // pointers (in/out/tw/aux) patched at launch
  int32_t elt_in, elt_out;  // element types of the pass input/output
  int64_t in_elems, out_elems;
  int32_t tw_id;            // index into Plan::tables (twiddles), -1 = none
  int32_t aux_id;           // index into Plan::tables (aux), -1 = none
  int32_t bs_chirp_id = -1; // Bluestein chirp / kernel spectrum tables, -1 = none
  int32_t bs_bhat_id = -1;
  int32_t fs_tw_id = -1;    // four-step twiddles, -1 = none
  int32_t split_part = 0;   // 1 / 2: first / second pass of a four-step split
  int32_t variant;          // kernel family chosen (0 = generic LDS pass)
  int32_t threads;
  int64_t lds_bytes;
  int64_t grid;
  int32_t src;              // 0 = user input, 1 = work0, 2 = work1
  int32_t dst;              // 0 = user output, 1 = work0, 2 = work1
  // the generic form of the pass as built (before a specialised variant was
  // applied): the starting point of the operator-carrying variants
  PassArgs gen_args;
  int32_t gen_threads;
  int64_t gen_lds, gen_grid;
  int32_t variant_op = -1;  // operator-carrying variant (-1: not tuned, 0: generic)
};

struct Table {
  int32_t real_bytes;            // 4 or 8
  std::vector<double> values;    // interleaved re, im (double precision)
  std::vector<uint8_t> host;     // rounded to the plan's precision
  void* dev = nullptr;           // device copy
};

// ---- stage of the pencil chain ----------------------------------------------
struct Stage {
  int32_t dim;          // logical dimension (0-based)
  int32_t kind_fwd, kind_bwd;
  int64_t n_in, n_out;  // length along `dim`, forward input / output
  int32_t elt_in, elt_out;
  int32_t perm[kMaxDims];  // memory order: memory axis i holds logical perm[i]
  int64_t dims_before[kMaxDims], dims_after[kMaxDims];
  int32_t mem_axis;        // memory axis of the transformed dimension
};

struct DeviceState;

struct Plan {
  int32_t N = 0;             // transformed dims
  int32_t Ne = 0;            // extra dims
  int64_t extra[kMaxDims] = {0};
  int64_t E = 1;             // product of extra dims
  int64_t size_in[kMaxDims] = {0}, size_out[kMaxDims] = {0};
  int32_t kinds[kMaxDims] = {0};
  int32_t real_type = 1;     // PFFT_F32 / PFFT_F64
  bool permute = true;
  bool inplace = false;
  int64_t pad_in0 = 0;     // in-place RFFT!: padded length 2 (n/2 + 1) of input memory axis 0
  bool adjoint_ok = true;  // no R2R stage: the adjoint directions are defined
  uint32_t flags = 0;
  double scale = 1.0;
  int32_t elt_in = 0, elt_out = 0;
  std::vector<Stage> stages;
  std::vector<PassPlan> fwd, bwd;  // pass lists per direction
  std::vector<Table> tables;
  size_t in_bytes = 0, out_bytes = 0, work_bytes = 0, work0_bytes = 0;
  DeviceState* dev = nullptr;
};

// Points a pass descriptor at the plan's tables: the device copies, or (host)
// the copies rounded to the plan precision that the emulator reads.
inline void bind_tables(const Plan& p, const PassPlan& pp, PassArgs& a, bool host = false) {
  auto at = [&](int32_t id) -> const void* {
    if (id < 0) return nullptr;
    return host ? (const void*)p.tables[id].host.data() : (const void*)p.tables[id].dev;
  };
  a.tw = at(pp.tw_id);
  a.aux = at(pp.aux_id);
  a.bs_chirp = at(pp.bs_chirp_id);
  a.bs_bhat = at(pp.bs_bhat_id);
  a.fs_tw = at(pp.fs_tw_id);
}

int plan_build(Plan& p, int32_t ndims, const int64_t* dims, const int32_t* kinds,
               int32_t real_type, int32_t permute, int32_t n_extra,
               const int64_t* extra, uint32_t flags);

// Spectral operator of pfft_execute_op: checks it against the plan and the
// direction, and names the pass that carries it (index in that direction's
// pass list) and its side (1: on the elements read, 2: on those written).
int op_setup(const Plan& p, int dir, const pfft_spectral_op& op, int* pass, int* side);
// Writes the operator into the arguments of the carrying pass.
void op_fill(const Plan& p, int dir, const pfft_spectral_op& op, int side, PassArgs& a);
std::string plan_describe(const Plan& p);

// device side (kernels.hip)
int device_prepare(Plan& p);
void device_release(Plan& p);
int device_execute(Plan& p, int dir, const void* in, void* out, void* work, size_t work_bytes,
                   hipStream_t stream, const pfft_spectral_op* op = nullptr);
// pfft_execute with the intermediates in the library's static device scratch
int device_execute_noalloc(Plan& p, int dir, const void* in, void* out, hipStream_t stream);
int device_set_timing(Plan& p, bool on);
// pfft_execute_batch: the plans of the items, in item order
int device_execute_batch(int n, Plan* const* plans, const pfft_batch_item* items,
                         hipStream_t stream);
int device_project(Plan& p, void* const* uh, const double* scale, hipStream_t stream);
int device_stage_times(Plan& p, int n, float* ms);
int device_pass_names(Plan& p, int dir, std::string* out);
// tune cache as text (pfft_tune_export / pfft_tune_import / pfft_tune_forget)
int tune_export(std::string* out);
int tune_import(const char* text);
void tune_forget();
size_t eltype_bytes(int32_t eltype);

}  // namespace pfft
