/*
 * pfft.h -- C ABI of the MI355X-native PencilFFTs transform pipeline.
 *
 * One library, libpencilffts_hip.so, replaces the execution layer of
 * PencilFFTs.jl v0.15.3 for the single-GPU case: the per-dimension FFTW calls
 * (src/operations.jl:250) and the PencilArrays transposes between them
 * (src/operations.jl:239-240) become one stream-ordered sequence of fused HIP
 * passes.  Each entry point below names the reference interface it stands in
 * for.  All functions return an int32 status (PFFT_OK == 0) and never throw;
 * the message of the last failure on the calling thread is returned by
 * pfft_last_error().
 *
 * Array layout follows Julia: column-major, memory dims listed fastest first.
 * The input of a plan is laid out in logical order (x, y, z, ..., extra...),
 * the output in the memory order of the last pencil of the reference's chain
 * (reversed dims (z, y, x, extra...) for permute_dims = true), exactly as
 * PencilFFTs' allocate_output returns it (src/allocate.jl:106-110).
 */
#ifndef PFFT_H
#define PFFT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
  PFFT_OK = 0,
  PFFT_ERR_ARGUMENT = 1,    /* the reference throws ArgumentError here */
  PFFT_ERR_UNSUPPORTED = 2, /* valid in the reference, not in this build */
  PFFT_ERR_DEVICE = 3,      /* a HIP runtime call failed */
  PFFT_ERR_ALLOC = 4        /* host or device allocation failed */
};

/* ---- transform kinds (one per dimension) -------------------------------
 * Mirrors the transform types of src/Transforms/{c2c,r2c,r2r,no_transform}.jl.
 * R2R kinds are PFFT_R2R + the FFTW r2r kind code (FFTW_R2HC = 0 ... FFTW_RODFT11
 * = 10); R2HC and HC2R are rejected like the reference does (r2r.jl:40-46).
 * OR-ing PFFT_INPLACE gives the in-place variant (FFT!, RFFT!, ...).        */
enum {
  PFFT_NOTRANSFORM = 0,
  PFFT_FFT = 1,
  PFFT_BFFT = 2,
  PFFT_RFFT = 3,
  PFFT_BRFFT_EVEN = 4,
  PFFT_BRFFT_ODD = 5,
  PFFT_R2R = 16,
  PFFT_INPLACE = 0x100
};
enum {
  PFFT_R2HC = 0, PFFT_HC2R = 1, PFFT_DHT = 2,
  PFFT_REDFT00 = 3, PFFT_REDFT01 = 4, PFFT_REDFT10 = 5, PFFT_REDFT11 = 6,
  PFFT_RODFT00 = 7, PFFT_RODFT01 = 8, PFFT_RODFT10 = 9, PFFT_RODFT11 = 10
};

/* ---- element type codes ------------------------------------------------ */
enum { PFFT_F32 = 0, PFFT_F64 = 1, PFFT_C32 = 2, PFFT_C64 = 3, PFFT_NOTHING = -1 };

/* ---- execute directions (src/operations.jl:24-52) ---------------------- */
enum {
  PFFT_FORWARD = 0,          /* mul!(dst, p, src)          operations.jl:24-31 */
  PFFT_BACKWARD = 1,         /* PencilFFTs.bmul!(...)      operations.jl:34-41 */
  PFFT_BACKWARD_SCALED = 2,  /* ldiv!(dst, p, src)         operations.jl:44-52 */
  /* Adjoints (transposes for the real inner product Re sum conj(a) b) of the
   * two unnormalised maps, for reverse-mode differentiation.  No counterpart
   * in the reference; they back the ChainRulesCore rrule of INTEGRATION.md.
   * The adjoint of an r2c stage is the c2r of the spectrum divided by the
   * Hermitian weights w_k (1 at DC and, for even n, Nyquist; 2 elsewhere);
   * the adjoint of a c2r stage is w_k times the r2c.  Plans with an R2R
   * stage return PFFT_ERR_UNSUPPORTED. */
  PFFT_ADJOINT_FORWARD = 3,  /* x_bar = (mul!)^T y_bar: output-shaped -> input-shaped */
  PFFT_ADJOINT_BACKWARD = 4  /* y_bar = (bmul!)^T x_bar: input-shaped -> output-shaped */
};

/* ---- plan flags -------------------------------------------------------- */
enum {
  PFFT_FLAG_GENERIC_ONLY = 1u << 0, /* use only the generic LDS kernels */
  PFFT_FLAG_NO_AUTOTUNE = 1u << 1,  /* take the heuristic tile shapes */
  PFFT_FLAG_MFMA_SPLIT16 = 1u << 2, /* radix-16 stages on MFMA, split-FP16 */
  PFFT_FLAG_MFMA_BF16X3 = 1u << 3,  /* radix-16 stages on MFMA, BF16x3 */
  PFFT_FLAG_REVERSE_TILES = 1u << 4, /* with NO_AUTOTUNE or CANDIDATE(k): every specialised
                                        pass walks its tiles in descending order (the
                                        autotuner otherwise chooses the direction per pass) */
  PFFT_FLAG_STREAM_STORES = 1u << 5,  /* likewise: every specialised pass stores its output
                                        with the nontemporal hint (otherwise tuned per pass);
                                        no counterpart in the reference, like the other
                                        kernel-choice flags */
  PFFT_FLAG_DENSE_WORK = 1u << 6,     /* intermediate buffers keep dense rows (by default a
                                        buffer loaded as tiles by the next pass, or of rows of
                                        8 KiB and more, has its rows padded by 32 elements);
                                        results are bit-identical either way */
  PFFT_FLAG_ESTIMATE = 1u << 7        /* plan without timing anything: each pass takes the
                                        choice already in the tune cache (from tuning earlier
                                        in the process, or pfft_tune_import), else its first
                                        candidate.  The counterpart of planning with
                                        fftw_flags = FFTW.ESTIMATE (plans.jl:241-250), the
                                        reference's default; without it a plan is tuned on
                                        its first execute, as FFTW.MEASURE / PATIENT plans
                                        are timed at creation */
};
/* Instead of autotuning, run candidate k (0-based, taken modulo the number
 * of candidates) of every pass that has specialised kernels: lets a test
 * exercise each kernel variant the autotuner may pick.  Bits 16..23. */
#define PFFT_FLAG_CANDIDATE(k) ((uint32_t)(((k) + 1) & 0xff) << 16)

typedef struct pfft_plan pfft_plan;

/* ---- library ----------------------------------------------------------- */
const char* pfft_version(void);
const char* pfft_last_error(void);

/* ---- 1D transform metadata: src/Transforms/Transforms.jl public API ------
 * length_output  c2c.jl:49, r2c.jl:74-75, r2r.jl:82, no_transform.jl:23
 * binv           c2c.jl:58-61, r2c.jl:99-102, r2r.jl:103-131, no_transform.jl:21
 * scale_factor   c2c.jl:71, r2c.jl:104-111, r2r.jl:96-101, no_transform.jl:25
 *                (for one transformed dimension of length n_in; for BRFFT n_in
 *                is the complex input length)
 * eltype_input   c2c.jl:51, r2c.jl:80-81, r2r.jl:83 (PFFT_NOTHING = `nothing`)
 * eltype_output  c2c.jl:50, r2c.jl:77-78, r2r.jl:84, Transforms.jl:263-264
 * is_inplace     Transforms.jl:139-147 (tuple form: 1, 0, or -1 = `nothing`)
 * expand_dims    Transforms.jl:285-291, r2c.jl:114-124
 * transform_name Base.show of a transform (Transforms.jl:293-297, r2c.jl:64,
 *                r2r.jl:75-76)                                              */
int pfft_transform_check(int32_t kind);
int pfft_transform_length_output(int32_t kind, int64_t n_in, int64_t* n_out);
int pfft_transform_binv(int32_t kind, int64_t d, int32_t* kind_inv);
int pfft_transform_scale_factor(int32_t kind, int64_t n_in, int64_t* scale);
int pfft_transform_eltype_input(int32_t kind, int32_t real_type, int32_t* eltype);
int pfft_transform_eltype_output(int32_t kind, int32_t eltype_in, int32_t* eltype);
int pfft_transforms_is_inplace(int32_t n, const int32_t* kinds, int32_t* result);
int pfft_transform_expand_dims(int32_t kind, int32_t n, int32_t* kinds_out);
int pfft_transform_name(int32_t kind, char* buf, size_t buflen);

/* ---- plans ---------------------------------------------------------------
 * pfft_plan_create stands in for
 *   PencilFFTPlan(dims_global, transforms, proc_dims, comm, real_type;
 *                 extra_dims, permute_dims)               plans.jl:306-314
 * and runs the PLAN layer: GlobalFFTParams (global_params.jl:50-62), input
 * eltype inference (global_params.jl:81-95), the pencil/permutation chain
 * (plans.jl:397-521), the per-dimension binv and scale factors
 * (plans.jl:523-553) and their product (plans.jl:268).  proc_dims must
 * multiply to one process (one GPU).  Creating a plan touches no GPU: device
 * tables are built on the first execute (or by pfft_plan_prepare).
 *
 * In-place plans (kinds OR-ed with PFFT_INPLACE) stand in for the
 * reference's in-place PencilFFTPlan over a ManyPencilArray (allocate.jl:57-83,
 * operations.jl:175-225): one user buffer of pfft_plan_nbytes' in_bytes
 * (== out_bytes) holds the input view (the first pencil) and the output view
 * (the last pencil), both at offset 0, and pfft_execute takes it as both
 * `in` and `out`.  Accepted chains: FFT!/BFFT!/R2R!/NoTransform!, or RFFT!
 * followed by FFT!s, whose real input view sits in a parent padded to
 * 2 (n/2 + 1) along memory axis 0 (ManyPencilArrayRFFT!,
 * multiarrays_r2c.jl:26-79; see pfft_plan_parent_dims).                     */
int pfft_plan_create(pfft_plan** out, int32_t ndims, const int64_t* dims_logical,
                     const int32_t* kinds, int32_t real_type, int32_t permute_dims,
                     int32_t n_extra, const int64_t* extra_dims, uint32_t flags);
int pfft_plan_destroy(pfft_plan* plan);
int pfft_plan_prepare(pfft_plan* plan, void* hip_stream);

/* Queries (plans.jl:555-591 accessors, allocate.jl shapes). */
int pfft_plan_ndims(const pfft_plan* plan, int32_t* ndims, int32_t* n_extra);
int pfft_plan_size_global(const pfft_plan* plan, int32_t which, int64_t* dims);
int pfft_plan_mem_dims(const pfft_plan* plan, int32_t which, int64_t* dims);
/* Memory dims of the parent array of the input (0) / output (1) view: the
 * memory dims, with memory axis 0 padded to 2 (n/2 + 1) for the real input of
 * an in-place RFFT! plan (multiarrays_r2c.jl:75-86, _make_real_array).      */
int pfft_plan_parent_dims(const pfft_plan* plan, int32_t which, int64_t* dims);
int pfft_plan_permutation(const pfft_plan* plan, int32_t which, int32_t* perm);
int pfft_plan_eltypes(const pfft_plan* plan, int32_t* eltype_in, int32_t* eltype_out);
int pfft_plan_is_inplace(const pfft_plan* plan, int32_t* inplace);
int pfft_plan_scale_factor(const pfft_plan* plan, double* scale);
int pfft_plan_kinds(const pfft_plan* plan, int32_t* kinds_fwd, int32_t* kinds_bwd);
int pfft_plan_nbytes(const pfft_plan* plan, size_t* in_bytes, size_t* out_bytes,
                     size_t* work_bytes);
/* Per-stage pencil chain: memory permutation (1-based logical dims, fastest
 * first) and logical global dims of the stage's input and output pencils,
 * and the memory axis the 1D transform runs along (plans.jl:531-538).     */
int pfft_plan_stage(const pfft_plan* plan, int32_t stage, int32_t* perm_in,
                    int64_t* dims_in, int32_t* perm_out, int64_t* dims_out,
                    int32_t* mem_axis);
int pfft_plan_describe(const pfft_plan* plan, char* buf, size_t buflen);

/* ---- execution -----------------------------------------------------------
 * pfft_execute stands in for mul! / PencilFFTs.bmul! / ldiv! on a
 * PencilFFTPlan (operations.jl:24-52): forward stages 1..N, backward stages
 * N..1 with the binv transforms, and for PFFT_BACKWARD_SCALED the
 * normalisation `dst .*= inv(scale_factor(p))` (operations.jl:50,64-66)
 * fused into the last pass.  `in` and `out` are device pointers sized by
 * pfft_plan_nbytes for the direction (forward: in = input array, out = output
 * array; backward: in = output-shaped array, out = input-shaped array).
 * `work` must hold work_bytes bytes, or be NULL to use a plan-owned buffer
 * (like the reference's ibuf/obuf, plans.jl:233-234, which makes the plan
 * non-reentrant).  Overlapping in/out ranges are rejected with the
 * reference's "out-of-place plan applied to aliased data"
 * (operations.jl:91-93).  All work is enqueued on `hip_stream` (NULL = the
 * null stream); the call does not synchronise.                              */
int pfft_execute(pfft_plan* plan, int32_t direction, const void* in, void* out,
                 void* work, size_t work_bytes, void* hip_stream);

/* pfft_execute that allocates nothing.  No counterpart in the reference: it
 * is for callers that cannot call hipMalloc -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation.  The
 * intermediate pencils go to a static device array of the library
 * (PFFT_NOALLOC_SCRATCH_BYTES, 256 MiB, held on every GPU the library runs
 * on until the process ends) instead of a workspace; the caller passes the
 * input and output arrays as for pfft_execute.  The plan must be prepared
 * (pfft_plan_prepare: tables are uploaded and kernels tuned there, which
 * allocates), and its work_bytes (pfft_plan_nbytes) must not exceed the
 * scratch size; either failure returns PFFT_ERR_ARGUMENT.  Spectral
 * operators are not available here.  All calls on one GPU share its array,
 * and its contents are whatever the last call left: two calls that overlap
 * in time on one GPU (other streams or threads) race.                      */
#define PFFT_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)
int pfft_execute_noalloc(pfft_plan* plan, int32_t direction, const void* in, void* out,
                         void* hip_stream);

/* Many arrays through one plan: the reference broadcasts mul!/ldiv!/bmul!
 * over tuples or arrays of PencilArrays one at a time (operations.jl:139-152,
 * with the length check of :130-137).  Here all n_arrays transforms are
 * issued back to back on one stream with one shared workspace.              */
int pfft_execute_many(pfft_plan* plan, int32_t direction, int32_t n_arrays,
                      const void* const* ins, void* const* outs, void* work,
                      size_t work_bytes, void* hip_stream);

/* ---- spectral operators fused into the passes ------------------------------
 * A diagonal multiplier alpha m(k) of the spectrum (the output of mul!), k the
 * integer wavenumbers of each spectral element: rfftfreq along the r2c
 * dimension, fftfreq along the c2c ones (test/include/Grids.jl:44-59 with a
 * 2 pi box).  The forms are those of the reference's Fourier-space operations
 * (test/include/FourierOperations.jl:18-86, docs/examples/gradient.jl:254-273,
 * docs/examples/navier_stokes.jl:292-344):
 *   PFFT_OP_DERIVATIVE     m = i k_dim                 (one gradient component)
 *   PFFT_OP_LAPLACIAN      m = -|k|^2
 *   PFFT_OP_INV_LAPLACIAN  m = -1/|k|^2, 0 at k = 0    (Poisson solve)
 *   PFFT_OP_DEALIAS        m = 0 where any |k_d| > (2/3) max|k_d|, else 1
 * pfft_execute_op fuses it into the pass that writes the spectrum
 *   PFFT_FORWARD            out = alpha m F(in)
 *   PFFT_ADJOINT_BACKWARD   out = alpha conj(m) B^T(in)
 * or reads it
 *   PFFT_BACKWARD[_SCALED]  out = alpha B(m in) [/ N]
 *   PFFT_ADJOINT_FORWARD    out = alpha F^T(conj(m) in)
 * so an operator costs no memory pass of its own.  On the writing side
 * accumulate = 1 adds into `out` instead of overwriting it (sums over
 * components, as in a divergence or a gradient's adjoint).  No counterpart in
 * the reference's API: it applies these operators in separate Julia loops
 * between transforms.  Plans need FFT/BFFT/RFFT dimensions, and the carrying
 * pass must be a c2c one (two or more transformed dimensions) whose lines
 * fit one LDS-resident pass (not split in two), else PFFT_ERR_UNSUPPORTED.  */
enum {
  PFFT_OP_NONE = 0,
  PFFT_OP_DERIVATIVE = 1,
  PFFT_OP_LAPLACIAN = 2,
  PFFT_OP_INV_LAPLACIAN = 3,
  PFFT_OP_DEALIAS = 4
};
typedef struct {
  int32_t op;          /* PFFT_OP_* */
  int32_t dim;         /* 0-based logical dimension of PFFT_OP_DERIVATIVE */
  double alpha;        /* scalar factor */
  int32_t accumulate;  /* writing side: out += result */
} pfft_spectral_op;
int pfft_execute_op(pfft_plan* plan, int32_t direction, const pfft_spectral_op* op,
                    const void* in, void* out, void* work, size_t work_bytes,
                    void* hip_stream);

/* ---- batches of plans of different shapes ----------------------------------
 * Transforms of several plans (different shapes, kinds or precisions) in one
 * call: the collection broadcast of operations.jl:139-152 for fields that do
 * not share a plan (the reference runs them one after the other).  The
 * result is that of pfft_execute_op (pfft_execute when op is NULL) for each
 * item in order on `hip_stream`, except that items of different plans may
 * run at the same time: they are enqueued on up to four library streams
 * forked from `hip_stream` by events and joined back into it before the call
 * returns, so the call may be captured in a hipGraph (the library's streams
 * and events are created on a device's first batch, which must then not be
 * under capture).  Items of one plan run in their order on one of those
 * streams, since they share the plan's workspace when `work` is NULL.  No
 * item may write memory that another item reads or writes.  Batches on one
 * device from several host threads at once are serialised by a lock.  An
 * item that fails stops the enqueueing (the items before it stay enqueued)
 * and its status is returned.                                              */
typedef struct {
  pfft_plan* plan;
  int32_t direction;              /* PFFT_FORWARD ... PFFT_ADJOINT_BACKWARD */
  const pfft_spectral_op* op;     /* NULL: no operator */
  const void* in;
  void* out;
  void* work;                     /* NULL: the plan's own workspace */
  size_t work_bytes;
} pfft_batch_item;
int pfft_execute_batch(int32_t n_items, const pfft_batch_item* items, void* hip_stream);

/* Divergence-free projection of the spectrum of a 3-component vector field,
 * in place: u_i <- u_i - k_i (k . u) / |k|^2, with the mean mode (k = 0) left
 * as is -- project_divergence_free! of the reference's Navier-Stokes example
 * (docs/examples/navier_stokes.jl:351-366).  k are the plan's wavenumbers
 * (those of the spectral operators) times scale[d] = 2 pi / L_d (scale NULL:
 * a 2 pi box).  uh_x, uh_y, uh_z are output-shaped arrays (extra dims
 * included) of a 3D plan of FFT/BFFT/RFFT dimensions.  It mixes the
 * components at each wavenumber, so it is one elementwise kernel, not an
 * operator carried by a transform pass.                                    */
int pfft_project_divergence_free(pfft_plan* plan, void* uh_x, void* uh_y, void* uh_z,
                                 const double* scale, void* hip_stream);

/* ---- tune data ("wisdom") -------------------------------------------------
 * The autotuner's choices of this process as text, for a later process to
 * plan the same shapes without timing anything: the counterpart of FFTW's
 * wisdom (FFTW.export_wisdom / import_wisdom / forget_wisdom of FFTW.jl),
 * which the reference's FFTW plans (plans.jl:523-553, built with fftw_flags,
 * plans.jl:241-250) draw on.  pfft_tune_export writes the text into buf
 * (NUL-terminated) when buf is non-NULL and large enough, and its size with
 * the NUL into *needed when needed is non-NULL.  pfft_tune_import merges
 * text exported by the same build on the same kind of device, and refuses
 * any other with PFFT_ERR_ARGUMENT.  Imported choices apply to plans
 * created afterwards, whatever their flags (PFFT_FLAG_NO_AUTOTUNE and
 * PFFT_FLAG_CANDIDATE(k) aside), so a plan whose every pass is covered
 * times nothing.  pfft_tune_forget empties the cache.                      */
int pfft_tune_export(char* buf, size_t buflen, size_t* needed);
int pfft_tune_import(const char* text);
int pfft_tune_forget(void);

/* Kernel timing of the last execute on a plan with profiling enabled: the
 * mirror of the reference's TimerOutputs sections "FFT" per stage
 * (operations.jl:250).  Enabled with pfft_plan_set_timing(plan, 1).        */
int pfft_plan_set_timing(pfft_plan* plan, int32_t enable);
/* Kernel chosen for each pass of a direction, ';'-separated in launch order
 * (prepares the plan on the current device). */
int pfft_plan_kernels(pfft_plan* plan, int32_t direction, char* buf, size_t buflen);
int pfft_plan_stage_times(pfft_plan* plan, int32_t n, float* ms);

#ifdef __cplusplus
}
#endif

#endif /* PFFT_H */
