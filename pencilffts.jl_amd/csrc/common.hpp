// Shared definitions of the transform pipeline: per-line operation kinds, the
// pass descriptor handed to every kernel, and a small complex type usable on
// host and device.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define PFFT_HD __host__ __device__ __forceinline__

namespace pfft {

constexpr int kMaxDims = 8;       // transformed dims + the flattened extra dim
constexpr int kMaxStages = 24;    // radix stages of one line FFT
constexpr int kBlockThreads = 256;       // generic pass workgroup
constexpr int kBlockThreadsHeavy = 512;  // ... for lines with direct prime stages or Bluestein
constexpr int64_t kMaxLds = 160 * 1024;  // LDS bytes per workgroup (gfx950)
// Line FFTs whose length has a prime factor above this run in the Bluestein
// form (direct stages would cost that factor per output point).
constexpr int64_t kBluesteinPrime = 64;

// What one pass does to each line.  A pass transforms one logical dimension
// of the array (the reference's `mul!(parent(v), r.fftw_plan, parent(u))`,
// operations.jl:250) and writes it in the memory order of the next pencil
// (the transpose of operations.jl:239-240, fused into the store).
enum LineKind : int32_t {
  L_COPY = 0,      // NoTransform / IdentityPlan (no_transform.jl:31)
  L_C2C = 1,       // FFT / BFFT (sign from PassArgs::bwd)
  L_R2C_HALF = 2,  // RFFT, even length: n/2-point complex FFT + split
  L_R2C_FULL = 3,  // RFFT, odd length: n-point complex FFT of real data
  L_C2R_HALF = 4,  // BRFFT{even}: Hermitian merge + n/2-point complex FFT
  L_C2R_FULL = 5,  // BRFFT{odd}: Hermitian extension + n-point FFT
  L_DCT1 = 6,      // REDFT00: even extension, 2(n-1)-point FFT
  L_DCT2 = 7,      // REDFT10: Makhoul reorder, n-point FFT, post-twiddle
  L_DCT3 = 8,      // REDFT01: pre-twiddle, n-point FFT, Makhoul un-reorder
  L_DCT4 = 9,      // REDFT11: pre-twiddle, zero-padded 2n-point FFT, post-twiddle
  L_DST1 = 10,     // RODFT00: odd extension, 2(n+1)-point FFT
  L_DST2 = 11,     // RODFT10: DCT2 of (-1)^j x, reversed output
  L_DST3 = 12,     // RODFT01: (-1)^k DCT3 of the reversed input
  L_DST4 = 13,     // RODFT11: (-1)^k DCT4 of the reversed input
  L_DHT = 14,      // DHT: Re F - Im F
  // REDFT11 / RODFT11 of even length n: the n/2-point FFTs of two real lines
  // a, b (the halves z_m = (x_2m + i x_{n-1-2m}) c_m) interleaved into one
  // n-point FFT W, split afterwards as Z_m = (W_m + W_{m+n/2}) / 2 and
  // w^m (W_m - W_{m+n/2}) / 2.  The generic pass feeds one real line per work
  // line (b = 0); the fast pass two (the parts of complex data, or a pair of
  // real lines).
  L_DCT4H = 15,
  L_DST4H = 16
};

// Descriptor of one pass over the whole array (passed by value to kernels).
//
// The array is viewed as a set of lines along the transformed axis.  A line is
// named by a multi-index m over `D` line dims; dim 0 is the tile axis, along
// which `T` consecutive lines are processed by one workgroup so that the
// transposed store writes T consecutive elements.  Element j of line m sits at
//   in  + (sum_d m_d * lsin[d]  + j * sk_in)   (input elements)
//   out + (sum_d m_d * lsout[d] + k * sk_out)  (output elements)
struct PassArgs {
  const void* in;
  void* out;
  const void* tw;      // M-point twiddles exp(-2 pi i k / M), k < M
  const void* aux;     // per-kind pre/post twiddles (may be null)
  const int32_t* radix_sched;  // unused by device code; kept for layout
  int32_t kind;        // LineKind
  int32_t bwd;         // 1: backward sign (conjugated twiddles)
  int32_t in_cplx;     // input elements complex
  int32_t out_cplx;    // output elements complex
  int32_t r2r_cplx;    // R2R/COPY applied to complex data (re and im parts)
  int32_t T;           // lines per tile
  int32_t D;           // number of line dims
  int32_t nstages;     // radix stages of the M-point FFT
  int32_t radix[kMaxStages];
  int64_t L_in, L_out; // elements along the transform axis (in/out types)
  int64_t n;           // logical transform length (real length for r2c/c2r)
  int64_t M;           // FFT length per line
  int64_t sk_in, sk_out;
  int64_t ldim[kMaxDims];
  int64_t lsin[kMaxDims];
  int64_t lsout[kMaxDims];
  int64_t ntiles;      // ceil(ldim[0]/T) * prod(ldim[1..D-1])
  int32_t tile_order;  // 0: tile index runs along line dim 0 first, 1: along dims 1.. first
  int32_t tile_rev;    // 1: tiles taken in descending order
  int64_t Mp;          // LDS line pitch (complex elements) of the work buffers
  int64_t raw_pitch;   // LDS line pitch (scalars) of the raw input tile
  int64_t lds_w1_off;  // byte offset of work buffer 1 (== raw tile) in LDS
  int32_t load_w0;     // c2c, even r2c / c2r: the load writes the work lines (no raw tile)
  int32_t pair_real;   // DCT/DST I-III, DHT: two real lines per complex work line (re, im)
  double scale;        // output multiplier (1 unless ldiv! is fused)
  int32_t do_scale;
  int32_t adjoint;     // fuse the Hermitian weights of the r2c/c2r adjoints
  // Spectral operator fused into this pass (pfft_execute_op): the elements it
  // reads (op_side 1) or writes (op_side 2) are multiplied by alpha m(k), k
  // the integer wavenumbers of the spectrum element; op_acc adds the written
  // value into the output.
  int32_t op;          // PFFT_OP_* (0: none)
  int32_t op_side;
  int32_t op_dim;      // logical dim of PFFT_OP_DERIVATIVE
  int32_t op_conj;     // conj(m): the transposes
  int32_t op_acc;
  int32_t pdim;        // logical dim transformed by this pass
  int32_t kN;          // logical dims of the spectrum
  int32_t llog[kMaxDims];   // logical dim of each line dim (-1: extra dims)
  int32_t kmode[kMaxDims];  // per logical dim: 0 fftfreq, 1 rfftfreq
  int64_t kn[kMaxDims];     // per logical dim: spectrum length
  double kmax[kMaxDims];    // per logical dim: largest stored |k|
  double op_alpha;
  // Bluestein (chirp-z) form of the M-point line FFT, for lengths with a
  // large prime factor: X_k = c_k sum_j (x_j c_j) conj(c_{k-j}) with
  // c_j = exp(-i pi j^2 / M), the convolution done by two bs_M-point FFTs
  // (bs_M a power of two >= 2M - 1).  bs_M = 0: off.  With it, radix / nstages
  // / tw describe the bs_M-point FFT.
  int64_t bs_M;
  const void* bs_chirp;   // c_j, j < M
  const void* bs_bhat;    // bs_M-point FFT of b_m = conj(c_m) (m mod bs_M, |m| < M), / bs_M,
                          // in the output order of the in-place forward FFT
  // First pass of the four-step split of a line too long for the LDS: output
  // k of the line with index j2 along line dim 0 is multiplied by
  // w^(j2 k), w = exp(-2 pi i / fs_M) (conjugated for a backward
  // transform).  fs_M = 0: off.
  int64_t fs_M;
  const void* fs_tw;      // fs_M-point twiddles
};

template <typename R>
struct cpx {
  R re, im;
};

template <typename R>
PFFT_HD cpx<R> mk(R a, R b) { cpx<R> z; z.re = a; z.im = b; return z; }
template <typename R>
PFFT_HD cpx<R> operator+(cpx<R> a, cpx<R> b) { return mk(a.re + b.re, a.im + b.im); }
template <typename R>
PFFT_HD cpx<R> operator-(cpx<R> a, cpx<R> b) { return mk(a.re - b.re, a.im - b.im); }
template <typename R>
PFFT_HD cpx<R> operator*(cpx<R> a, cpx<R> b) {
  return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <typename R>
PFFT_HD cpx<R> scl(cpx<R> a, R s) { return mk(a.re * s, a.im * s); }
template <typename R>
PFFT_HD cpx<R> conj(cpx<R> a) { return mk(a.re, -a.im); }
// multiply by -i (forward) or +i (backward)
template <typename R>
PFFT_HD cpx<R> mul_mi(cpx<R> a) { return mk(a.im, -a.re); }
template <typename R>
PFFT_HD cpx<R> mul_pi(cpx<R> a) { return mk(-a.im, a.re); }

}  // namespace pfft
