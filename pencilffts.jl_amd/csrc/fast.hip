// Specialised passes for power-of-two line lengths (64 ... 1024 complex
// points per FFT): the hot path of the pipeline.
//
// Work split: one line is owned by L lanes of one wavefront (L = M / P), each
// lane holding P points in registers in the strided order x[l + L p].  The
// Stockham stages run radix-8/16/4 butterflies in registers; between stages
// the line goes through a padded LDS region private to the line, ordered by
// wavefront-scope fences (no workgroup barrier).  The first stage takes its
// inputs straight from global memory (each load instruction of a wave reads
// consecutive elements) and the last stage leaves its outputs in registers in
// the same strided order, so a contiguous line store is coalesced as well.
//
// A workgroup owns T lines.  When the output (or input) of the pass is
// transposed -- the fused PencilArrays transpose of the reference -- the T
// lines meet in an LDS tile and are written (read) as runs of T consecutive
// elements with 16-byte accesses.
//
// Kinds: complex FFT (the backward sign as conj(FFT(conj x))), r2c and c2r of
// an even real length n via the n/2-point complex FFT with the split/merge
// twiddle w_n^k fused in.
//
// The file is compiled twice.  As itself it holds the kernels and the host
// side (candidates, launch, names); as fast_nt.hip (PFFT_FAST_NT = 1) it
// holds the same kernels with every output store carrying the nontemporal
// hint and exports only that kernel table, entry for entry in the same
// order.  The hint is a compile-time choice: a run-time branch around each
// store changes register allocation and scheduling of every kernel.
#ifndef PFFT_FAST_NT
#define PFFT_FAST_NT 0
#endif
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <string>

#include "../../include/pfft.h"
#include "fast_pass.hpp"
#include "generic_pass.hpp"

namespace pfft {
namespace fk {

template <typename R>
using C = cpx<R>;

// one padding complex every 8 spreads the radix-8 scatter over the banks
__device__ __forceinline__ constexpr int pad8(int q) { return q + (q >> 3); }
// pad8(q + off) for a compile-time offset that is a multiple of 8: the pad
// is additive, so a lane keeps one base address pad8(q) for all its accesses
__device__ __forceinline__ constexpr int pad8_off(int off) { return off + (off >> 3); }
template <int M>
constexpr int line_pitch() { return M + (M >> 3) + 2; }  // holds indices 0..M

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Orders the LDS accesses of one line: a wavefront fence when the line's L
// lanes are one wavefront, a workgroup barrier when they span several
// (every thread of the workgroup reaches every line_sync).
template <int L>
__device__ __forceinline__ void line_sync() {
  if constexpr (L <= 64) wave_sync();
  else __syncthreads();
}

// ---- forward small DFTs in registers ----------------------------------------
template <typename R>
__device__ __forceinline__ void dft16(C<R>* v) {
  // 16 = 4 x 4: n = 4a + b, k = k1 + 4 k2
  const R c1 = R(0.92387953251128675612818318939679);  // cos(pi/8)
  const R s1 = R(0.38268343236508977172845998403040);  // sin(pi/8)
  const R h = R(0.70710678118654752440084436210485);
  C<R> y[16];
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    C<R> t[4] = {v[b], v[4 + b], v[8 + b], v[12 + b]};
    dft4(t);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) y[4 * b + k1] = t[k1];
  }
  // twiddles w16^(b k1)
  y[5] = mk(y[5].re * c1 + y[5].im * s1, y[5].im * c1 - y[5].re * s1);      // w^1
  y[6] = mk((y[6].re + y[6].im) * h, (y[6].im - y[6].re) * h);              // w^2
  y[7] = mk(y[7].re * s1 + y[7].im * c1, y[7].im * s1 - y[7].re * c1);      // w^3
  y[9] = mk((y[9].re + y[9].im) * h, (y[9].im - y[9].re) * h);              // w^2
  y[10] = mul_mi(y[10]);                                                    // w^4
  y[11] = mk((y[11].im - y[11].re) * h, -(y[11].re + y[11].im) * h);        // w^6
  y[13] = mk(y[13].re * s1 + y[13].im * c1, y[13].im * s1 - y[13].re * c1); // w^3
  y[14] = mk((y[14].im - y[14].re) * h, -(y[14].re + y[14].im) * h);        // w^6
  y[15] = mk(-y[15].re * c1 - y[15].im * s1, -y[15].im * c1 + y[15].re * s1); // w^9
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    C<R> t[4] = {y[k1], y[4 + k1], y[8 + k1], y[12 + k1]};
    dft4(t);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = t[k2];
  }
}

// Good-Thomas prime-factor DFT of N1 N2 points, N1 and N2 coprime: with the
// index maps n = (N2 n1 + N1 n2) mod N and k = (N2 <1/N2 mod N1> k1 + N1 <1/N1
// mod N2> k2) mod N it is N1 DFTs of N2 points, then N2 DFTs of N1 points,
// with no twiddles; the maps are compile-time register renamings.
constexpr int inv_mod(int a, int m) {
  for (int x = 1; x < m; ++x)
    if ((a * x) % m == 1) return x;
  return 1;
}
template <typename R, int RAD>
__device__ __forceinline__ void dft(C<R>* v);
template <typename R, int N1, int N2>
__device__ __forceinline__ void dft_pfa(C<R>* v) {
  constexpr int N = N1 * N2;
  constexpr int E1 = N2 * inv_mod(N2 % N1, N1), E2 = N1 * inv_mod(N1 % N2, N2);
  C<R> t[N1][N2];
#pragma unroll
  for (int n1 = 0; n1 < N1; ++n1)
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) t[n1][n2] = v[(N2 * n1 + N1 * n2) % N];
#pragma unroll
  for (int n1 = 0; n1 < N1; ++n1) dft<R, N2>(t[n1]);
#pragma unroll
  for (int k2 = 0; k2 < N2; ++k2) {
    C<R> u[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) u[n1] = t[n1][k2];
    dft<R, N1>(u);
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) v[(E1 * k1 + E2 * k2) % N] = u[k1];
  }
}

template <typename R, int RAD>
__device__ __forceinline__ void dft(C<R>* v) {
  if constexpr (RAD == 2) dft2(v);
  else if constexpr (RAD == 3) dft3(v);
  else if constexpr (RAD == 4) dft4(v);
  else if constexpr (RAD == 5) dft5(v);
  else if constexpr (RAD == 8) dft8(v);
  else if constexpr (RAD == 16) dft16(v);
  else if constexpr (RAD == 6) dft_pfa<R, 2, 3>(v);
  else if constexpr (RAD == 10) dft_pfa<R, 2, 5>(v);
  else if constexpr (RAD == 12) dft_pfa<R, 4, 3>(v);
  else if constexpr (RAD == 20) dft_pfa<R, 4, 5>(v);
  else static_assert(RAD == 0, "no register DFT of this radix");
}

// ---- one Stockham stage over one line ------------------------------------------
// Butterfly j (0 <= j < M/RAD) of a stage with Ns = product of earlier radices:
//   inputs  x[j + (M/RAD) r] * w_M^(r (j mod Ns) M/(Ns RAD))
//   outputs y[(j / Ns) Ns RAD + (j mod Ns) + r Ns]
// Lane l owns butterflies j = l + L i.  With FROM_REG the inputs come from
// the strided registers v (index i + BPL r); with TO_REG (last stage, Ns RAD
// == M) the outputs land in v in the same strided order.
template <typename R, int M, int L, int P, int RAD, int NS, bool FROM_REG, bool TO_REG,
          bool BFLY_REG = false>
__device__ __forceinline__ void stage(C<R>* v, C<R>* line, int l, const C<R>* __restrict__ tw) {
  constexpr int NB = M / RAD;
  constexpr int BPL = NB / L;
  static_assert(BPL * RAD == P, "radix must divide the points per lane");
  // TO_REG leaves output r of the lane's butterfly i in register i + BPL r:
  // the natural strided order after the last stage; BFLY_REG asks for that
  // butterfly order after an inner stage (a cross-lane exchange follows)
  static_assert(!TO_REG || BFLY_REG || NS * RAD == M,
                "registers hold the output of the last stage only");
  C<R> u[P];
  const C<R>* lp = line + pad8(l);
#pragma unroll
  for (int i = 0; i < BPL; ++i)
#pragma unroll
    for (int r = 0; r < RAD; ++r) {
      if constexpr (FROM_REG) u[i * RAD + r] = v[i + BPL * r];
      else if constexpr (L % 8 == 0 && NB % 8 == 0) u[i * RAD + r] = lp[pad8_off(L * i + NB * r)];
      else u[i * RAD + r] = line[pad8(l + L * i + NB * r)];
    }
#pragma unroll
  for (int i = 0; i < BPL; ++i) {
    C<R>* w = u + i * RAD;
    if constexpr (NS > 1) {
      // one table load per butterfly: w^1 = w_M^(jj M/(Ns RAD)); the other
      // powers by a product tree at most three multiplies deep
      const int jj = (l + L * i) % NS;
      const C<R> w1 = tw[jj * (M / (NS * RAD))];
      C<R> pw[RAD];
      pw[1] = w1;
#pragma unroll
      for (int r = 2; r < RAD; ++r) pw[r] = pw[r / 2] * pw[r - r / 2];
#pragma unroll
      for (int r = 1; r < RAD; ++r) w[r] = w[r] * pw[r];
    }
    dft<R, RAD>(w);
  }
  // several wavefronts share the line: all of them have read it before any
  // overwrites it
  if constexpr (L > 64 && !FROM_REG && !TO_REG) __syncthreads();
#pragma unroll
  for (int i = 0; i < BPL; ++i) {
    if constexpr (TO_REG) {
#pragma unroll
      for (int r = 0; r < RAD; ++r) v[i + BPL * r] = u[i * RAD + r];
    } else {
      const int j = l + L * i;
      const int base = (j / NS) * NS * RAD + (j % NS);
#pragma unroll
      for (int r = 0; r < RAD; ++r) line[pad8(base + r * NS)] = u[i * RAD + r];
    }
  }
}

template <typename R, int M, int L, int P, int NS, bool FIRST, int RAD, int... REST>
__device__ __forceinline__ void fft_rec(C<R>* v, C<R>* line, int l, const C<R>* tw) {
  constexpr bool LAST = sizeof...(REST) == 0;
  stage<R, M, L, P, RAD, NS, FIRST, LAST>(v, line, l, tw);
  if constexpr (!LAST) {
    line_sync<L>();
    fft_rec<R, M, L, P, NS * RAD, false, REST...>(v, line, l, tw);
  }
}

// ---- 512-point lines with the last exchange across lanes (MF = 4) -------------
// The Stockham form moves data between stages through the line's LDS slot.
// Between stages 2 and 3 of 512 = 8 . 8 . 8 (one wavefront per line, 8
// points per lane) that exchange is a pure digit transpose: point
// 64 a + 8 b + c sits in lane 8 a + c, register b after stage 2 and is needed
// in lane 8 b + c, register a by stage 3.  Swapping lane bit 3 + k with
// register bit k for k = 0, 1, 2 does it in registers: bit 3 by a DPP row
// rotation by 8 (lane l ^ 8 within a row of 16), bits 4 and 5 by gfx950's
// v_permlane16_swap / v_permlane32_swap, which exchange a register pair's
// odd rows (upper half) with the other's even rows (lower half) directly.
__device__ __forceinline__ void xl_pair_row8(float& r0, float& r1, bool hi) {
  // lane bit 3 set: keeps r1, sends r0, receives the partner's r1 into r0
  const float send = hi ? r0 : r1;
  const float recv = __int_as_float(
      __builtin_amdgcn_update_dpp(0, __float_as_int(send), 0x128 /* row_ror:8 */, 0xf, 0xf, false));
  if (hi) r0 = recv;
  else r1 = recv;
}
__device__ __forceinline__ void xl_pair_swap16(float& r0, float& r1) {
  const auto s = __builtin_amdgcn_permlane16_swap(__float_as_uint(r0), __float_as_uint(r1),
                                                  false, false);
  r0 = __uint_as_float(s[0]);
  r1 = __uint_as_float(s[1]);
}
__device__ __forceinline__ void xl_pair_swap32(float& r0, float& r1) {
  const auto s = __builtin_amdgcn_permlane32_swap(__float_as_uint(r0), __float_as_uint(r1),
                                                  false, false);
  r0 = __uint_as_float(s[0]);
  r1 = __uint_as_float(s[1]);
}
__device__ __forceinline__ void xl_transpose8(C<float>* v, int l) {
  const bool hi = (l >> 3) & 1;
#pragma unroll
  for (int r = 0; r < 8; r += 2) {
    xl_pair_row8(v[r].re, v[r + 1].re, hi);
    xl_pair_row8(v[r].im, v[r + 1].im, hi);
  }
#pragma unroll
  for (int r : {0, 1, 4, 5}) {
    xl_pair_swap16(v[r].re, v[r + 2].re);
    xl_pair_swap16(v[r].im, v[r + 2].im);
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    xl_pair_swap32(v[r].re, v[r + 4].re);
    xl_pair_swap32(v[r].im, v[r + 4].im);
  }
}
// v: x[l + 64 p] in register p on entry, X[l + 64 p] on exit (as Cfg<512>)
__device__ __forceinline__ void fft512_xl(C<float>* v, C<float>* line, int l,
                                          const C<float>* __restrict__ tw) {
  stage<float, 512, 64, 8, 8, 1, true, false>(v, line, l, tw);
  wave_sync();
  stage<float, 512, 64, 8, 8, 8, false, true, true>(v, line, l, tw);
  xl_transpose8(v, l);
  stage<float, 512, 64, 8, 8, 64, true, true>(v, line, l, tw);
}
// 256 = 8 . 8 . 4 with 32 lanes per line (two lines per wavefront): after
// stage 2, point 64 a + 8 b + c (a < 4) sits in lane 8 a + c, register b;
// stage 3 (radix 4, two butterflies per lane) wants it in lane
// 8 (b mod 4) + c, register (b div 4) + 2 a.  Lane bits 3 and 4 swap with
// register bits 0 and 1 (DPP rotation by 8, v_permlane16_swap, both within
// the line's 32 lanes), then a compile-time renaming of the registers
// (a0 a1 b2 -> b2 a0 a1) finishes it.
__device__ __forceinline__ void fft256_xl(C<float>* v, C<float>* line, int l,
                                          const C<float>* __restrict__ tw) {
  stage<float, 256, 32, 8, 8, 1, true, false>(v, line, l, tw);
  wave_sync();
  stage<float, 256, 32, 8, 8, 8, false, true, true>(v, line, l, tw);
  const bool hi = (l >> 3) & 1;
#pragma unroll
  for (int r = 0; r < 8; r += 2) {
    xl_pair_row8(v[r].re, v[r + 1].re, hi);
    xl_pair_row8(v[r].im, v[r + 1].im, hi);
  }
#pragma unroll
  for (int r : {0, 1, 4, 5}) {
    xl_pair_swap16(v[r].re, v[r + 2].re);
    xl_pair_swap16(v[r].im, v[r + 2].im);
  }
  C<float> u[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) u[((r << 1) & 6) | (r >> 2)] = v[r];
#pragma unroll
  for (int r = 0; r < 8; ++r) v[r] = u[r];
  stage<float, 256, 32, 8, 4, 64, true, true>(v, line, l, tw);
}
// 1024 = 16 . 16 . 4, one wavefront per line, 16 points per lane: after
// stage 2, point 256 a + 16 b + c (a < 4, c < 16) sits in lane 16 a + c,
// register b; stage 3 (radix 4, four butterflies per lane) wants it in lane
// 16 (b mod 4) + c, register (b div 4) + 4 a.  Lane bits 4 and 5 swap with
// register bits 0 and 1 by the two row-swap instructions alone (no DPP,
// no selects), then the registers are renamed (a0 a1 b2 b3 -> b2 b3 a0 a1).
__device__ __forceinline__ void fft1024_xl(C<float>* v, C<float>* line, int l,
                                           const C<float>* __restrict__ tw) {
  stage<float, 1024, 64, 16, 16, 1, true, false>(v, line, l, tw);
  wave_sync();
  stage<float, 1024, 64, 16, 16, 16, false, true, true>(v, line, l, tw);
#pragma unroll
  for (int r = 0; r < 16; r += 2) {   // lane bit 4 <-> register bit 0
    xl_pair_swap16(v[r].re, v[r + 1].re);
    xl_pair_swap16(v[r].im, v[r + 1].im);
  }
#pragma unroll
  for (int r = 0; r < 16; ++r)        // lane bit 5 <-> register bit 1
    if ((r & 2) == 0) {
      xl_pair_swap32(v[r].re, v[r + 2].re);
      xl_pair_swap32(v[r].im, v[r + 2].im);
    }
  C<float> u[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) u[((r & 3) << 2) | (r >> 2)] = v[r];
#pragma unroll
  for (int r = 0; r < 16; ++r) v[r] = u[r];
  stage<float, 1024, 64, 16, 4, 256, true, true>(v, line, l, tw);
}

// radix schedules: M = L * P, every radix divides P; W wavefronts per line
template <int M, int W = 1> struct Cfg;
template <> struct Cfg<64> {
  static constexpr int L = 8, P = 8;
  template <typename R>
  __device__ static void fft(C<R>* v, C<R>* line, int l, const C<R>* tw) {
    fft_rec<R, 64, L, P, 1, true, 8, 8>(v, line, l, tw);
  }
};
template <> struct Cfg<128> {
  static constexpr int L = 16, P = 8;
  template <typename R>
  __device__ static void fft(C<R>* v, C<R>* line, int l, const C<R>* tw) {
    fft_rec<R, 128, L, P, 1, true, 8, 4, 4>(v, line, l, tw);
  }
};
template <> struct Cfg<256> {
  static constexpr int L = 32, P = 8;
  template <typename R>
  __device__ static void fft(C<R>* v, C<R>* line, int l, const C<R>* tw) {
    fft_rec<R, 256, L, P, 1, true, 8, 8, 4>(v, line, l, tw);
  }
};
template <> struct Cfg<512> {
  static constexpr int L = 64, P = 8;
  template <typename R>
  __device__ static void fft(C<R>* v, C<R>* line, int l, const C<R>* tw) {
    fft_rec<R, 512, L, P, 1, true, 8, 8, 8>(v, line, l, tw);
  }
};
template <> struct Cfg<1024> {
  static constexpr int L = 64, P = 16;
  template <typename R>
  __device__ static void fft(C<R>* v, C<R>* line, int l, const C<R>* tw) {
    fft_rec<R, 1024, L, P, 1, true, 16, 16, 4>(v, line, l, tw);
  }
};
// two wavefronts per line: half the registers per lane, stages synchronised
// by workgroup barriers.  (Twice the lanes on 128- and 256-point lines, at 4
// points per lane with radix-4 stages, measured 5-25 % slower than W = 1:
// one more LDS round trip per line outweighs the extra waves.)
template <> struct Cfg<2048> {
  static constexpr int L = 64, P = 32;
  template <typename R>
  __device__ static void fft(C<R>* v, C<R>* line, int l, const C<R>* tw) {
    fft_rec<R, 2048, L, P, 1, true, 16, 16, 8>(v, line, l, tw);
  }
};
template <> struct Cfg<2048, 2> {
  static constexpr int L = 128, P = 16;
  template <typename R>
  __device__ static void fft(C<R>* v, C<R>* line, int l, const C<R>* tw) {
    fft_rec<R, 2048, L, P, 1, true, 16, 16, 8>(v, line, l, tw);
  }
};
template <> struct Cfg<1024, 2> {
  static constexpr int L = 128, P = 8;
  template <typename R>
  __device__ static void fft(C<R>* v, C<R>* line, int l, const C<R>* tw) {
    fft_rec<R, 1024, L, P, 1, true, 8, 8, 4, 4>(v, line, l, tw);
  }
};

// ---- radix-16 stages on MFMA (1024-point lines) ------------------------------------
// A radix-16 butterfly is the 16x16 DFT matrix times the twiddled inputs; in
// real form a 32x32 matrix G times [Re x; Im x].  One wavefront owns a line's
// 64 butterflies as two 32-column blocks of v_mfma_f32_32x32x16 (gfx950; two
// K-chunks of 16: the real parts, then the imaginary parts).  Lane (j = l % 32,
// h = l / 32) feeds points 8h..8h+7 of butterfly j (its B-operand slice) and
// receives points {4h..4h+3, 8+4h..11+4h} (its accumulator rows); the stage
// goes LDS to LDS, so the two sets need not coincide.
//   MF = 1  split-FP16: x = x_hi + x_lo and G = G_hi + G_lo in FP16,
//           G_hi x_hi + G_hi x_lo + G_lo x_hi accumulated in FP32.  The line
//           is scaled once, on entry, by a power of two (exact) so that
//           |x| < 2^10: the first stage's outputs stay below 2^14 and both
//           stages' FP16 operands in range; the scale leaves with the
//           output scaling of the pass.
//   MF = 2  BF16x3: three-term BF16 splits, the six products whose order is at
//           most the third term, accumulated in FP32 (BF16 has the FP32
//           exponent range: no scaling).
// The data splits round the magnitude to the top 11 (FP16) or 8 (BF16)
// significant bits with integer arithmetic on the bit pattern (nearest,
// ties away), so hi converts exactly and the remainder x - hi is exact in
// FP32.
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

static __constant__ float kCos16[16] = {
    1.0f, 0.92387953251128675613f, 0.70710678118654752440f, 0.38268343236508977173f,
    0.0f, -0.38268343236508977173f, -0.70710678118654752440f, -0.92387953251128675613f,
    -1.0f, -0.92387953251128675613f, -0.70710678118654752440f, -0.38268343236508977173f,
    0.0f, 0.38268343236508977173f, 0.70710678118654752440f, 0.92387953251128675613f};
static __constant__ float kSin16[16] = {
    0.0f, 0.38268343236508977173f, 0.70710678118654752440f, 0.92387953251128675613f,
    1.0f, 0.92387953251128675613f, 0.70710678118654752440f, 0.38268343236508977173f,
    0.0f, -0.38268343236508977173f, -0.70710678118654752440f, -0.92387953251128675613f,
    -1.0f, -0.92387953251128675613f, -0.70710678118654752440f, -0.38268343236508977173f};

// G[i][s]: rows (Re y_q, Im y_q), columns (Re x_r, Im x_r), y = F x with
// F_qr = exp(-2 pi i q r / 16) = c - i sn:  Re y = c Re x + sn Im x,
// Im y = -sn Re x + c Im x.
__device__ __forceinline__ float g16(int i, int s) {
  const int m = ((i & 15) * (s & 15)) & 15;
  const float c = kCos16[m], sn = kSin16[m];
  if (i < 16) return s < 16 ? c : sn;
  return s < 16 ? -sn : c;
}

__device__ __forceinline__ short f2bf(float x) {   // round to nearest even
  unsigned u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}
__device__ __forceinline__ float bf2f(short b) {
  return __uint_as_float(((unsigned)(unsigned short)b) << 16);
}

template <int MF>
struct MfmaA {   // this lane's A operands, two K-chunks
  f16x8 hi[2], lo[2];
  bf16x8 b1[2], b2[2], b3[2];
  __device__ void init(int l) {
    const int i = l & 31, h = l >> 5;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      s16x8 t1, t2, t3;
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const float g = g16(i, 16 * c + 8 * h + t);
        if constexpr (MF == 1) {
          const _Float16 gh = (_Float16)g;
          hi[c][t] = gh;
          lo[c][t] = (_Float16)(g - (float)gh);
        } else {
          const short a1 = f2bf(g);
          const float r1 = g - bf2f(a1);
          const short a2 = f2bf(r1);
          t1[t] = a1;
          t2[t] = a2;
          t3[t] = f2bf(r1 - bf2f(a2));
        }
      }
      if constexpr (MF == 2) {
        b1[c] = __builtin_bit_cast(bf16x8, t1);
        b2[c] = __builtin_bit_cast(bf16x8, t2);
        b3[c] = __builtin_bit_cast(bf16x8, t3);
      }
    }
  }
};

__device__ __forceinline__ float wave_max(float x) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) x = fmaxf(x, __shfl_xor(x, o, 64));
  return x;
}

// G [Re x; Im x] for one 32-butterfly block; x[t] is point 8h + t of this
// lane's butterfly.  K-chunk c = 0 holds the real parts, c = 1 the imaginary
// parts.
template <int MF>
__device__ __forceinline__ f32x16 mfma_block(const MfmaA<MF>& A, const C<float> (&x)[8]) {
  f32x16 d = {};
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    if constexpr (MF == 1) {
      f16x8 bh, bl;
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const float b = c == 0 ? x[t].re : x[t].im;
        const float hf = __uint_as_float((__float_as_uint(b) + 0x1000u) & 0xffffe000u);
        bh[t] = (_Float16)hf;
        bl[t] = (_Float16)(b - hf);
      }
      d = __builtin_amdgcn_mfma_f32_32x32x16_f16(A.hi[c], bh, d, 0, 0, 0);
      d = __builtin_amdgcn_mfma_f32_32x32x16_f16(A.hi[c], bl, d, 0, 0, 0);
      d = __builtin_amdgcn_mfma_f32_32x32x16_f16(A.lo[c], bh, d, 0, 0, 0);
    } else {
      s16x8 t1, t2, t3;
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const float b = c == 0 ? x[t].re : x[t].im;
        const unsigned u1 = (__float_as_uint(b) + 0x8000u) & 0xffff0000u;
        const float r1 = b - __uint_as_float(u1);
        const unsigned u2 = (__float_as_uint(r1) + 0x8000u) & 0xffff0000u;
        const float r2 = r1 - __uint_as_float(u2);
        t1[t] = (short)(u1 >> 16);
        t2[t] = (short)(u2 >> 16);
        t3[t] = (short)((__float_as_uint(r2) + 0x8000u) >> 16);
      }
      const bf16x8 x1 = __builtin_bit_cast(bf16x8, t1);
      const bf16x8 x2 = __builtin_bit_cast(bf16x8, t2);
      const bf16x8 x3 = __builtin_bit_cast(bf16x8, t3);
      d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A.b1[c], x1, d, 0, 0, 0);
      d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A.b1[c], x2, d, 0, 0, 0);
      d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A.b2[c], x1, d, 0, 0, 0);
      d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A.b1[c], x3, d, 0, 0, 0);
      d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A.b3[c], x1, d, 0, 0, 0);
      d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A.b2[c], x2, d, 0, 0, 0);
    }
  }
  return d;
}

// One radix-16 Stockham stage of a 1024-point line on MFMA, LDS to LDS.
// With one wavefront per line (L = 64) the wave multiplies both 32-butterfly
// blocks; with two (L = 128) wavefront w multiplies block w.
// Addresses in closed form (butterfly b = 32 blk + jj, point r = 8h + t):
//   read   pad8(b + 64 r)            = pad8(jj) + 576 h + 36 blk + 72 t
//   write  NS = 1:  pad8(16 b + q)   = 18 jj + 4 h + 576 blk + 9 g + t'
//          NS = 16: pad8(256 (b / 16) + b % 16 + 16 q)
//                   = 288 (jj / 16) + pad8(jj % 16) + 72 h + 576 blk + 144 g + 18 t'
// with output point q = 4h + t' + 8g (t' = t % 4, g = t / 4), so each lane
// keeps one base address and every access takes a constant offset.
template <int MF, int NS, int L>
__device__ __forceinline__ void stage16_mfma(C<float>* line, int l, const C<float>* __restrict__ tw,
                                             const MfmaA<MF>& A) {
  static_assert(NS == 1 || NS == 16, "the two radix-16 stages of 1024 = 16 x 16 x 4");
  static_assert(L == 64 || L == 128, "one or two wavefronts per line");
  constexpr int NBLK = L == 64 ? 2 : 1;   // blocks per wavefront
  const int jj = l & 31, h = (l >> 5) & 1, w = L == 64 ? 0 : (l >> 6);
  const C<float>* rd = line + pad8(jj) + 576 * h + 36 * w;
  C<float> x[NBLK][8];
#pragma unroll
  for (int k = 0; k < NBLK; ++k)
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      C<float> v = rd[36 * k + 72 * t];
      if constexpr (NS > 1) v = v * tw[(8 * h + t) * (jj & 15) * 4];
      x[k][t] = v;
    }
  f32x16 d[NBLK];
#pragma unroll
  for (int k = 0; k < NBLK; ++k) d[k] = mfma_block<MF>(A, x[k]);
  // in place: one wavefront's LDS reads precede its writes in program order
  // and complete in order; a second wavefront on the line waits at a barrier
  if constexpr (L > 64) __syncthreads();
  C<float>* wr = NS == 1 ? line + 18 * jj + 4 * h + 576 * w
                         : line + 288 * (jj >> 4) + pad8(jj & 15) + 72 * h + 576 * w;
#pragma unroll
  for (int k = 0; k < NBLK; ++k)
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int off = NS == 1 ? 576 * k + 9 * (t >> 2) + (t & 3)
                              : 576 * k + 144 * (t >> 2) + 18 * (t & 3);
      wr[off] = mk(d[k][t], d[k][8 + t]);
    }
}

// 1024 = 16 (MFMA) x 16 (MFMA) x 4 (registers); P = 1024 / L points per lane.
// Returns e: the outputs in v are the transform times 2^e.
template <int MF, int L>
__device__ __forceinline__ int fft1024_mfma(C<float>* v, C<float>* line, int l,
                                            const C<float>* tw, const MfmaA<MF>& A) {
  constexpr int P = 1024 / L;
  int e = 0;
  if constexpr (MF == 1) {
    // largest magnitude as bits: |x| orders like its bit pattern
    unsigned mb = 0;
#pragma unroll
    for (int p = 0; p < P; ++p)
      mb = max(mb, max(__float_as_uint(v[p].re) & 0x7fffffffu,
                       __float_as_uint(v[p].im) & 0x7fffffffu));
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) mb = max(mb, (unsigned)__shfl_xor((int)mb, o, 64));
    if constexpr (L > 64) {
      // the line's two wavefronts exchange their maxima through the spare
      // slot at index 1152 (c2c lines use indices up to pad8(1023) = 1150)
      unsigned* ex = (unsigned*)(line + 1152);
      if ((l & 63) == 0) ex[l >> 6] = mb;
      __syncthreads();
      mb = max(ex[0], ex[1]);
    }
    // |x| 2^e < 2^10 (a subnormal maximum counts as exponent -127: still below)
    e = mb ? 9 - ((int)(mb >> 23) - 127) : 0;
#pragma unroll
    for (int p = 0; p < P; ++p) v[p] = mk(ldexpf(v[p].re, e), ldexpf(v[p].im, e));
  }
  C<float>* lp = line + pad8(l);
#pragma unroll
  for (int p = 0; p < P; ++p) lp[pad8_off(L * p)] = v[p];
  line_sync<L>();
  stage16_mfma<MF, 1, L>(line, l, tw, A);
  line_sync<L>();
  stage16_mfma<MF, 16, L>(line, l, tw, A);
  line_sync<L>();
  stage<float, 1024, L, P, 4, 256, false, true>(v, line, l, tw);
  return e;
}

// ---- 1024 = 32 x 32 on MFMA, split-FP16, the line in registers -----------------------
// Four-step form of a 1024-point line x[n1 + 32 n2] on one wavefront:
//   Y[n1, k2] = sum_n2 x[n1 + 32 n2] w32^(n2 k2)       (32-point DFTs, MFMA)
//   Y'        = Y[n1, k2] w1024^(n1 k2)                (twiddle, VALU)
//   X[k2 + 32 k1] = sum_n1 Y'[n1, k2] w32^(n1 k1)      (32-point DFTs, MFMA)
// Each step is one complex 32x32x32 product, as four real products on
// v_mfma_f32_32x32x16_f16 over two K-chunks.  The first step takes the
// data as the A operand (row n1 = lane % 32, K = n2 in registers), so its
// accumulator holds Y with k2 across the lanes and n1 in registers -- the
// layout the second step takes as its B operand (K = n1 in registers, column
// k2 = lane % 32).  No LDS exchange and no transposition are needed inside
// the line.  K slot (h = lane / 32, t, chunk c) stands for index
// pi(8h + t + 16c) = 4h + (t & 3) + 8 (t >> 2) + 16c in both steps, which
// is the accumulator row order; so the DFT-matrix slices the two steps need
// are the same registers, and register p of lane l holds point
//   J(l, p) = (l & 31) + 128 (l >> 5) + 32 (p & 3) + 256 (p >> 2)
// both on input and on output.  Operands are split x = hi + lo in FP16
// (round to nearest both times; hi_A hi_B + hi_A lo_B + lo_A hi_B in FP32).
// The line is scaled on entry by 2^e with |x| 2^e < 2^9, so the first
// step's outputs (sums of 32 complex terms) stay below 2^15 and in FP16
// range for the second step.
__device__ __forceinline__ constexpr int jmap_lane(int l) { return (l & 31) + 128 * (l >> 5); }
__device__ __forceinline__ constexpr int jmap_off(int p) { return 32 * (p & 3) + 256 * (p >> 2); }

static __constant__ float kCos32[32] = {
    1.0f, 0.98078528040323044913f, 0.92387953251128675613f, 0.83146961230254523708f,
    0.70710678118654752440f, 0.55557023301960222474f, 0.38268343236508977173f,
    0.19509032201612826785f, 0.0f, -0.19509032201612826785f, -0.38268343236508977173f,
    -0.55557023301960222474f, -0.70710678118654752440f, -0.83146961230254523708f,
    -0.92387953251128675613f, -0.98078528040323044913f, -1.0f, -0.98078528040323044913f,
    -0.92387953251128675613f, -0.83146961230254523708f, -0.70710678118654752440f,
    -0.55557023301960222474f, -0.38268343236508977173f, -0.19509032201612826785f, 0.0f,
    0.19509032201612826785f, 0.38268343236508977173f, 0.55557023301960222474f,
    0.70710678118654752440f, 0.83146961230254523708f, 0.92387953251128675613f,
    0.98078528040323044913f};

typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

// hi + lo split of 8 floats into two FP16 operand slices, both rounded to
// nearest: pairwise, so that one packed conversion (v_cvt_pk_f16_f32) rounds
// two values and the hi halves convert back from that packed result -- 3
// VALU per value instead of 4 (a separate scalar conversion for the
// back-conversion), with the same bits
__device__ __forceinline__ void split8(const float (&x)[8], f16x8& hi, f16x8& lo) {
#pragma unroll
  for (int t = 0; t < 8; t += 2) {
    const f32x2 x2 = {x[t], x[t + 1]};
    const f16x2 h = __builtin_convertvector(x2, f16x2);
    const f32x2 hb = __builtin_convertvector(h, f32x2);
    const f16x2 l2 = __builtin_convertvector(x2 - hb, f16x2);
    hi[t] = h[0];
    hi[t + 1] = h[1];
    lo[t] = l2[0];
    lo[t + 1] = l2[1];
  }
}
__device__ __forceinline__ f16x8 neg8(f16x8 v) {
  return __builtin_bit_cast(f16x8, __builtin_bit_cast(s16x8, v) ^ (short)0x8000);
}

// The DFT-matrix slices of the two steps live in LDS, 8 KB per workgroup,
// rather than in 32 registers of every lane: entry 64 s + l is lane l's
// slice s = 2 part + c of part (Re hi, Re lo, Im hi, Im lo) and K-chunk c,
// i.e. w32^((l mod 32) pi(8 (l div 32) + t + 16 c)) for t = 0..7.  Each
// operand is read (16 bytes per lane, conflict-free) just before its MFMAs.
constexpr int kMfma32TableBytes = 8 * 64 * 16;
__device__ __forceinline__ void mfma32_table_init(f16x8* tab, int tid, int nthr) {
  for (int idx = tid; idx < 128; idx += nthr) {
    const int l = idx & 63, c = idx >> 6;
    const int j = l & 31, h = l >> 5;
    f16x8 rh, rl, ih, il;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int n = 4 * h + (t & 3) + 8 * (t >> 2) + 16 * c;
      const int m = (j * n) & 31;
      const float fr = kCos32[m];
      const float fi = -kCos32[(m + 24) & 31];   // -sin(2 pi m / 32)
      const _Float16 a = (_Float16)fr, b = (_Float16)fi;
      rh[t] = a;
      rl[t] = (_Float16)(fr - (float)a);
      ih[t] = b;
      il[t] = (_Float16)(fi - (float)b);
    }
    tab[(0 + c) * 64 + l] = rh;
    tab[(2 + c) * 64 + l] = rl;
    tab[(4 + c) * 64 + l] = ih;
    tab[(6 + c) * 64 + l] = il;
  }
}

// D += A B with A, B split in hi + lo
__device__ __forceinline__ f32x16 mm3(f16x8 ah, f16x8 al, f16x8 bh, f16x8 bl, f32x16 d) {
  d = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh, d, 0, 0, 0);
  d = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl, d, 0, 0, 0);
  d = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh, d, 0, 0, 0);
  return d;
}

// v: point J(l, p) in register p on entry, X[J(l, p)] on exit (times 2^e).
// tw2: the step twiddles, entry 64 i + l = w1024^((l & 31) n1), n1 the row
// of accumulator register i; ftab: the DFT-matrix slices (LDS).  Returns e.
// Chunk by chunk, each operand is built (and each slice read) right before
// the MFMAs that use it, so that few values are live at once.
__device__ __forceinline__ int fft1024_mfma32(C<float>* v, int l, const C<float>* __restrict__ tw2,
                                              const f16x8* ftab) {
  // scale: largest magnitude as bits (|x| orders like its bit pattern)
  unsigned mb = 0;
#pragma unroll
  for (int p = 0; p < 16; ++p)
    mb = max(mb, max(__float_as_uint(v[p].re) & 0x7fffffffu,
                     __float_as_uint(v[p].im) & 0x7fffffffu));
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) mb = max(mb, (unsigned)__shfl_xor((int)mb, o, 64));
  const int e = mb ? 8 - ((int)(mb >> 23) - 127) : 0;
  auto F = [&](int part, int c) { return ftab[(2 * part + c) * 64 + l]; };
  f32x16 dr = {}, di = {};
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    float xr[8], xi[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      xr[t] = ldexpf(v[8 * c + t].re, e);
      xi[t] = ldexpf(v[8 * c + t].im, e);
    }
    f16x8 rh, rl, ih, il;
    split8(xr, rh, rl);
    split8(xi, ih, il);
    // Re Y = Xr Fr - Xi Fi, Im Y = Xr Fi + Xi Fr (data as the A operand)
    const f16x8 frh = F(0, c), frl = F(1, c);
    dr = mm3(rh, rl, frh, frl, dr);
    di = mm3(ih, il, frh, frl, di);
    const f16x8 fih = F(2, c), fil = F(3, c);
    dr = mm3(neg8(ih), neg8(il), fih, fil, dr);
    di = mm3(rh, rl, fih, fil, di);
    __builtin_amdgcn_sched_barrier(0);
  }
  f32x16 zr = {}, zi = {};
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    // twiddle w1024^(n1 k2) on this chunk's rows, then the split
    float yr[8], yi[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int i = 8 * c + t;
      const C<float> w = tw2[64 * i + l];
      const float a = dr[i], b = di[i];
      yr[t] = a * w.re - b * w.im;
      yi[t] = a * w.im + b * w.re;
    }
    f16x8 rh, rl, ih, il;
    split8(yr, rh, rl);
    split8(yi, ih, il);
    // the DFT matrix as the A operand now
    const f16x8 frh = F(0, c), frl = F(1, c);
    zr = mm3(frh, frl, rh, rl, zr);
    zi = mm3(frh, frl, ih, il, zi);
    const f16x8 fih = F(2, c), fil = F(3, c);
    zr = mm3(fih, fil, neg8(ih), neg8(il), zr);
    zi = mm3(fih, fil, rh, rl, zi);
    __builtin_amdgcn_sched_barrier(0);
  }
#pragma unroll
  for (int p = 0; p < 16; ++p) v[p] = mk(zr[p], zi[p]);
  return e;
}

// mixed radix: P is a multiple of every radix of the schedule
#define PFFT_CFG(MM, LL, PP, ...)                                        \
  template <> struct Cfg<MM> {                                           \
    static constexpr int L = LL, P = PP;                                 \
    template <typename R>                                                \
    __device__ static void fft(C<R>* v, C<R>* line, int l, const C<R>* tw) { \
      fft_rec<R, MM, L, P, 1, true, __VA_ARGS__>(v, line, l, tw);        \
    }                                                                    \
  };
PFFT_CFG(32, 4, 8, 8, 4)
// composite radices 6, 10, 12, 20 (prime-factor DFTs in registers) save
// LDS round trips over schedules of 2, 3, 4 and 5
PFFT_CFG(48, 4, 12, 4, 12)
PFFT_CFG(96, 8, 12, 4, 4, 6)
PFFT_CFG(160, 8, 20, 4, 4, 10)
PFFT_CFG(192, 16, 12, 4, 4, 12)
PFFT_CFG(320, 16, 20, 4, 4, 20)
PFFT_CFG(384, 32, 12, 4, 4, 4, 6)
PFFT_CFG(640, 32, 20, 4, 4, 4, 10)
PFFT_CFG(768, 64, 12, 4, 4, 4, 12)
#undef PFFT_CFG

// kinds of the fast pass:
//   K_C2C  complex FFT (BFFT as conj(FFT(conj)))
//   K_R2C  r2c of real length 2M via the M-point FFT + split twiddle
//   K_C2R  c2r to real length 2M via merge twiddle + M-point FFT
//   K_DCT2C REDFT10 of complex data: Makhoul reorder, M-point FFT, Hermitian
//           split of the real/imaginary parts and the post-twiddle
//   K_DCT3C REDFT01 of complex data: pre-twiddle over (k, M-k), backward
//           M-point FFT, inverse Makhoul reorder
//   K_DST2C RODFT10 of complex data: y_k = DCT-II((-1)^j x_j)_{M-1-k}, the
//           sign folded into the Makhoul gather, the reversal into the
//           post-processing's reads
//   K_DST3C RODFT01 of complex data: y_k = (-1)^k DCT-III(x_{M-1-j})_k, the
//           reversal folded into the pre-twiddle's gather, the sign into the
//           output factor
//   K_DHTC  DHT of complex data: one M-point FFT of x = a + i b, then
//           A = (Z_k + conj Z_{M-k}) / 2, B = (Z_k - conj Z_{M-k}) / 2i and
//           y_k = (Re A - Im A) + i (Re B - Im B)
//   K_DCT4C REDFT11 of complex data: the M/2-point FFTs of the halves
//           z_m = (x_2m + i x_{M-1-2m}) c_m of the real part (even points)
//           and of the imaginary part (odd points) interleaved into one
//           M-point FFT W; then Z_m = (W_m + W_{m+M/2}) / 2 for the real part,
//           w_M^-m (W_m - W_{m+M/2}) / 2 for the imaginary part, and
//           y_2m = 2 Re(d_m Z_m), y_{M-1-2m} = -2 Im(d_m Z_m), d_m = e^{-i pi m / M}
//   K_DST4C RODFT11 of complex data: y_k = (-1)^k DCT-IV(x_{M-1-j})_k, the
//           reversal folded into the halves' gather, the sign into the output
//   K_DCT1C REDFT00 of complex data of n = M/2 + 1 points: the M-point FFT of
//           the even extension (x_j, j <= M/2; x_{M-j} above) is the DCT-I of
//           the real part plus i times that of the imaginary part, rows 0..M/2
//   K_DST1C RODFT00 of complex data of n = M/2 - 1 points: the M-point FFT W
//           of the odd extension (0, x_0 .. x_{n-1}, 0, -x_{n-1} .. -x_0)
//           gives y_k = i W_{k+1}, k < n
// K_REAL_PAIRS (OR-ed into the R2R kinds): the same transforms of real data,
// two real lines (2j, 2j + 1) forming the real and imaginary parts of one
// complex line.  On the side where the pass is transposed the two lines are
// adjacent in memory, so the tile moves them as complex elements; on a line
// side the pair's real and imaginary parts are read from (written to) two
// lines, a.lsin[0] / 2 (a.lsout[0] / 2) elements apart.
enum { K_C2C = 0, K_R2C = 1, K_C2R = 2, K_DCT2C = 3, K_DCT3C = 4, K_DST2C = 5, K_DST3C = 6,
       K_DHTC = 7, K_DCT4C = 8, K_DST4C = 9, K_DCT1C = 10, K_DST1C = 11 };
constexpr int K_REAL_PAIRS = 16;
enum { IO_LINE = 0, IO_TILE = 1 };

// points per line read and written (the r2c output and c2r input carry
// M + 1, the DCT-I and DST-I lines M/2 + 1 and M/2 - 1)
template <int KIND, int M>
constexpr int rows_in() {
  return KIND == K_C2R ? M + 1 : KIND == K_DCT1C ? M / 2 + 1 : KIND == K_DST1C ? M / 2 - 1 : M;
}
template <int KIND, int M>
constexpr int rows_out() {
  return KIND == K_R2C ? M + 1 : KIND == K_DCT1C ? M / 2 + 1 : KIND == K_DST1C ? M / 2 - 1 : M;
}

// Output stores, plain or with the nontemporal (streaming) hint: the next
// pass reads the output only after the whole pass, and whether keeping it in
// the caches helps depends on the pass and its neighbours, so the autotuner
// chooses per pass.
template <bool NT, typename R>
__device__ __forceinline__ void st_line(C<R>* p, C<R> v) {
  if constexpr (NT) {
    typedef R v2 __attribute__((ext_vector_type(2)));
    __builtin_nontemporal_store(v2{v.re, v.im}, (v2*)p);
  } else {
    *p = v;
  }
}
template <bool NT, typename R>
__device__ __forceinline__ void st_real(R* p, R v) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}
template <bool NT>
__device__ __forceinline__ void st_vec(float4* p, float4 v) {
  if constexpr (NT) {
    typedef float v4 __attribute__((ext_vector_type(4)));
    __builtin_nontemporal_store(v4{v.x, v.y, v.z, v.w}, (v4*)p);
  } else {
    *p = v;
  }
}

// tile element (row k, line t) sits at column t ^ tile_xor(k) of row k.
// LDS has 64 banks of 4 bytes; a 64-bit access is served 32 lanes per cycle,
// a 128-bit one 16 lanes per cycle.
// SW = 1 (lines of 32 or more lanes (float) / 16 or more (double)): the XOR
// is (k / RPC) mod T, RPC = rows per 256 bytes, so a column written or read by
// 32 (float) / 16 (double) consecutive rows meets 32 / 16 distinct bank slots.
// A float pair (t, t+1), t even, stays in one aligned 16-byte slot, in swapped
// order when the XOR is odd (see tile_vec).
// SW = 0: the XOR keeps float pairs in order ((k mod T/2) * 2, (k mod T)
// for double): 2- to 4-way column conflicts, but it repeats every T/2 (T) rows,
// which lines of fewer lanes need for lane-base addressing.
template <typename R, int T, int SW>
__device__ __forceinline__ int tile_xor(int k) {
  if constexpr (SW) {
    constexpr int RPC = (sizeof(R) == 4 ? 32 : 16) / T;
    static_assert(RPC >= 1, "tile row wider than 256 bytes");
    return (k / RPC) & (T - 1);
  } else if constexpr (sizeof(R) == 4) {
    return (k & (T / 2 - 1)) << 1;
  } else {
    return k & (T - 1);
  }
}
template <typename R, int T, int SW>
__device__ __forceinline__ int tile_idx(int k, int t) {
  return k * T + (t ^ tile_xor<R, T, SW>(k));
}
// 16-byte vector of row k holding columns (t, t + EPV - 1), t a multiple of
// EPV: its element offset, and whether its two float elements are swapped.
template <typename R, int T, int SW>
__device__ __forceinline__ int tile_vec(int k, int t, bool* swap) {
  const int x = tile_xor<R, T, SW>(k);
  if constexpr (sizeof(R) == 4) {
    *swap = (x & 1) != 0;
    return k * T + ((t ^ x) & ~1);
  } else {
    *swap = false;
    return k * T + (t ^ x);
  }
}
typedef float tile_v4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ tile_v4 swap_pair(tile_v4 v, bool sw) {
  return sw ? tile_v4{v.z, v.w, v.x, v.y} : v;
}

// Cooperative tile copies.  Full rounds (every thread busy) are unrolled
// with compile-time trip counts, so a thread issues all its global loads
// before the first LDS write; the partial last round is a separate guarded
// step that keeps nothing live across its branch.
template <typename R, int T, int ROWS, int NTHR, int SW, bool NT, bool ACC = false>
__device__ __forceinline__ void tile_store(const C<R>* tile, const PassArgs& a, int64_t base,
                                           int tid) {
  C<R>* out = (C<R>*)a.out;
  // ACC: add into the output (spectral operators with accumulate)
  auto put = [&](float4* dst, float4 v) {
    if constexpr (ACC) {
      const float4 o = *dst;
      if constexpr (sizeof(R) == 4) {
        v.x += o.x; v.y += o.y; v.z += o.z; v.w += o.w;
      } else {
        double2 vd = *(double2*)&v;
        const double2 od = *(const double2*)&o;
        vd.x += od.x; vd.y += od.y;
        v = *(float4*)&vd;
      }
    }
    st_vec<NT>(dst, v);
  };
  constexpr int EPV = sizeof(R) == 4 ? 2 : 1;      // complex per 16-byte vector
  constexpr int VPR = T / EPV;                     // vectors per row
  constexpr int TOTAL = ROWS * VPR;
  constexpr int FULL = TOTAL / NTHR;
  auto get = [&](int k, int t) -> float4 {
    bool sw;
    const int o = tile_vec<R, T, SW>(k, t, &sw);
    const tile_v4 v = swap_pair(*(const tile_v4*)&tile[o], sw);
    return make_float4(v.x, v.y, v.z, v.w);
  };
#pragma unroll
  for (int i = 0; i < FULL; ++i) {
    const int e = tid + i * NTHR;
    const int k = e / VPR;
    const int t = EPV * (e - k * VPR);
    put((float4*)&out[base + (int64_t)k * a.sk_out + t], get(k, t));
  }
  if constexpr (TOTAL % NTHR != 0) {
    const int e = tid + FULL * NTHR;
    if (e < TOTAL) {
      const int k = e / VPR;
      const int t = EPV * (e - k * VPR);
      put((float4*)&out[base + (int64_t)k * a.sk_out + t], get(k, t));
    }
  }
}

template <typename R, int T, int ROWS, int NTHR, int SW>
__device__ __forceinline__ void tile_load(C<R>* tile, const PassArgs& a, int64_t base, int tid) {
  const C<R>* in = (const C<R>*)a.in;
  constexpr int EPV = sizeof(R) == 4 ? 2 : 1;
  constexpr int VPR = T / EPV;
  constexpr int TOTAL = ROWS * VPR;
  constexpr int FULL = TOTAL / NTHR;
  // global and LDS accesses never alias, so the unrolled loads are hoisted
  // ahead of the LDS writes
  auto set = [&](int k, int t, tile_v4 v) {
    bool sw;
    const int o = tile_vec<R, T, SW>(k, t, &sw);
    *(tile_v4*)&tile[o] = swap_pair(v, sw);
  };
#pragma unroll
  for (int i = 0; i < FULL; ++i) {
    const int e = tid + i * NTHR;
    const int k = e / VPR;
    const int t = EPV * (e - k * VPR);
    set(k, t, *(const tile_v4*)&in[base + (int64_t)k * a.sk_in + t]);
  }
  if constexpr (TOTAL % NTHR != 0) {
    const int e = tid + FULL * NTHR;
    if (e < TOTAL) {
      const int k = e / VPR;
      const int t = EPV * (e - k * VPR);
      set(k, t, *(const tile_v4*)&in[base + (int64_t)k * a.sk_in + t]);
    }
  }
}

// Lines of a tile: T; lines in flight per pass over the workgroup: TQ =
// T / LPW (one per L lanes), so each line slot handles LPW lines in turn.
// A transposed store goes out in NSP row chunks of M / NSP rows, so the
// tile needs 1 / NSP of the rows in LDS and more workgroups fit on a CU.
// PF: the prefetching variants keep the twiddles (and the c2r merge factors)
// in LDS after the tile / line slots.  The 32 x 32 MFMA lines (MF = 3) keep
// the whole line in registers and use no line slot; their DFT-matrix table
// follows the tile.
template <typename R, int M, int T, int LPW, int KIND, int LD, int ST, int NSP, int MF = 0>
constexpr size_t lds_base() {   // T / LPW line slots whatever the lanes per line
  constexpr size_t lines = MF == 3 ? 0 : (size_t)(T / LPW) * line_pitch<M>() * sizeof(C<R>);
  constexpr int rows_ld = rows_in<KIND, M>();
  constexpr int rows_st = NSP == 1 ? rows_out<KIND, M>() : M / NSP + (KIND == K_R2C ? 1 : 0);
  constexpr size_t tld = LD == IO_TILE ? (size_t)rows_ld * T * sizeof(C<R>) : 0;
  constexpr size_t tst = ST == IO_TILE ? (size_t)rows_st * T * sizeof(C<R>) : 0;
  constexpr size_t tile = tld > tst ? tld : tst;
  return lines > tile ? lines : tile;
}
template <typename R, int M, int T, int LPW, int KIND, int LD, int ST, int NSP, int PF = 0,
          int MF = 0>
constexpr size_t lds_bytes() {
  constexpr size_t tables =
      PF ? (size_t)(M + (KIND == K_C2R || KIND == K_R2C ? M + 1 : 0)) * sizeof(C<R>) : 0;
  constexpr size_t mf3 = MF == 3 ? kMfma32TableBytes : 0;
  return lds_base<R, M, T, LPW, KIND, LD, ST, NSP, MF>() + tables + mf3;
}

// Register staging of a transposed input tile (prefetching variants): the
// loads of the next tile are issued before the current tile is transformed
// and land in LDS at the top of the next round.
template <typename R, int T, int ROWS, int NTHR, int SW>
struct TileStage {
  static constexpr int EPV = sizeof(R) == 4 ? 2 : 1;   // complex per 16-byte vector
  static constexpr int VPR = T / EPV;
  static constexpr int TOTAL = ROWS * VPR;
  static constexpr int FULL = TOTAL / NTHR;
  static constexpr int N = (TOTAL + NTHR - 1) / NTHR;
  // a native vector type: HIP's float4 struct is copied by memcpy, which
  // keeps the array in scratch memory
  typedef float v4 __attribute__((ext_vector_type(4)));
  v4 r[N];
  __device__ __forceinline__ void fetch(const PassArgs& a, int64_t base, int tid) {
    const C<R>* in = (const C<R>*)a.in;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const int e = tid + i * NTHR;
      if (i < FULL || e < TOTAL) {
        const int k = e / VPR;
        const int t = EPV * (e - k * VPR);
        r[i] = *(const v4*)&in[base + (int64_t)k * a.sk_in + t];
      }
    }
  }
  __device__ __forceinline__ void put(C<R>* tile, int tid) const {
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const int e = tid + i * NTHR;
      if (i < FULL || e < TOTAL) {
        const int k = e / VPR;
        const int t = EPV * (e - k * VPR);
        bool sw;
        const int o = tile_vec<R, T, SW>(k, t, &sw);
        *(v4*)&tile[o] = swap_pair(r[i], sw);
      }
    }
  }
};

// OCC > 0 asks the compiler for at least OCC waves per SIMD (a register
// budget of 512 / OCC): more waves in flight at the price of possible spills;
// the autotuner keeps such a variant only where it is faster.
// OPK = 1 compiles in the spectral operator of pfft_execute_op (c2c only):
// the elements read (op_side 1) or written (op_side 2) are multiplied by
// alpha m(k); those variants run only the passes that carry an operator.
// PF = 1: a persistent grid where each workgroup issues the loads of its
// next tile (into registers: the transposed tile, or its lines) before
// transforming the current one, with the twiddles and split/merge factors
// read from LDS so that no other global load waits behind the prefetch.
template <typename R, int M, int T, int LPW, int KINDF, int LD, int ST, int MF, int NSP, int W,
          int OCC, int OPK, int PF, int NT>
__global__ __launch_bounds__((T / LPW * Cfg<M, W>::L))
__attribute__((amdgpu_waves_per_eu(OCC > 0 ? OCC : 1))) void k_fast(PassArgs a) {
  constexpr int KIND = KINDF & 15;
  constexpr bool RP = (KINDF & K_REAL_PAIRS) != 0;   // real data, two lines per complex line
  constexpr int L = Cfg<M, W>::L, P = Cfg<M, W>::P;
  constexpr int TQ = T / LPW;
  constexpr int NTHR = TQ * L;
  static_assert(TQ * LPW == T, "LPW must divide T");
  static_assert(P % NSP == 0 && (NSP == 1 || ST == IO_TILE), "store chunks split P");
  static_assert(OPK == 0 || KIND == K_C2C, "spectral operators ride on c2c passes");
  static_assert(!RP || KIND >= K_DCT2C, "real pairs: R2R kinds");
  // inputs / outputs per line (the r2c output and c2r input carry M + 1)
  constexpr int ROWS_IN = rows_in<KIND, M>();
  constexpr int ROWS_OUT = rows_out<KIND, M>();
  // DCT-I / DST-I: an extended line of ROWS_IN points, rows OFF .. OFF +
  // ROWS_OUT - 1 of its FFT kept
  constexpr bool EXT = KIND == K_DCT1C || KIND == K_DST1C;
  constexpr int OFF = KIND == K_DST1C ? 1 : 0;
  static_assert(!EXT || NSP == 1, "extended lines: whole-tile stores");
  constexpr bool DCTLIKE = KIND == K_DCT2C || KIND == K_DCT3C || KIND == K_DST2C ||
                           KIND == K_DST3C || KIND == K_DCT4C || KIND == K_DST4C;
  constexpr bool GATHER = KIND == K_C2R || DCTLIKE || EXT;          // pre reads (k, M-k)
  // DCT-III / DST-III into a transposed tile: the inverse Makhoul reorder is
  // a row permutation, applied by the tile writes themselves (no pass of the
  // line through its LDS slot)
  constexpr bool ROWPERM = (KIND == K_DCT3C || KIND == K_DST3C) && ST == IO_TILE && NSP == 1;
  constexpr bool POST =
      (KIND == K_R2C || DCTLIKE || KIND == K_DHTC) && !ROWPERM;   // post reads LDS
  constexpr int NX = KIND == K_R2C ? P + 1 : P;   // outputs per lane per line
  extern __shared__ __align__(16) char smem[];
  C<R>* lds = (C<R>*)smem;
  const int tid = threadIdx.x;
  const int g = tid / L;         // line slot
  const int l = tid - g * L;     // lane within the line
  C<R>* line = lds + g * line_pitch<M>();
  C<R>* tile = lds;
  // Lane bases for the points k = l + L p and M - k = (M - L (p+1)) + (L - l):
  // pad8 is additive over multiples of 8 and the tile swizzle repeats every
  // SWP rows, so with L a multiple of those each access is base + constant.
  // conflict-free tile swizzle when the line's lanes cover its period
  constexpr int SW = L % (sizeof(R) == 4 ? 32 : 16) == 0 && T * sizeof(C<R>) <= 256 ? 1 : 0;
  constexpr int SWP = SW ? (sizeof(R) == 4 ? 32 : 16) : (sizeof(R) == 4 ? T / 2 : T);
  // MF = 3 (1024 = 32 x 32 on MFMA): register p of lane l holds point
  // J(l, p) = jmap_lane(l) + jmap_off(p) instead of l + L p; the offsets are
  // multiples of 32 rows, so the tile swizzle (period SWP | 32) keeps the
  // lane-base addressing
  constexpr bool JM = MF == 3;
  static_assert(!JM || (KIND == K_C2C && OPK == 0 && PF == 0 && L == 64 && NSP <= 4),
                "32 x 32 MFMA lines: c2c, one wavefront per line");
  const int lrow = JM ? jmap_lane(l) : l;
  auto ROW = [&](int p) -> int { return JM ? jmap_off(p) : L * p; };
  constexpr bool ADD8 = L % 8 == 0, ADDT = (JM ? 32 : L) % SWP == 0;
  C<R>* const lp = line + pad8(l);
  C<R>* const lm = line + pad8(L - l);
  auto LK = [&](int p) -> C<R>& {   // line[pad8(l + L p)]
    if constexpr (ADD8) return lp[pad8_off(L * p)];
    else return line[pad8(l + L * p)];
  };
  auto LMK = [&](int p) -> C<R>& {  // line[pad8(M - l - L p)], for l + L p > 0
    if constexpr (ADD8) return lm[pad8_off(M - L * (p + 1))];
    else return line[pad8(M - l - L * p)];
  };
  auto TK = [&](int p, int t, int k0) -> C<R>& {   // tile[tile_idx(k0 + l + L p, t)]
    if constexpr (ADDT) return tile[tile_idx<R, T, SW>(lrow, t) + (k0 + ROW(p)) * T];
    else return tile[tile_idx<R, T, SW>(k0 + lrow + ROW(p), t)];
  };
  auto TMK = [&](int p, int t) -> C<R>& {          // tile[tile_idx(M - l - L p, t)]
    if constexpr (ADDT) return tile[tile_idx<R, T, SW>(L - l, t) + (M - L * (p + 1)) * T];
    else return tile[tile_idx<R, T, SW>(M - l - L * p, t)];
  };
  static_assert(PF == 0 || ((MF == 0 || MF == 4) && L <= 64 && OPK == 0 &&
                            (LD == IO_TILE ? KIND == K_C2C || KIND == K_C2R
                                           : KIND == K_C2C || KIND == K_R2C)),
                "prefetching variants: c2c, c2r with transposed loads, r2c with line loads");
  constexpr bool PFT = PF && LD == IO_TILE, PFL = PF && LD == IO_LINE;
  constexpr size_t TBL = lds_base<R, M, T, LPW, KIND, LD, ST, NSP, MF>();
  if constexpr (PF) {
    C<R>* st = (C<R>*)(smem + TBL);
    for (int i = tid; i < M; i += NTHR) st[i] = ((const C<R>*)a.tw)[i];
    if constexpr (KIND == K_C2R || KIND == K_R2C)
      for (int i = tid; i <= M; i += NTHR) st[M + i] = ((const C<R>*)a.aux)[i];
    if constexpr (PFL) __syncthreads();   // line loads reach the FFT without a barrier
  }
  const C<R>* __restrict__ tw = PF ? (const C<R>*)(smem + TBL) : (const C<R>*)a.tw;
  const C<R>* __restrict__ aux = PF ? (const C<R>*)(smem + TBL) + M : (const C<R>*)a.aux;
  static_assert(!DCTLIKE || L % 4 == 0, "DCT-IV twiddles by lane base");
  // backward FFTs run as conj(FFT(conj x))
  const bool conj_in =
      KIND == K_C2C ? a.bwd != 0 : (KIND == K_C2R || KIND == K_DCT3C || KIND == K_DST3C);
  const bool conj_out = KIND == K_C2C && a.bwd != 0;
  const R sc = a.do_scale ? (R)a.scale : R(1);
  static_assert(MF == 0 || (MF == 4 ? (M == 256 || M == 512 || M == 1024) && W == 1 : M == 1024) &&
                               sizeof(R) == 4,
                "MFMA stages: 1024-point f32 lines; cross-lane exchange: 256/512/1024-point "
                "f32 lines, one wavefront per line");
  MfmaA<MF == 1 || MF == 2 ? MF : 1> mfA;
  if constexpr (MF == 1 || MF == 2) mfA.init(l % 64);
  // the 32 x 32 MFMA form's DFT-matrix slices, after the tile
  [[maybe_unused]] f16x8* const ftab = (f16x8*)(smem + TBL);
  if constexpr (JM) {
    mfma32_table_init(ftab, tid, NTHR);
    __syncthreads();
  }

  [[maybe_unused]] TileStage<R, T, PFT ? ROWS_IN : 1, NTHR, SW> stg;
  // the next tile's lines (contiguous loads), as native vectors
  typedef R cv2 __attribute__((ext_vector_type(2)));
  [[maybe_unused]] cv2 nl[PFL ? LPW : 1][PFL ? P : 1];
  auto fetch_lines = [&](const TileCtx& cc) {
    if constexpr (PFL) {
#pragma unroll
      for (int q = 0; q < LPW; ++q) {
        const int t = q * TQ + g;
        const cv2* in = (const cv2*)((const R*)a.in +
                                     (KIND == K_R2C ? 1 : 2) * (cc.in_base + (int64_t)t * a.lsin[0]));
        if (t < cc.Tact) {
#pragma unroll
          for (int p = 0; p < P; ++p) nl[q][p] = in[l + L * p];
        }
      }
    }
  };
  [[maybe_unused]] TileCtx cn;
  if constexpr (PF) {
    if (blockIdx.x < a.ntiles) {
      tile_init(a, blockIdx.x, cn);
      if constexpr (PFT) stg.fetch(a, cn.in_base, tid);
      else fetch_lines(cn);
    }
  }
  for (int64_t tl = blockIdx.x; tl < a.ntiles; tl += gridDim.x) {
    TileCtx c;
    if constexpr (PF) c = cn;
    else tile_init(a, tl, c);
    C<R> v[LPW][NX];
    // spectral operator: the state of the tile's shared line dims, once
    [[maybe_unused]] OpLine<R> otile;
    [[maybe_unused]] int64_t om0 = 0;
    if constexpr (OPK) otile = op_tile<R>(a, tl, &om0);

    // ---------------- load and pre-process ----------------
    if constexpr (PFT) {
      // the previous round ended with a barrier: the tile is free (and, on
      // the first round, the tables are written)
      stg.put(tile, tid);
      __syncthreads();
      const int64_t nx = tl + gridDim.x;
      if (nx < a.ntiles) {
        tile_init(a, nx, cn);
        stg.fetch(a, cn.in_base, tid);
      }
    } else if constexpr (LD == IO_TILE) {
      tile_load<R, T, ROWS_IN, NTHR, SW>(tile, a, c.in_base, tid);
      __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < LPW; ++q) {
      // the 32 x 32 MFMA lines load each line right before its transform
      // (below), so that one line's registers are live at a time
      if constexpr (JM) break;
      const int t = q * TQ + g;
      const bool active = t < c.Tact;
      if constexpr (GATHER) {
        // element x_k of this line: from the tile; with line loads, from the
        // line's LDS slot (filled first), or -- for the gathers of x_k with
        // x_{M-k} (c2r merge, DCT-III / DST-III pre-twiddles, the DCT-I /
        // DST-I extensions) -- straight
        // from global memory where the gather needs it: each wavefront load
        // covers one contiguous run, ascending or descending, and the second
        // touch of an element comes from the caches.  (The Makhoul reorder's
        // stride-2 reads measured slower that way: those lines go through
        // the slot.)
        constexpr bool DIRECT =
            LD == IO_LINE && (KIND == K_C2R || KIND == K_DCT3C || KIND == K_DST3C || EXT);
        [[maybe_unused]] const C<R>* lin =
            (const C<R>*)a.in + c.in_base + (int64_t)t * a.lsin[0];
        [[maybe_unused]] const R* lia = (const R*)a.in + c.in_base + (int64_t)t * a.lsin[0];
        if constexpr (LD == IO_LINE && !DIRECT) {
          if (q > 0) line_sync<L>();
          if (active) {
            // real pairs: the pair's two real lines, lsin[0] / 2 apart
            if constexpr (RP)
              for (int k = l; k < ROWS_IN; k += L) line[pad8(k)] = mk(lia[k], lia[a.lsin[0] / 2 + k]);
            else
              for (int k = l; k < ROWS_IN; k += L) line[pad8(k)] = lin[k];
          }
          line_sync<L>();
        }
        auto X = [&](int k) -> C<R> {
          if constexpr (LD == IO_TILE) {
            return tile[tile_idx<R, T, SW>(k, t)];
          } else if constexpr (!DIRECT) {
            return line[pad8(k)];
          } else {
            if (!active) return mk(R(0), R(0));
            if constexpr (RP) return mk(lia[k], lia[a.lsin[0] / 2 + k]);
            else return lin[k];
          }
        };
        auto XK = [&](int p) -> C<R> {    // x_k, k = l + L p
          if constexpr (LD == IO_TILE) return TK(p, t, 0);
          else if constexpr (!DIRECT) return LK(p);
          else return X(l + L * p);
        };
        auto XMK = [&](int p) -> C<R> {   // x_{M-k} for k > 0 (x_M when k = 0)
          if constexpr (LD == IO_TILE) return TMK(p, t);
          else if constexpr (!DIRECT) return LMK(p);
          else return X(M - l - L * p);
        };
#pragma unroll
        for (int p = 0; p < P; ++p) {
          const int k = l + L * p;
          C<R> z;
          if constexpr (KIND == K_C2R) {
            // Hermitian merge: Z_k = (X_k + conj X_{M-k}) + i (X_k - conj X_{M-k}) w_n^-k
            // (k = 0 pairs X_0 with X_M, the row M of the input)
            C<R> A = XK(p), B = conj(XMK(p));
            if (k == 0) { A.im = R(0); B.im = R(0); }
            else if (a.adjoint) { A = scl(A, R(0.5)); B = scl(B, R(0.5)); }  // X_k / w_k
            const C<R> E = A + B;
            const C<R> O = (A - B) * conj(aux[k]);
            z = E + mul_pi(O);
          } else if constexpr (KIND == K_DCT2C) {
            // Makhoul reorder: v_j = x_{2j} (j < ceil(M/2)), v_{M-1-j} = x_{2j+1}
            z = X(2 * k < M ? 2 * k : 2 * M - 2 * k - 1);
          } else if constexpr (KIND == K_DST2C) {
            // the reorder of (-1)^j x_j: the odd-indexed points change sign
            z = 2 * k < M ? X(2 * k) : scl(X(2 * M - 2 * k - 1), R(-1));
          } else if constexpr (KIND == K_DCT1C) {
            // even extension
            z = X(2 * k <= M ? k : M - k);
          } else if constexpr (KIND == K_DST1C) {
            // odd extension, zeros at 0 and M/2
            z = 2 * k < M ? (k == 0 ? mk(R(0), R(0)) : X(k - 1))
                          : (2 * k == M ? mk(R(0), R(0)) : scl(X(M - k - 1), R(-1)));
          } else if constexpr (KIND == K_DCT4C || KIND == K_DST4C) {
            // point 2m + s: the half z_m of part s (0 real, 1 imaginary);
            // the DST reads the line reversed
            // the DST reads the line reversed.  m = l/2 + (L/2) p, so
            // c_m = c_{l/2} w_M^{(L/4) p}: one table entry per lane, the
            // others wave-uniform (scalar loads)
            const int m2 = k & ~1;
            const C<R> x0 = X(m2), x1 = X(M - 1 - m2);
            const C<R> xa = KIND == K_DCT4C ? x0 : x1, xb = KIND == K_DCT4C ? x1 : x0;
            const C<R> cm = aux[l >> 1] * tw[(L / 4) * p];
            z = cm * ((k & 1) ? mk(xa.im, xb.im) : mk(xa.re, xb.re));
          } else if constexpr (KIND == K_DST3C) {
            // the pre-twiddle of the reversed line x'_j = x_{M-1-j}
            const C<R> xk = X(M - 1 - k);
            const C<R> xr = k == 0 ? mk(R(0), R(0)) : X(k - 1);
            z = conj(aux[k]) * (xk + mul_mi(xr));
          } else {
            // V_k = w_4M^-k (x_k - i x_{M-k}), x_M = 0
            const C<R> xk = XK(p);
            const C<R> xr = k == 0 ? mk(R(0), R(0)) : XMK(p);
            z = conj(aux[k]) * (xk + mul_mi(xr));
          }
          v[q][p] = conj_in ? conj(z) : z;
        }
      } else if constexpr (LD == IO_TILE) {
#pragma unroll
        for (int p = 0; p < P; ++p) {
          const C<R> x = TK(p, t, 0);
          v[q][p] = conj_in ? conj(x) : x;
        }
        if constexpr (OPK) {
          if (a.op_side == 1) {   // m x, applied before the conjugation of BFFT
            const OpLine<R> ol = op_line<R>(a, otile, om0 + t);
#pragma unroll
            for (int p = 0; p < P; ++p) {
              const C<R> f = op_factor<R>(a, ol, l + L * p);
              v[q][p] = v[q][p] * (conj_in ? conj(f) : f);
            }
          }
        }
      } else {
        // contiguous line: complex elements, or real pairs (x[2m], x[2m+1]) for r2c
        const C<R>* in = KIND == K_R2C
                             ? (const C<R>*)((const R*)a.in + c.in_base + (int64_t)t * a.lsin[0])
                             : (const C<R>*)a.in + c.in_base + (int64_t)t * a.lsin[0];
        if (active) {
          // real pairs: the pair's two real lines, lsin[0] / 2 apart
          [[maybe_unused]] const R* ia = (const R*)a.in + c.in_base + (int64_t)t * a.lsin[0];
#pragma unroll
          for (int p = 0; p < P; ++p) {
            C<R> x;
            if constexpr (PFL) x = mk(nl[q][p].x, nl[q][p].y);
            else if constexpr (RP) x = mk(ia[l + L * p], ia[a.lsin[0] / 2 + l + L * p]);
            else x = in[lrow + ROW(p)];
            v[q][p] = conj_in ? conj(x) : x;
          }
          if constexpr (OPK) {
            if (a.op_side == 1) {   // m x, applied before the conjugation of BFFT
              const OpLine<R> ol = op_line<R>(a, otile, om0 + t);
#pragma unroll
              for (int p = 0; p < P; ++p) {
                const C<R> f = op_factor<R>(a, ol, l + L * p);
                v[q][p] = v[q][p] * (conj_in ? conj(f) : f);
              }
            }
          }
        }
      }
    }
    if constexpr (PFL) {
      const int64_t nx = tl + gridDim.x;
      if (nx < a.ntiles) {
        tile_init(a, nx, cn);
        fetch_lines(cn);
      }
    }
    if constexpr (LD == IO_TILE) __syncthreads();   // line slots alias the tile
    else if constexpr (GATHER) line_sync<L>();

    // ---------------- FFT, post-process, line stores ----------------
#pragma unroll
    for (int q = 0; q < LPW; ++q) {
      const int t = q * TQ + g;
      const bool active = t < c.Tact;
      if (q > 0) line_sync<L>();
      [[maybe_unused]] R scq = sc;   // output scaling of this line
      if constexpr (JM) {
        // this line's points J(l, p), from the tile or from global memory
        if constexpr (LD == IO_TILE) {
#pragma unroll
          for (int p = 0; p < P; ++p) {
            const C<R> x = TK(p, t, 0);
            v[q][p] = conj_in ? conj(x) : x;
          }
        } else {
          const C<R>* in = (const C<R>*)a.in + c.in_base + (int64_t)t * a.lsin[0];
          if (active) {
#pragma unroll
            for (int p = 0; p < P; ++p) {
              const C<R> x = in[lrow + ROW(p)];
              v[q][p] = conj_in ? conj(x) : x;
            }
          }
        }
      }
      if constexpr (MF == 4 && M == 512) {
        fft512_xl((C<float>*)v[q], (C<float>*)line, l, (const C<float>*)tw);
      } else if constexpr (MF == 4 && M == 1024) {
        fft1024_xl((C<float>*)v[q], (C<float>*)line, l, (const C<float>*)tw);
      } else if constexpr (MF == 4) {
        fft256_xl((C<float>*)v[q], (C<float>*)line, l, (const C<float>*)tw);
      } else if constexpr (MF != 0) {
        int e;
        if constexpr (JM)
          e = fft1024_mfma32((C<float>*)v[q], l, (const C<float>*)aux, ftab);
        else
          e = fft1024_mfma<MF, L>((C<float>*)v[q], (C<float>*)line, l, (const C<float>*)tw, mfA);
        // remove the entry scaling 2^e: folded into the output factor when
        // that stays a normal number (sc >= 2^-60 here), else applied apart
        if (e >= -60 && e <= 60) {
          scq = ldexpf(sc, -e);
        } else {
#pragma unroll
          for (int p = 0; p < P; ++p) v[q][p] = mk(ldexpf(v[q][p].re, -e), ldexpf(v[q][p].im, -e));
        }
      } else {
        Cfg<M, W>::template fft<R>(v[q], line, l, tw);
      }
      if constexpr (POST) {
        // the transform lands in natural order in the line's LDS slot
        line_sync<L>();
#pragma unroll
        for (int p = 0; p < P; ++p)
          LK(p) = KIND == K_DCT3C || KIND == K_DST3C ? conj(v[q][p]) : v[q][p];
        line_sync<L>();
        C<R> Y[NX];
#pragma unroll
        for (int p = 0; p < NX; ++p) {
          const int k = p < P ? l + L * p : M;
          if constexpr (KIND == K_DST2C) {
            // DCT-II post-processing at the reversed index kk = M - 1 - k
            const int kk = M - 1 - k;
            const C<R> Vk = line[pad8(kk)];
            const C<R> Vn = conj(kk != 0 ? line[pad8(k + 1)] : line[0]);
            const C<R> Va = scl(Vk + Vn, R(0.5));
            const C<R> Vb = scl(mul_mi(Vk - Vn), R(0.5));
            const C<R> w = aux[kk];
            Y[p] = scl(mk(w.re * Va.re - w.im * Va.im, w.re * Vb.re - w.im * Vb.im), R(2) * sc);
          } else if constexpr (KIND == K_DST3C) {
            // inverse Makhoul reorder, then the sign (-1)^k
            Y[p] = scl(line[pad8((k & 1) ? M - 1 - (k >> 1) : (k >> 1))], (k & 1) ? -sc : sc);
          } else if constexpr (KIND == K_DCT4C || KIND == K_DST4C) {
            // output k = 2m (k even) or M - 1 - 2m (k odd) of each part, from
            // W_m and W_{m+M/2}; e_m = aux[M/2 + m] = e^{i pi m / M}, formed
            // as e_{m(l)} w_M^{-+(L/4) p} (m runs up by (L/2) p for even k,
            // down for odd k)
            const int m = (k & 1) ? M / 2 - 1 - (k >> 1) : (k >> 1);
            const C<R> Wm = line[pad8(m)], Wh = line[pad8(m + M / 2)];
            const C<R> wq = tw[(L / 4) * p];
            const C<R> e = aux[M / 2 + ((l & 1) ? M / 2 - 1 - (l >> 1) : (l >> 1))] *
                           ((l & 1) ? wq : conj(wq));
            const C<R> ya = conj(e) * (Wm + Wh), yb = e * (Wm - Wh);
            const R s = (k & 1) ? (KIND == K_DST4C ? sc : -sc) : sc;
            Y[p] = scl((k & 1) ? mk(ya.im, yb.im) : mk(ya.re, yb.re), s);
          } else if constexpr (KIND == K_DHTC) {
            // the transforms A, B of the real and imaginary parts, then
            // Re - Im of each
            const C<R> Z = LK(p);
            const C<R> Zc = conj(k != 0 ? LMK(p) : line[0]);
            const C<R> A = scl(Z + Zc, R(0.5));
            const C<R> B = scl(mul_mi(Z - Zc), R(0.5));
            Y[p] = scl(mk(A.re - A.im, B.re - B.im), sc);
          } else if constexpr (KIND == K_R2C) {
            // X_k = (Z_k + conj Z_{M-k})/2 - i w_n^k (Z_k - conj Z_{M-k})/2, k = 0..M
            const C<R> Z = p < P ? LK(p) : line[0];
            const C<R> Zc = conj(p < P && k != 0 ? LMK(p) : line[0]);
            const C<R> E = scl(Z + Zc, R(0.5));
            const C<R> O = scl(mul_mi(Z - Zc), R(0.5));
            Y[p] = E + aux[k] * O;
            if (a.adjoint && k != 0 && k != M) Y[p] = scl(Y[p], R(2));    // w_k X_k
          } else if constexpr (KIND == K_DCT2C) {
            // split V into the transforms of the real and imaginary parts,
            // then y_k = 2 Re(w_4M^k V_k) for each
            const C<R> Vk = LK(p);
            const C<R> Vn = conj(k != 0 ? LMK(p) : line[0]);
            const C<R> Va = scl(Vk + Vn, R(0.5));
            const C<R> Vb = scl(mul_mi(Vk - Vn), R(0.5));
            const C<R> w = aux[k];
            Y[p] = scl(mk(w.re * Va.re - w.im * Va.im, w.re * Vb.re - w.im * Vb.im), R(2) * sc);
          } else {
            // inverse Makhoul reorder: y_{2m} = v_m, y_{2m+1} = v_{M-1-m}
            Y[p] = scl(line[pad8((k & 1) ? M - 1 - (k >> 1) : (k >> 1))], sc);
          }
        }
#pragma unroll
        for (int p = 0; p < NX; ++p) v[q][p] = Y[p];
        if constexpr (ST == IO_LINE && RP) {
          // the pair's two real lines, lsout[0] / 2 apart
          R* oa = (R*)a.out + c.out_base + (int64_t)t * a.lsout[0];
          R* ob = oa + a.lsout[0] / 2;
          if (active) {
#pragma unroll
            for (int p = 0; p < P; ++p) {
              st_real<NT != 0>(&oa[l + L * p], Y[p].re);
              st_real<NT != 0>(&ob[l + L * p], Y[p].im);
            }
          }
        } else if constexpr (ST == IO_LINE) {
          C<R>* out = (C<R>*)a.out + c.out_base + (int64_t)t * a.lsout[0];
          if (active) {
#pragma unroll
            for (int p = 0; p < P; ++p) st_line<NT != 0>(&out[l + L * p], Y[p]);
            if constexpr (KIND == K_R2C)
              if (l == 0) st_line<NT != 0>(&out[M], Y[P]);
          }
        }
      } else if constexpr (KIND == K_C2R) {
        // z_m -> real pairs x[2m] = Re z_m, x[2m+1] = Im z_m
        C<R>* out = (C<R>*)((R*)a.out + c.out_base + (int64_t)t * a.lsout[0]);
        if (active) {
#pragma unroll
          for (int p = 0; p < P; ++p)
            st_line<NT != 0>(&out[l + L * p], scl(conj(v[q][p]), sc));
        }
      } else {
#pragma unroll
        for (int p = 0; p < P; ++p)
          v[q][p] = scl(conj_out || ROWPERM ? conj(v[q][p]) : v[q][p], scq);
        if constexpr (KIND == K_DST1C) {
#pragma unroll
          for (int p = 0; p < P; ++p) v[q][p] = mul_pi(v[q][p]);
        }
        if constexpr (OPK) {
          if (a.op_side == 2) {
            const OpLine<R> ol = op_line<R>(a, otile, om0 + t);
#pragma unroll
            for (int p = 0; p < P; ++p) v[q][p] = v[q][p] * op_factor<R>(a, ol, l + L * p);
          }
        }
        if constexpr (EXT && ST == IO_LINE) {
          // rows OFF .. OFF + ROWS_OUT - 1 of the extended line's transform
          C<R>* out = (C<R>*)a.out + c.out_base + (int64_t)t * a.lsout[0];
          R* oa = (R*)a.out + c.out_base + (int64_t)t * a.lsout[0];
          R* ob = oa + a.lsout[0] / 2;
          if (active) {
#pragma unroll
            for (int p = 0; p < P; ++p) {
              const int k = l + L * p;
              if (k >= OFF && k < OFF + ROWS_OUT) {
                if constexpr (RP) {
                  st_real<NT != 0>(&oa[k - OFF], v[q][p].re);
                  st_real<NT != 0>(&ob[k - OFF], v[q][p].im);
                } else {
                  st_line<NT != 0>(&out[k - OFF], v[q][p]);
                }
              }
            }
          }
        } else if constexpr (ST == IO_LINE) {
          C<R>* out = (C<R>*)a.out + c.out_base + (int64_t)t * a.lsout[0];
          if (active) {
            if (OPK && a.op_acc) {
#pragma unroll
              for (int p = 0; p < P; ++p) out[l + L * p] = v[q][p] + out[l + L * p];
            } else {
#pragma unroll
              for (int p = 0; p < P; ++p) st_line<NT != 0>(&out[lrow + ROW(p)], v[q][p]);
            }
          }
        } else if constexpr (JM && NSP == 1) {
          // the whole tile is in LDS: the line goes there now and its
          // registers are free for the next line
#pragma unroll
          for (int p = 0; p < P; ++p) TK(p, t, 0) = v[q][p];
        }
      }
    }

    // ---------------- transposed store ----------------
    if constexpr (ST == IO_TILE) {
      // chunk h holds rows k = l + L p with p in [h PCH, (h+1) PCH), i.e.
      // k in [h RCH, (h+1) RCH); the r2c row M rides with the last chunk
      constexpr int RCH = M / NSP, PCH = P / NSP;
#pragma unroll
      for (int h = 0; h < NSP; ++h) {
        __syncthreads();   // line slots / the previous chunk are read
#pragma unroll
        for (int q = 0; q < LPW; ++q) {
          if constexpr (JM && NSP == 1) break;   // written line by line above
          const int t = q * TQ + g;
          if constexpr (EXT) {
#pragma unroll
            for (int p = 0; p < P; ++p) {
              const int k = l + L * p;
              if (k >= OFF && k < OFF + ROWS_OUT) tile[tile_idx<R, T, SW>(k - OFF, t)] = v[q][p];
            }
            continue;
          }
          if constexpr (ROWPERM) {
            // point m to row 2m (m < M/2) or 2M - 1 - 2m; DST-III: odd rows negated
#pragma unroll
            for (int p = 0; p < P; ++p) {
              const int m = l + L * p;
              const int row = 2 * m < M ? 2 * m : 2 * M - 1 - 2 * m;
              tile[tile_idx<R, T, SW>(row, t)] =
                  KIND == K_DST3C && (row & 1) ? scl(v[q][p], R(-1)) : v[q][p];
            }
            continue;
          }
#pragma unroll
          for (int p = h * PCH; p < (h + 1) * PCH; ++p) TK(p, t, -h * RCH) = v[q][p];
          if constexpr (KIND == K_R2C)
            if (h == NSP - 1 && l == 0) tile[tile_idx<R, T, SW>(M - h * RCH, t)] = v[q][P];
        }
        __syncthreads();
        const int64_t ob = c.out_base + (int64_t)(h * RCH) * a.sk_out;
        constexpr int LAST_ROWS = ROWS_OUT - (NSP - 1) * RCH;
        if (OPK && a.op_acc) {
          if (h == NSP - 1) tile_store<R, T, LAST_ROWS, NTHR, SW, NT, true>(tile, a, ob, tid);
          else tile_store<R, T, RCH, NTHR, SW, NT, true>(tile, a, ob, tid);
        } else {
          if (h == NSP - 1) tile_store<R, T, LAST_ROWS, NTHR, SW, NT>(tile, a, ob, tid);
          else tile_store<R, T, RCH, NTHR, SW, NT>(tile, a, ob, tid);
        }
      }
    }
    __syncthreads();
  }
}

// ---- registry --------------------------------------------------------------------
struct Entry {
  int real_bytes, M, kind, ld, st, T, lpw, threads, mf, nsp, w, occ, opk, pf;
  size_t lds;
  const void* fn;
  void (*launch)(dim3, dim3, size_t, hipStream_t, const PassArgs&);
};

template <typename R, int M, int T, int LPW, int KIND, int LD, int ST, int MF, int NSP, int W,
          int OCC, int OPK, int PF>
static void launch_one(dim3 g, dim3 b, size_t lds, hipStream_t s, const PassArgs& a) {
  hipLaunchKernelGGL((k_fast<R, M, T, LPW, KIND, LD, ST, MF, NSP, W, OCC, OPK, PF, PFFT_FAST_NT>),
                     g, b, lds, s, a);
}

template <typename R, int M, int T, int LPW, int KIND, int LD, int ST, int MF = 0, int NSP = 1,
          int W = 1, int OCC = 0, int OPK = 0, int PF = 0>
static void add(std::vector<Entry>& v) {
  Entry e;
  e.real_bytes = sizeof(R);
  e.M = M;
  e.kind = KIND;
  e.ld = LD;
  e.st = ST;
  e.T = T;
  e.lpw = LPW;
  e.mf = MF;
  e.nsp = NSP;
  e.w = W;
  e.occ = OCC;
  e.opk = OPK;
  e.pf = PF;
  e.threads = T / LPW * Cfg<M, W>::L;
  e.lds = lds_bytes<R, M, T, LPW, KIND, LD, ST, NSP, PF, MF>();
  e.fn = (const void*)k_fast<R, M, T, LPW, KIND, LD, ST, MF, NSP, W, OCC, OPK, PF, PFFT_FAST_NT>;
  e.launch = launch_one<R, M, T, LPW, KIND, LD, ST, MF, NSP, W, OCC, OPK, PF>;
  v.push_back(e);
}

// Candidates per (precision, length, kind, load, store).  The first of each
// group is the default when autotuning is off; the autotuner times them all.
// c2c 1024-point passes with the radix-16 stages on MFMA (precision flags)
template <int MF>
static void add_mfma(std::vector<Entry>& v) {
  add<float, 1024, 4, 1, K_C2C, IO_LINE, IO_LINE, MF>(v);
  add<float, 1024, 8, 1, K_C2C, IO_LINE, IO_TILE, MF>(v);
  add<float, 1024, 16, 2, K_C2C, IO_LINE, IO_TILE, MF, 2>(v);
  add<float, 1024, 8, 1, K_C2C, IO_TILE, IO_LINE, MF>(v);
  // two wavefronts per line: one 32-butterfly block per wavefront
  add<float, 1024, 2, 1, K_C2C, IO_LINE, IO_LINE, MF, 1, 2>(v);
  add<float, 1024, 4, 1, K_C2C, IO_LINE, IO_LINE, MF, 1, 2>(v);
  add<float, 1024, 16, 4, K_C2C, IO_LINE, IO_TILE, MF, 4, 2>(v);
  add<float, 1024, 16, 4, K_C2C, IO_LINE, IO_TILE, MF, 2, 2>(v);
  add<float, 1024, 8, 2, K_C2C, IO_LINE, IO_TILE, MF, 2, 2>(v);
  add<float, 1024, 8, 2, K_C2C, IO_TILE, IO_LINE, MF, 1, 2>(v);
}

// c2c 1024-point passes with 1024 = 32 x 32 on MFMA in split-FP16, the line
// in registers (one wavefront per line, no line slots in LDS)
static void add_mfma32(std::vector<Entry>& v) {
  add<float, 1024, 4, 1, K_C2C, IO_LINE, IO_LINE, 3>(v);
  add<float, 1024, 2, 1, K_C2C, IO_LINE, IO_LINE, 3>(v);
  add<float, 1024, 8, 1, K_C2C, IO_LINE, IO_LINE, 3>(v);
  // register budgets for 4 and 3 waves per SIMD
  add<float, 1024, 4, 1, K_C2C, IO_LINE, IO_LINE, 3, 1, 1, 4>(v);
  add<float, 1024, 4, 1, K_C2C, IO_LINE, IO_LINE, 3, 1, 1, 3>(v);
  add<float, 1024, 8, 1, K_C2C, IO_LINE, IO_TILE, 3, 2, 1, 4>(v);
  add<float, 1024, 8, 1, K_C2C, IO_LINE, IO_TILE, 3, 2, 1, 3>(v);
  // 16-line tiles (128-byte runs) with several lines per wavefront, one line
  // in registers at a time: each goes into the whole-tile LDS buffer as soon
  // as it is transformed
  add<float, 1024, 16, 2, K_C2C, IO_LINE, IO_TILE, 3>(v);
  add<float, 1024, 16, 4, K_C2C, IO_LINE, IO_TILE, 3>(v);
  add<float, 1024, 8, 2, K_C2C, IO_LINE, IO_TILE, 3>(v);
  add<float, 1024, 16, 2, K_C2C, IO_TILE, IO_LINE, 3>(v);
  add<float, 1024, 16, 1, K_C2C, IO_LINE, IO_TILE, 3, 2>(v);
  add<float, 1024, 16, 1, K_C2C, IO_LINE, IO_TILE, 3, 4>(v);
  add<float, 1024, 8, 1, K_C2C, IO_LINE, IO_TILE, 3>(v);
  add<float, 1024, 8, 1, K_C2C, IO_LINE, IO_TILE, 3, 2>(v);
  add<float, 1024, 8, 1, K_C2C, IO_TILE, IO_LINE, 3>(v);
  add<float, 1024, 16, 1, K_C2C, IO_TILE, IO_LINE, 3>(v);
}

// c2c passes that carry a spectral operator (pfft_execute_op): one tile shape
// per load/store form
template <typename R, int M>
static void add_op(std::vector<Entry>& v) {
  constexpr int TL = 256 / Cfg<M>::L;
  constexpr int TT = sizeof(R) == 4 && M < 1024 ? 16 : 8;
  add<R, M, TL, 1, K_C2C, IO_LINE, IO_LINE, 0, 1, 1, 0, 1>(v);
  add<R, M, TL * 2, 2, K_C2C, IO_LINE, IO_LINE, 0, 1, 1, 0, 1>(v);
  add<R, M, TT, 1, K_C2C, IO_LINE, IO_TILE, 0, 1, 1, 0, 1>(v);
  add<R, M, TT, 2, K_C2C, IO_LINE, IO_TILE, 0, 1, 1, 0, 1>(v);
  if constexpr (M >= 256 && (M & (M - 1)) == 0)
    add<R, M, TT, 1, K_C2C, IO_LINE, IO_TILE, 0, 2, 1, 0, 1>(v);
  add<R, M, TT, 1, K_C2C, IO_TILE, IO_LINE, 0, 1, 1, 0, 1>(v);
}

// R2R kinds (complex data, and real data as line pairs): contiguous lines,
// transposed stores, transposed loads
template <typename R, int M, int K, bool MORE>
static void add_r2r_kind(std::vector<Entry>& v) {
  constexpr int TL = 256 / Cfg<M>::L;
  constexpr int TT = sizeof(R) == 4 ? 16 : 8;
  add<R, M, TL, 1, K, IO_LINE, IO_LINE>(v);
  add<R, M, TT, 1, K, IO_LINE, IO_TILE>(v);
  if constexpr (MORE) add<R, M, TT, 2, K, IO_LINE, IO_TILE>(v);
  add<R, M, TT, 1, K, IO_TILE, IO_LINE>(v);
}

template <typename R, int M>
static void add_dct(std::vector<Entry>& v) {
  add_r2r_kind<R, M, K_DCT2C, true>(v);
  add_r2r_kind<R, M, K_DCT3C, true>(v);
  add_r2r_kind<R, M, K_DCT2C | K_REAL_PAIRS, true>(v);
  add_r2r_kind<R, M, K_DCT3C | K_REAL_PAIRS, true>(v);
  add_r2r_kind<R, M, K_DST2C, false>(v);
  add_r2r_kind<R, M, K_DST3C, false>(v);
  add_r2r_kind<R, M, K_DHTC, false>(v);
  add_r2r_kind<R, M, K_DST2C | K_REAL_PAIRS, false>(v);
  add_r2r_kind<R, M, K_DST3C | K_REAL_PAIRS, false>(v);
  add_r2r_kind<R, M, K_DHTC | K_REAL_PAIRS, false>(v);
  add_r2r_kind<R, M, K_DCT4C, true>(v);
  add_r2r_kind<R, M, K_DST4C, false>(v);
  add_r2r_kind<R, M, K_DCT4C | K_REAL_PAIRS, true>(v);
  add_r2r_kind<R, M, K_DST4C | K_REAL_PAIRS, false>(v);
}

// 512- and 1024-point extended lines: the cross-lane exchange, and for 1024
// points 16-line tiles (128-byte runs) with two lines per wavefront
template <int M, int K>
static void add_ext_more(std::vector<Entry>& v) {
  constexpr int TL = 256 / Cfg<M>::L;
  constexpr int TT = M < 1024 ? 16 : 8;
  add<float, M, TL, 1, K, IO_LINE, IO_LINE, 4>(v);
  add<float, M, TT, 1, K, IO_LINE, IO_TILE, 4>(v);
  add<float, M, TT, 1, K, IO_TILE, IO_LINE, 4>(v);
  if constexpr (M == 1024) {
    add<float, M, 16, 2, K, IO_LINE, IO_TILE>(v);
    add<float, M, 16, 2, K, IO_TILE, IO_LINE>(v);
    add<float, M, 16, 2, K, IO_LINE, IO_TILE, 4>(v);
    add<float, M, 16, 2, K, IO_TILE, IO_LINE, 4>(v);
  }
}

// DCT-I / DST-I (complex data and real line pairs) as M-point FFTs of the
// extended lines
template <typename R, int M>
static void add_ext(std::vector<Entry>& v) {
  constexpr int TL = 256 / Cfg<M>::L;
  constexpr int TT = sizeof(R) == 4 && M < 1024 ? 16 : 8;
  add<R, M, TL, 1, K_DCT1C, IO_LINE, IO_LINE>(v);
  add<R, M, TT, 1, K_DCT1C, IO_LINE, IO_TILE>(v);
  add<R, M, TT, 1, K_DCT1C, IO_TILE, IO_LINE>(v);
  add<R, M, TL, 1, K_DST1C, IO_LINE, IO_LINE>(v);
  add<R, M, TT, 1, K_DST1C, IO_LINE, IO_TILE>(v);
  add<R, M, TT, 1, K_DST1C, IO_TILE, IO_LINE>(v);
  add<R, M, TL, 1, K_DCT1C | K_REAL_PAIRS, IO_LINE, IO_LINE>(v);
  add<R, M, TT, 1, K_DCT1C | K_REAL_PAIRS, IO_LINE, IO_TILE>(v);
  add<R, M, TT, 1, K_DCT1C | K_REAL_PAIRS, IO_TILE, IO_LINE>(v);
  add<R, M, TL, 1, K_DST1C | K_REAL_PAIRS, IO_LINE, IO_LINE>(v);
  add<R, M, TT, 1, K_DST1C | K_REAL_PAIRS, IO_LINE, IO_TILE>(v);
  add<R, M, TT, 1, K_DST1C | K_REAL_PAIRS, IO_TILE, IO_LINE>(v);
  if constexpr (sizeof(R) == 4 && M >= 512) {
    add_ext_more<M, K_DCT1C>(v);
    add_ext_more<M, K_DST1C>(v);
    add_ext_more<M, K_DCT1C | K_REAL_PAIRS>(v);
    add_ext_more<M, K_DST1C | K_REAL_PAIRS>(v);
  }
}

template <typename R, int M>
static void add_size(std::vector<Entry>& v) {
  constexpr int TL = 256 / Cfg<M>::L;       // contiguous passes: 256 threads
  constexpr int TT = sizeof(R) == 4 ? 16 : 8;  // transposed runs of 128 bytes
  add<R, M, TL, 1, K_C2C, IO_LINE, IO_LINE>(v);
  add<R, M, TL * 2, 2, K_C2C, IO_LINE, IO_LINE>(v);
  add<R, M, TL, 1, K_R2C, IO_LINE, IO_LINE>(v);
  add<R, M, TL, 1, K_C2R, IO_LINE, IO_LINE>(v);
  if constexpr (M < 1024) {
    add<R, M, TT, 2, K_C2C, IO_LINE, IO_TILE>(v);
    add<R, M, TT, 1, K_C2C, IO_LINE, IO_TILE>(v);
    add<R, M, TT, 4, K_C2C, IO_LINE, IO_TILE>(v);
    add<R, M, TT, 2, K_C2C, IO_TILE, IO_LINE>(v);
    add<R, M, TT, 1, K_C2C, IO_TILE, IO_LINE>(v);
    add<R, M, TT, 4, K_C2C, IO_TILE, IO_LINE>(v);
    add<R, M, TT, 2, K_R2C, IO_LINE, IO_TILE>(v);
    add<R, M, TT, 1, K_R2C, IO_LINE, IO_TILE>(v);
    add<R, M, TT, 2, K_C2R, IO_TILE, IO_LINE>(v);
    add<R, M, TT, 1, K_C2R, IO_TILE, IO_LINE>(v);
    // transposed loads with the next tile prefetched (persistent grid)
    add<R, M, TT, 1, K_C2R, IO_TILE, IO_LINE, 0, 1, 1, 0, 0, 1>(v);
    if constexpr (sizeof(R) == 4 && (M == 256 || M == 512)) {
      // register budgets for more waves in flight on the r2c / c2r of 512 and
      // 1024 reals (small spills, kept only where the tuner finds them faster)
      add<R, M, TT, 1, K_C2R, IO_TILE, IO_LINE, 0, 1, 1, 6>(v);
      add<R, M, TT, 1, K_R2C, IO_LINE, IO_TILE, 0, 1, 1, 8>(v);
    }
    // line loads of the next tile prefetched (persistent grid); on c2c lines
    // the loop-top wait also drains the previous tile's stores, which costs
    // more than the prefetch gains, so only the r2c form is kept
    if constexpr (M >= 64 && (M & (M - 1)) == 0)
      add<R, M, TT, 1, K_R2C, IO_LINE, IO_TILE, 0, 1, 1, 0, 0, 1>(v);
    if constexpr (M >= 256 && (M & (M - 1)) == 0) {
      // row-chunked transposed stores: half (quarter) of the tile in LDS
      add<R, M, TT, 2, K_C2C, IO_LINE, IO_TILE, 0, 2>(v);
      add<R, M, TT, 1, K_C2C, IO_LINE, IO_TILE, 0, 2>(v);
      add<R, M, TT, 2, K_R2C, IO_LINE, IO_TILE, 0, 2>(v);
      add<R, M, TT, 1, K_R2C, IO_LINE, IO_TILE, 0, 2>(v);
      if constexpr (M == 512) add<R, M, TT, 2, K_C2C, IO_LINE, IO_TILE, 0, 4>(v);
    }
    if constexpr (sizeof(R) == 4) {
      add<R, M, 32, 4, K_C2C, IO_LINE, IO_TILE>(v);
      add<R, M, 32, 4, K_C2C, IO_TILE, IO_LINE>(v);
      if constexpr (M <= 256) {   // larger lines spill at 128 VGPRs
        add<R, M, 32, 2, K_R2C, IO_LINE, IO_TILE>(v);
        add<R, M, 32, 2, K_C2R, IO_TILE, IO_LINE>(v);
      }
      if constexpr (M == 256 || M == 512) {
        // 256-byte runs (32-line tiles) in row chunks, so the tile takes a
        // quarter (a half) of the LDS of the whole-tile form
        add<R, M, 32, 4, K_C2C, IO_LINE, IO_TILE, 0, 4>(v);
        add<R, M, 32, 4, K_C2C, IO_LINE, IO_TILE, 0, 2>(v);
        add<R, M, 32, 2, K_R2C, IO_LINE, IO_TILE, 0, 2>(v);
        add<R, M, 32, 2, K_R2C, IO_LINE, IO_TILE, 0, 4>(v);
      }
    }
  } else {
    add<R, M, 8, 2, K_C2C, IO_LINE, IO_TILE>(v);
    add<R, M, 8, 1, K_C2C, IO_LINE, IO_TILE>(v);
    add<R, M, 8, 2, K_C2C, IO_LINE, IO_TILE, 0, 2>(v);
    add<R, M, 8, 1, K_C2C, IO_LINE, IO_TILE, 0, 2>(v);
    // 128-byte runs: a 16-line tile in row chunks
    add<R, M, 16, 2, K_C2C, IO_LINE, IO_TILE, 0, 2>(v);
    add<R, M, 16, 4, K_C2C, IO_LINE, IO_TILE, 0, 2>(v);
    add<R, M, 16, 4, K_C2C, IO_LINE, IO_TILE, 0, 4>(v);
    if constexpr (sizeof(R) == 4) {
      // two wavefronts per line (8 points per lane)
      add<R, M, 2, 1, K_C2C, IO_LINE, IO_LINE, 0, 1, 2>(v);
      add<R, M, 4, 1, K_C2C, IO_LINE, IO_LINE, 0, 1, 2>(v);
      add<R, M, 16, 4, K_C2C, IO_LINE, IO_TILE, 0, 2, 2>(v);
      add<R, M, 16, 4, K_C2C, IO_LINE, IO_TILE, 0, 4, 2>(v);
      add<R, M, 8, 2, K_C2C, IO_LINE, IO_TILE, 0, 2, 2>(v);
      add<R, M, 8, 2, K_C2C, IO_TILE, IO_LINE, 0, 1, 2>(v);
    }
    add<R, M, 8, 2, K_C2C, IO_TILE, IO_LINE>(v);
    add<R, M, 8, 1, K_C2C, IO_TILE, IO_LINE>(v);
    add<R, M, 8, 2, K_R2C, IO_LINE, IO_TILE>(v);
    add<R, M, 8, 2, K_C2R, IO_TILE, IO_LINE>(v);
  }
}

// FP32 passes with the last stage exchange across lanes (MF = 4): twins of
// the Stockham variants the tuner picks at 512^3
static void add_xlane(std::vector<Entry>& v) {
  add<float, 512, 4, 1, K_C2C, IO_LINE, IO_LINE, 4>(v);
  add<float, 512, 8, 2, K_C2C, IO_LINE, IO_LINE, 4>(v);
  add<float, 512, 16, 1, K_C2C, IO_LINE, IO_TILE, 4>(v);
  add<float, 512, 16, 2, K_C2C, IO_LINE, IO_TILE, 4>(v);
  add<float, 512, 16, 1, K_C2C, IO_LINE, IO_TILE, 4, 2>(v);
  add<float, 512, 16, 2, K_C2C, IO_LINE, IO_TILE, 4, 2>(v);
  add<float, 512, 16, 1, K_C2C, IO_TILE, IO_LINE, 4>(v);
  add<float, 512, 16, 2, K_C2C, IO_TILE, IO_LINE, 4>(v);
  // the r2c / c2r passes of 512 reals (256-point lines) and 256-point c2c
  add<float, 256, 16, 1, K_R2C, IO_LINE, IO_TILE, 4>(v);
  add<float, 256, 16, 2, K_R2C, IO_LINE, IO_TILE, 4>(v);
  add<float, 256, 16, 1, K_C2R, IO_TILE, IO_LINE, 4>(v);
  add<float, 256, 16, 2, K_C2R, IO_TILE, IO_LINE, 4>(v);
  add<float, 256, 16, 1, K_C2R, IO_TILE, IO_LINE, 4, 1, 1, 0, 0, 1>(v);   // prefetching
  add<float, 256, 16, 1, K_R2C, IO_LINE, IO_TILE, 4, 1, 1, 0, 0, 1>(v);
  add<float, 256, 8, 1, K_R2C, IO_LINE, IO_LINE, 4>(v);
  add<float, 256, 8, 1, K_C2R, IO_LINE, IO_LINE, 4>(v);
  add<float, 256, 8, 1, K_C2C, IO_LINE, IO_LINE, 4>(v);
  add<float, 256, 16, 1, K_C2C, IO_LINE, IO_TILE, 4>(v);
  add<float, 256, 16, 1, K_C2C, IO_TILE, IO_LINE, 4>(v);
  // 1024-point c2c (the plain FP32 precision baseline): the tuned shapes
  add<float, 1024, 16, 2, K_C2C, IO_LINE, IO_TILE, 4, 2>(v);
  add<float, 1024, 16, 1, K_C2C, IO_LINE, IO_TILE, 4, 2>(v);
  add<float, 1024, 8, 1, K_C2C, IO_LINE, IO_TILE, 4>(v);
  add<float, 1024, 4, 1, K_C2C, IO_LINE, IO_LINE, 4>(v);
  add<float, 1024, 16, 2, K_C2C, IO_TILE, IO_LINE, 4>(v);
  add<float, 1024, 8, 1, K_C2C, IO_TILE, IO_LINE, 4>(v);
}

// Row-chunked transposed stores for the mixed-radix FP32 lengths: a 16-line
// tile of 320 or 384 rows takes 41-49 KB of LDS beside the line slots, one
// workgroup per CU; in 2 or 4 chunks it takes a half or a quarter of that
template <int M>
static void add_mixed_chunks(std::vector<Entry>& v) {
  add<float, M, 16, 1, K_C2C, IO_LINE, IO_TILE, 0, 2>(v);
  add<float, M, 16, 1, K_C2C, IO_LINE, IO_TILE, 0, 4>(v);
  add<float, M, 16, 2, K_C2C, IO_LINE, IO_TILE, 0, 2>(v);
  add<float, M, 16, 1, K_R2C, IO_LINE, IO_TILE, 0, 2>(v);
  add<float, M, 16, 1, K_R2C, IO_LINE, IO_TILE, 0, 4>(v);
}

// 2048-point FP32 c2c lines (32 points per lane, or 16 with two wavefronts
// per line): 4-line tiles, since the line slots (18 KB each) and the tile
// share the LDS
static void add_2048(std::vector<Entry>& v) {
  add<float, 2048, 4, 1, K_C2C, IO_LINE, IO_LINE>(v);
  add<float, 2048, 2, 1, K_C2C, IO_LINE, IO_LINE, 0, 1, 2>(v);
  add<float, 2048, 4, 1, K_C2C, IO_LINE, IO_TILE>(v);
  add<float, 2048, 4, 1, K_C2C, IO_LINE, IO_TILE, 0, 2>(v);
  add<float, 2048, 4, 1, K_C2C, IO_LINE, IO_TILE, 0, 2, 2>(v);
  add<float, 2048, 4, 1, K_C2C, IO_TILE, IO_LINE>(v);
  add<float, 2048, 4, 1, K_C2C, IO_TILE, IO_LINE, 0, 1, 2>(v);
  // the r2c / c2r of 4096 reals
  add<float, 2048, 4, 1, K_R2C, IO_LINE, IO_LINE>(v);
  add<float, 2048, 4, 1, K_C2R, IO_LINE, IO_LINE>(v);
  add<float, 2048, 4, 1, K_R2C, IO_LINE, IO_TILE>(v);
  add<float, 2048, 4, 1, K_C2R, IO_TILE, IO_LINE>(v);
}

// Float64 1024-point c2c lines with a transposed side: 4-line tiles (64-byte
// runs), since the line slots of 1024 complex doubles (18 KB each) and the
// tile share the LDS
static void add_f64_1024(std::vector<Entry>& v) {
  add<double, 1024, 4, 1, K_C2C, IO_LINE, IO_TILE>(v);
  add<double, 1024, 4, 1, K_C2C, IO_LINE, IO_TILE, 0, 2>(v);
  add<double, 1024, 4, 1, K_C2C, IO_TILE, IO_LINE>(v);
}

// DCT-II / DCT-III (the Chebyshev pair) on the mixed-radix FP32 lengths, and
// on the Float64 lengths without the full R2R set
template <int M, typename R = float>
static void add_dct_mixed(std::vector<Entry>& v) {
  add_r2r_kind<R, M, K_DCT2C, true>(v);
  add_r2r_kind<R, M, K_DCT3C, true>(v);
  add_r2r_kind<R, M, K_DCT2C | K_REAL_PAIRS, true>(v);
  add_r2r_kind<R, M, K_DCT3C | K_REAL_PAIRS, true>(v);
}

// 256- and 512-point DCT-II / III / IV lines with the cross-lane exchange
template <int M>
static void add_dct_xlane(std::vector<Entry>& v) {
  constexpr int TL = 256 / Cfg<M>::L;
  add<float, M, TL, 1, K_DCT2C, IO_LINE, IO_LINE, 4>(v);
  add<float, M, 16, 1, K_DCT2C, IO_LINE, IO_TILE, 4>(v);
  add<float, M, 16, 2, K_DCT2C, IO_LINE, IO_TILE, 4>(v);
  add<float, M, TL, 1, K_DCT3C, IO_LINE, IO_LINE, 4>(v);
  add<float, M, 16, 1, K_DCT3C, IO_LINE, IO_TILE, 4>(v);
  add<float, M, 16, 2, K_DCT3C, IO_LINE, IO_TILE, 4>(v);
  add<float, M, TL, 1, K_DCT2C | K_REAL_PAIRS, IO_LINE, IO_LINE, 4>(v);
  add<float, M, 16, 1, K_DCT2C | K_REAL_PAIRS, IO_LINE, IO_TILE, 4>(v);
  add<float, M, 16, 2, K_DCT2C | K_REAL_PAIRS, IO_LINE, IO_TILE, 4>(v);
  add<float, M, TL, 1, K_DCT3C | K_REAL_PAIRS, IO_LINE, IO_LINE, 4>(v);
  add<float, M, 16, 1, K_DCT3C | K_REAL_PAIRS, IO_LINE, IO_TILE, 4>(v);
  add<float, M, 16, 2, K_DCT3C | K_REAL_PAIRS, IO_LINE, IO_TILE, 4>(v);
  add<float, M, TL, 1, K_DCT4C | K_REAL_PAIRS, IO_LINE, IO_LINE, 4>(v);
  add<float, M, 16, 1, K_DCT4C | K_REAL_PAIRS, IO_LINE, IO_TILE, 4>(v);
  add<float, M, 16, 2, K_DCT4C | K_REAL_PAIRS, IO_LINE, IO_TILE, 4>(v);
  add<float, M, TL, 1, K_DCT4C, IO_LINE, IO_LINE, 4>(v);
  add<float, M, 16, 1, K_DCT4C, IO_LINE, IO_TILE, 4>(v);
  add<float, M, 16, 2, K_DCT4C, IO_LINE, IO_TILE, 4>(v);
}

// Row-chunked transposed stores for the DCT-II / DCT-III of complex data
// (the mixed workload's third dimension)
template <typename R, int M>
static void add_dct_chunks(std::vector<Entry>& v) {
  constexpr int TT = sizeof(R) == 4 ? 16 : 8;
  add<R, M, TT, 1, K_DCT2C, IO_LINE, IO_TILE, 0, 2>(v);
  add<R, M, TT, 1, K_DCT3C, IO_LINE, IO_TILE, 0, 2>(v);
  add<R, M, TT, 2, K_DCT3C, IO_LINE, IO_TILE, 0, 2>(v);
}

static const std::vector<Entry>& registry() {
  static std::vector<Entry> v = [] {
    std::vector<Entry> r;
    add_size<float, 64>(r);
    add_size<float, 128>(r);
    add_size<float, 256>(r);
    add_size<float, 512>(r);
    add_size<float, 1024>(r);
    add_mfma32(r);
    add_mfma<1>(r);
    add_mfma<2>(r);
    add_size<float, 32>(r);
    add_size<float, 48>(r);
    add_size<float, 96>(r);
    add_size<float, 160>(r);
    add_size<float, 192>(r);
    add_size<float, 320>(r);
    add_size<float, 384>(r);
    add_dct<float, 256>(r);
    add_dct<float, 128>(r);
    add_dct<float, 64>(r);
    add_dct<float, 512>(r);
    add_dct<double, 256>(r);
    add_dct<double, 64>(r);
    add_size<double, 32>(r);   // r2c / c2r of 64 reals (the 64^3 Float64 case)
    add_size<double, 64>(r);
    add_size<double, 128>(r);
    add_size<double, 256>(r);
    add_size<double, 512>(r);
    add_op<float, 64>(r);
    add_op<float, 96>(r);
    add_op<float, 128>(r);
    add_op<float, 192>(r);
    add_op<float, 256>(r);
    add_op<float, 320>(r);
    add_op<float, 384>(r);
    add_op<float, 512>(r);
    add_op<float, 1024>(r);
    add_op<double, 64>(r);
    add_op<double, 128>(r);
    add_op<double, 256>(r);
    add_op<double, 512>(r);
    add_xlane(r);
    add_mixed_chunks<96>(r);
    add_mixed_chunks<160>(r);
    add_mixed_chunks<192>(r);
    add_mixed_chunks<320>(r);
    add_mixed_chunks<384>(r);
    add_dct_chunks<float, 256>(r);
    add_dct_chunks<float, 512>(r);
    add_ext<float, 128>(r);
    add_ext<float, 256>(r);
    add_ext<float, 512>(r);
    add_ext<float, 1024>(r);
    add_ext<double, 128>(r);
    add_ext<double, 256>(r);
    add_ext<double, 512>(r);
    add_dct_xlane<256>(r);
    add_dct_xlane<512>(r);
    add_size<float, 640>(r);
    add_size<float, 768>(r);
    add_mixed_chunks<640>(r);
    add_mixed_chunks<768>(r);
    add_dct_mixed<96>(r);
    add_dct_mixed<160>(r);
    add_dct_mixed<192>(r);
    add_dct_mixed<320>(r);
    add_dct_mixed<384>(r);
    add_size<double, 1024>(r);
    add_f64_1024(r);
    add_dct_mixed<128, double>(r);
    add_dct_mixed<512, double>(r);
    add_size<double, 96>(r);
    add_size<double, 192>(r);
    add_size<double, 384>(r);
    add_2048(r);
    return r;
  }();
  return v;
}

}  // namespace fk

using namespace fk;

#if PFFT_FAST_NT
// the kernel table with nontemporal output stores, read by the other object
const std::vector<Entry>& fast_registry_nt() { return registry(); }
#else
const std::vector<Entry>& fast_registry_nt();

// Variant ids: registry index + 1, with kOrderBit set for the alternative
// tile order (tile_split order 1) of a transposing kernel, kVariantReverse
// for the descending tile direction and kVariantStream for the twin kernel
// whose output stores carry the nontemporal hint.
constexpr int kOrderBit = 1 << 20;
static int variant_index(int variant) { return (variant & (kOrderBit - 1)) - 1; }
static_assert(kVariantReverse > kOrderBit && kVariantStream > kVariantReverse,
              "variant id bits");

// The twin table matches the plain one entry for entry (both objects build it
// from the same list); checked once, as a mismatch would launch the wrong
// kernel shape.
static bool twins_match() {
  static const bool ok = [] {
    const std::vector<Entry>& a = registry();
    const std::vector<Entry>& b = fast_registry_nt();
    if (a.size() != b.size()) return false;
    for (size_t i = 0; i < a.size(); ++i) {
      const Entry& x = a[i];
      const Entry& y = b[i];
      if (x.real_bytes != y.real_bytes || x.M != y.M || x.kind != y.kind || x.ld != y.ld ||
          x.st != y.st || x.T != y.T || x.lpw != y.lpw || x.threads != y.threads ||
          x.mf != y.mf || x.nsp != y.nsp || x.w != y.w || x.occ != y.occ || x.opk != y.opk ||
          x.pf != y.pf || x.lds != y.lds || y.fn == x.fn)
        return false;
    }
    return true;
  }();
  return ok;
}

// the table entry that runs a variant (the caller has range-checked it)
static const Entry& entry_of(int variant) {
  const std::vector<Entry>& reg =
      (variant & kVariantStream) && twins_match() ? fast_registry_nt() : registry();
  return reg[variant_index(variant)];
}

static bool all_even(const int64_t* s, int from, int D) {
  for (int d = from; d < D; ++d)
    if (s[d] & 1) return false;
  return true;
}

// Specialised kernels that can run a pass (first = default choice).
// The arguments of an R2R pass on real data seen as complex lines, each made
// of the real lines (2j, 2j + 1) along line dim 0 (K_REAL_PAIRS).  On a side
// where the lines are contiguous (element stride 1) the pair's lines stay
// apart: the side keeps real units with the pair stride 2 ls[0], and the
// kernel finds the second line ls[0] / 2 after the first.  On a transposed
// side the two lines are adjacent (ls[0] = 1), so one complex element holds
// both, and the side is addressed in complex units.  False where the pairs
// do not fit the layout.
static bool pair_view(PassArgs& a) {
  if (a.ldim[0] % 2) return false;
  const bool in_line = a.sk_in == 1, out_line = a.sk_out == 1;
  if ((!in_line && a.lsin[0] != 1) || (!out_line && a.lsout[0] != 1)) return false;
  auto side = [&](bool line, int64_t* ls, int64_t* sk) {
    if (line) {
      ls[0] *= 2;
      return true;
    }
    if (*sk % 2) return false;
    for (int d = 1; d < a.D; ++d)
      if (ls[d] % 2) return false;
    *sk /= 2;
    for (int d = 1; d < a.D; ++d) ls[d] /= 2;
    return true;
  };
  if (!side(in_line, a.lsin, &a.sk_in) || !side(out_line, a.lsout, &a.sk_out)) return false;
  a.ldim[0] /= 2;
  return true;
}

std::vector<int> fast_candidates(const Plan& p, const PassPlan& pp, bool op) {
  std::vector<int> out;
  if (p.flags & PFFT_FLAG_GENERIC_ONLY) return out;
  PassArgs a = pp.args;
  // the Bluestein form and the four-step twiddle run on the generic pass only
  if (a.bs_M || a.fs_M) return out;
  int kind;
  switch (a.kind) {
    case L_C2C: kind = K_C2C; break;
    case L_R2C_HALF: kind = K_R2C; break;
    case L_C2R_HALF: kind = K_C2R; break;
    case L_DCT2: kind = K_DCT2C; break;
    case L_DCT3: kind = K_DCT3C; break;
    case L_DST2: kind = K_DST2C; break;
    case L_DST3: kind = K_DST3C; break;
    case L_DHT: kind = K_DHTC; break;
    case L_DCT4H: kind = K_DCT4C; break;
    case L_DST4H: kind = K_DST4C; break;
    case L_DCT1: kind = K_DCT1C; break;
    case L_DST1: kind = K_DST1C; break;
    default: return out;
  }
  // R2R of real data: pairs of lines as complex lines
  if (kind >= K_DCT2C && !a.r2r_cplx) {
    if (!pair_view(a)) return out;
    kind |= K_REAL_PAIRS;
  }
  const int rb = p.real_type == PFFT_F32 ? 4 : 8;
  const int M = (int)a.M;
  int ld, st;
  if (a.sk_in == 1) ld = IO_LINE;
  else if (a.lsin[0] == 1 && kind != K_R2C) ld = IO_TILE;
  else return out;
  if (a.sk_out == 1) st = IO_LINE;
  else if (a.lsout[0] == 1 && kind != K_C2R) st = IO_TILE;
  else return out;
  if (ld == IO_TILE && st == IO_TILE) return out;
  // real pairs of r2c input / c2r output need even line offsets
  if (kind == K_R2C && !all_even(a.lsin, 0, a.D)) return out;
  if (kind == K_C2R && !all_even(a.lsout, 0, a.D)) return out;
  // precision flags put the radix-16 stages of 1024-point c2c lines on MFMA
  int want_mf = 0, alt_mf = -1;
  if (kind == K_C2C && M == 1024 && rb == 4) {
    // split-FP16: the 32 x 32 register form (first) and the radix-16 LDS form
    if (p.flags & PFFT_FLAG_MFMA_SPLIT16) want_mf = 3, alt_mf = 1;
    else if (p.flags & PFFT_FLAG_MFMA_BF16X3) want_mf = 2;
  }
  // 256-, 512- and 1024-point FP32 lines: the cross-lane exchange variants
  // beside the Stockham ones (same arithmetic, bit-identical results with
  // the one-wavefront Stockham kernels of the same length)
  if ((M == 512 || M == 256 || (M == 1024 && want_mf == 0)) && rb == 4) alt_mf = 4;
  const std::vector<Entry>& reg = registry();
  for (size_t i = 0; i < reg.size(); ++i) {
    const Entry& x = reg[i];
    if (x.real_bytes != rb || x.M != M || x.kind != kind || x.ld != ld || x.st != st) continue;
    if (x.mf != want_mf && x.mf != alt_mf) continue;
    if (x.opk != (op ? 1 : 0)) continue;
    if (x.lds > 160 * 1024) continue;
    if (ld == IO_TILE || st == IO_TILE) {
      // transposed tiles: full tiles, 16-byte aligned runs
      if (a.ldim[0] % x.T) continue;
      const int64_t sk = ld == IO_TILE ? a.sk_in : a.sk_out;
      const int64_t* ls = ld == IO_TILE ? a.lsin : a.lsout;
      if (rb == 4 && ((sk & 1) || !all_even(ls, 1, a.D))) continue;
      // both tile orders: a far-strided transposed side favours walking the
      // other line dims first
      out.push_back((int)i + 1);
      if (a.D > 1) out.push_back(((int)i + 1) | kOrderBit);
      continue;
    }
    out.push_back((int)i + 1);
  }
  // The first candidate is what an untuned plan (PFFT_FLAG_ESTIMATE,
  // PFFT_FLAG_NO_AUTOTUNE) runs: the cross-lane forms, then one line per
  // wavefront, which the tuner picks on most FP32 passes (the precision
  // forms keep their order)
  if (want_mf == 0) {
    auto pref = [&](int v) {
      const Entry& e = reg[variant_index(v)];
      return (e.mf == 4 ? 0 : 2) + (e.lpw == 1 ? 0 : 1);
    };
    std::stable_sort(out.begin(), out.end(), [&](int x, int y) { return pref(x) < pref(y); });
  }
  return out;
}

void fast_apply(PassPlan& pp, int variant) {
  // without a matching twin table the plain kernel runs (and the name says so)
  if (!twins_match()) variant &= ~kVariantStream;
  const Entry& e = entry_of(variant);
  PassArgs& a = pp.args;
  a.tile_order = (variant & kOrderBit) ? 1 : 0;
  a.tile_rev = (variant & kVariantReverse) ? 1 : 0;
  // real data of an R2R pass as line pairs (the candidates checked it fits)
  const bool rp = (e.kind & K_REAL_PAIRS) != 0;
  if (rp) pair_view(a);
  // contiguous passes: keep the input-adjacent line dim first (its partial
  // last tile costs little) unless it is shorter than a tile (not for line
  // pairs, which run along line dim 0)
  if (!rp && e.ld == IO_LINE && e.st == IO_LINE && a.ldim[0] < e.T) {
    for (int d = 1; d < a.D; ++d) {
      if (a.ldim[d] % e.T == 0) {
        std::swap(a.ldim[0], a.ldim[d]);
        std::swap(a.lsin[0], a.lsin[d]);
        std::swap(a.lsout[0], a.lsout[d]);
        std::swap(a.llog[0], a.llog[d]);
        break;
      }
    }
  }
  a.T = e.T;
  int64_t nt = (a.ldim[0] + e.T - 1) / e.T;
  for (int d = 1; d < a.D; ++d) nt *= a.ldim[d];
  a.ntiles = nt;
  pp.grid = std::min<int64_t>(nt, 1 << 20);
  if (e.pf) {
    // persistent grid: the workgroups resident at once on the device
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, e.fn, e.threads, e.lds) ==
            hipSuccess &&
        cus > 0 && per_cu > 0)
      pp.grid = std::min<int64_t>(nt, (int64_t)cus * per_cu);
    (void)hipGetLastError();
  }
  pp.threads = e.threads;
  pp.lds_bytes = (int64_t)e.lds;
  pp.variant = variant;
}

int fast_prepare(const PassPlan& pp) {
  const Entry& e = entry_of(pp.variant);
  if (e.lds > 64 * 1024) {
    hipError_t r = hipFuncSetAttribute(e.fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)e.lds);
    if (r != hipSuccess)
      return fail(PFFT_ERR_DEVICE, std::string("hipFuncSetAttribute: ") + hipGetErrorString(r));
  }
  return PFFT_OK;
}

int fast_launch(const PassPlan& pp, const PassArgs& a, hipStream_t stream) {
  const std::vector<Entry>& reg = registry();
  const int idx = variant_index(pp.variant);
  if (pp.variant < 1 || idx < 0 || idx >= (int)reg.size())
    return fail(PFFT_ERR_UNSUPPORTED, "no specialised kernel variant " +
                                          std::to_string(pp.variant));
  const Entry& e = entry_of(pp.variant);
  e.launch(dim3((unsigned)pp.grid), dim3(e.threads), e.lds, stream, a);
  return PFFT_OK;
}

const char* fast_name(int variant) {
  static thread_local std::string s;
  const std::vector<Entry>& reg = registry();
  const int idx = variant_index(variant);
  if (variant < 1 || idx < 0 || idx >= (int)reg.size()) return "generic";
  const Entry& e = reg[idx];
  // R2R kinds end in c (complex data) or r (real data, line pairs)
  static const char* kinds[12] = {"c2c",  "r2c",  "c2r", "dct2", "dct3", "dst2",
                                  "dst3", "dht", "dct4", "dst4", "dct1", "dst1"};
  const int kb = e.kind & 15;
  const std::string kname =
      std::string(kinds[kb]) + (kb >= K_DCT2C ? ((e.kind & K_REAL_PAIRS) ? "r" : "c") : "");
  s = std::string("k_fast<") + (e.real_bytes == 4 ? "f32" : "f64") + "," +
      std::to_string(e.M) + "," + kname + "," + (e.ld ? "tile" : "line") + "->" +
      (e.st ? "tile" : "line") + ",T=" + std::to_string(e.T) + ",LPW=" +
      std::to_string(e.lpw) + (e.nsp > 1 ? ",S=" + std::to_string(e.nsp) : std::string()) +
      (e.w > 1 ? ",W=" + std::to_string(e.w) : std::string()) +
      (e.occ > 0 ? ",O=" + std::to_string(e.occ) : std::string()) +
      (e.opk ? ",op" : "") + (e.pf ? ",PF" : "") +
      ((variant & kOrderBit) ? ",ord1" : "") + ((variant & kVariantReverse) ? ",rev" : "") +
      ((variant & kVariantStream) ? ",nt" : "") +
      (e.mf == 1   ? ",mfma-split16"
       : e.mf == 2 ? ",mfma-bf16x3"
       : e.mf == 3 ? ",mfma32-split16"
       : e.mf == 4 ? ",xlane"
                   : "") +
      ">";
  return s.c_str();
}

std::string fast_build_signature() {
  const std::vector<Entry>& reg = registry();
  uint64_t h = 1469598103934665603ull;   // FNV-1a, 64 bit
  for (size_t i = 0; i < reg.size(); ++i) {
    const std::string n = fast_name((int)i + 1);
    for (char ch : n) {
      h ^= (uint8_t)ch;
      h *= 1099511628211ull;
    }
    h ^= 0xff;
    h *= 1099511628211ull;
  }
  char buf[64];
  snprintf(buf, sizeof(buf), "%zu-%016llx", reg.size(), (unsigned long long)h);
  return buf;
}

#endif  // PFFT_FAST_NT

}  // namespace pfft
