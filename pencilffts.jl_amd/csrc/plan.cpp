// PLAN layer of the transform pipeline, restated in C++.
//
// Follows PencilFFTs.jl v0.15.3:
//   transform metadata      src/Transforms/{Transforms,c2c,r2c,r2r,no_transform}.jl
//   GlobalFFTParams         src/global_params.jl:50-62, input type :81-95
//   pencil chain            src/plans.jl:397-521
//   1D plans, binv, scale   src/plans.jl:523-553, product :268
// and then lowers every stage to one fused pass (transform + the transpose
// into the next pencil's memory order).
#include "plan.hpp"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <sstream>

#include "../../include/pfft.h"

namespace pfft {

// ---- errors ----------------------------------------------------------------
static thread_local std::string g_last_error;

void set_error(int code, const std::string& msg) {
  (void)code;
  g_last_error = msg;
}
int fail(int code, const std::string& msg) {
  set_error(code, msg);
  return code;
}
const char* last_error() { return g_last_error.c_str(); }
void clear_error() { g_last_error.clear(); }

// ---- transform metadata ------------------------------------------------------
static const char* kR2RNames[11] = {"R2HC",    "HC2R",    "DHT",     "REDFT00",
                                    "REDFT01", "REDFT10", "REDFT11", "RODFT00",
                                    "RODFT01", "RODFT10", "RODFT11"};

int32_t kind_base(int32_t kind) { return kind & ~PFFT_INPLACE; }
bool kind_inplace(int32_t kind) { return (kind & PFFT_INPLACE) != 0; }

static bool is_r2r(int32_t b) { return b >= PFFT_R2R && b <= PFFT_R2R + PFFT_RODFT11; }
static bool is_c2c(int32_t b) { return b == PFFT_FFT || b == PFFT_BFFT; }
static bool is_c2r(int32_t b) { return b == PFFT_BRFFT_EVEN || b == PFFT_BRFFT_ODD; }

bool kind_valid(int32_t kind, std::string* why) {
  if (kind & ~(PFFT_INPLACE | 0xff)) {
    if (why) *why = "unknown transform kind " + std::to_string(kind);
    return false;
  }
  const int32_t b = kind_base(kind);
  if (b == PFFT_NOTRANSFORM || is_c2c(b) || b == PFFT_RFFT || is_c2r(b)) return true;
  if (is_r2r(b)) {
    const int32_t k = b - PFFT_R2R;
    if (k == PFFT_R2HC || k == PFFT_HC2R) {
      // r2r.jl:40-46
      if (why) *why = std::string("unsupported r2r transform kind: ") + kR2RNames[k];
      return false;
    }
    return true;
  }
  if (why) *why = "unknown transform kind " + std::to_string(kind);
  return false;
}

std::string kind_name(int32_t kind) {
  const int32_t b = kind_base(kind);
  const char* bang = kind_inplace(kind) ? "!" : "";
  switch (b) {
    case PFFT_NOTRANSFORM: return std::string("NoTransform") + bang;
    case PFFT_FFT: return std::string("FFT") + bang;
    case PFFT_BFFT: return std::string("BFFT") + bang;
    case PFFT_RFFT: return std::string("RFFT") + bang;
    case PFFT_BRFFT_EVEN: return std::string("BRFFT") + bang + "{even}";
    case PFFT_BRFFT_ODD: return std::string("BRFFT") + bang + "{odd}";
    default: break;
  }
  if (is_r2r(b)) return std::string("R2R") + bang + "{" + kR2RNames[b - PFFT_R2R] + "}";
  return "?";
}

std::string kind_type_name(int32_t kind) {
  const int32_t b = kind_base(kind);
  const char* bang = kind_inplace(kind) ? "!" : "";
  std::string base;
  switch (b) {
    case PFFT_NOTRANSFORM: base = "NoTransform"; break;
    case PFFT_FFT: base = "FFT"; break;
    case PFFT_BFFT: base = "BFFT"; break;
    case PFFT_RFFT: base = "RFFT"; break;
    case PFFT_BRFFT_EVEN:
    case PFFT_BRFFT_ODD: base = "BRFFT"; break;
    default:
      if (is_r2r(b))
        return std::string("PencilFFTs.Transforms.R2R") + bang + "{" +
               std::to_string(b - PFFT_R2R) + "}";
      base = "?";
  }
  return "PencilFFTs.Transforms." + base + bang;
}

std::string eltype_name(int32_t e) {
  switch (e) {
    case PFFT_F32: return "Float32";
    case PFFT_F64: return "Float64";
    case PFFT_C32: return "ComplexF32";
    case PFFT_C64: return "ComplexF64";
    default: return "Nothing";
  }
}

size_t eltype_bytes(int32_t e) {
  switch (e) {
    case PFFT_F32: return 4;
    case PFFT_F64: return 8;
    case PFFT_C32: return 8;
    case PFFT_C64: return 16;
    default: return 0;
  }
}

static bool eltype_is_complex(int32_t e) { return e == PFFT_C32 || e == PFFT_C64; }
static int32_t eltype_real(int32_t e) { return (e == PFFT_C32 || e == PFFT_F32) ? PFFT_F32 : PFFT_F64; }
static int32_t eltype_complex(int32_t e) { return eltype_real(e) == PFFT_F32 ? PFFT_C32 : PFFT_C64; }

// c2c.jl:49, r2c.jl:74-75, r2r.jl:82, no_transform.jl:23
int length_output(int32_t kind, int64_t n, int64_t* out) {
  std::string why;
  if (!kind_valid(kind, &why)) return fail(PFFT_ERR_ARGUMENT, why);
  const int32_t b = kind_base(kind);
  if (b == PFFT_RFFT) *out = n / 2 + 1;
  else if (b == PFFT_BRFFT_EVEN) *out = 2 * n - 2;
  else if (b == PFFT_BRFFT_ODD) *out = 2 * n - 1;
  else *out = n;
  return PFFT_OK;
}

// c2c.jl:58-61, r2c.jl:99-102 (BRFFT(d) = BRFFT(iseven(d)), r2c.jl:66),
// r2r.jl:103-131, no_transform.jl:21
int binv(int32_t kind, int64_t d, int32_t* inv) {
  std::string why;
  if (!kind_valid(kind, &why)) return fail(PFFT_ERR_ARGUMENT, why);
  const int32_t ip = kind & PFFT_INPLACE;
  const int32_t b = kind_base(kind);
  int32_t r;
  switch (b) {
    case PFFT_FFT: r = PFFT_BFFT; break;
    case PFFT_BFFT: r = PFFT_FFT; break;
    case PFFT_RFFT: r = (d % 2 == 0) ? PFFT_BRFFT_EVEN : PFFT_BRFFT_ODD; break;
    case PFFT_BRFFT_EVEN:
    case PFFT_BRFFT_ODD: r = PFFT_RFFT; break;
    case PFFT_NOTRANSFORM: r = PFFT_NOTRANSFORM; break;
    default: {
      const int32_t k = b - PFFT_R2R;
      int32_t ki = k;
      if (k == PFFT_REDFT01) ki = PFFT_REDFT10;
      else if (k == PFFT_REDFT10) ki = PFFT_REDFT01;
      else if (k == PFFT_RODFT01) ki = PFFT_RODFT10;
      else if (k == PFFT_RODFT10) ki = PFFT_RODFT01;
      r = PFFT_R2R + ki;
    }
  }
  *inv = r | ip;
  return PFFT_OK;
}

// Scale factor of a transform over one dimension whose input length is n:
// c2c.jl:71, r2c.jl:104-111 (c2r: the real output length), r2r.jl:96-101,
// no_transform.jl:25.
int scale_factor_1d(int32_t kind, int64_t n, int64_t* s) {
  std::string why;
  if (!kind_valid(kind, &why)) return fail(PFFT_ERR_ARGUMENT, why);
  const int32_t b = kind_base(kind);
  if (b == PFFT_NOTRANSFORM) *s = 1;
  else if (is_c2c(b) || b == PFFT_RFFT) *s = n;
  else if (is_c2r(b)) return length_output(kind, n, s);
  else {
    const int32_t k = b - PFFT_R2R;
    if (k == PFFT_REDFT00) *s = 2 * (n - 1);
    else if (k == PFFT_RODFT00) *s = 2 * (n + 1);
    else if (k == PFFT_DHT) *s = n;
    else *s = 2 * n;
  }
  return PFFT_OK;
}

// c2c.jl:51, r2c.jl:80-81, r2r.jl:83, no_transform.jl:24
int eltype_input(int32_t kind, int32_t real_type, int32_t* out) {
  std::string why;
  if (!kind_valid(kind, &why)) return fail(PFFT_ERR_ARGUMENT, why);
  if (real_type != PFFT_F32 && real_type != PFFT_F64)
    return fail(PFFT_ERR_ARGUMENT, "real type must be Float32 or Float64");
  const int32_t b = kind_base(kind);
  if (is_c2c(b) || is_c2r(b)) *out = eltype_complex(real_type);
  else if (b == PFFT_RFFT) *out = real_type;
  else *out = PFFT_NOTHING;
  return PFFT_OK;
}

// c2c.jl:50, r2c.jl:77-78, r2r.jl:84, no_transform.jl:22, error Transforms.jl:263-264
int eltype_output(int32_t kind, int32_t t, int32_t* out) {
  std::string why;
  if (!kind_valid(kind, &why)) return fail(PFFT_ERR_ARGUMENT, why);
  if (t < PFFT_F32 || t > PFFT_C64)
    return fail(PFFT_ERR_ARGUMENT, "invalid element type code " + std::to_string(t));
  const int32_t b = kind_base(kind);
  bool ok = true;
  int32_t r = t;
  if (is_c2c(b)) ok = eltype_is_complex(t);
  else if (b == PFFT_RFFT) { ok = !eltype_is_complex(t); r = eltype_complex(t); }
  else if (is_c2r(b)) { ok = eltype_is_complex(t); r = eltype_real(t); }
  if (!ok)
    return fail(PFFT_ERR_ARGUMENT, "invalid input data type for " + kind_type_name(kind) +
                                       ": " + eltype_name(t));
  *out = r;
  return PFFT_OK;
}

// ---- tables ------------------------------------------------------------------
// exp(-2 pi i num / den) with the argument reduced exactly in integers and the
// trigonometric functions evaluated in long double before rounding.
static void expm(int64_t num, int64_t den, double* re, double* im) {
  int64_t r = num % den;
  if (r < 0) r += den;
  // fold into [0, den/8] with symmetries so the long-double argument is small
  const long double tau = 6.283185307179586476925286766559005768L;
  const long double th = tau * (long double)r / (long double)den;
  *re = (double)cosl(th);
  *im = (double)(-sinl(th));
  // exact values at multiples of quarter turns
  if ((4 * r) % den == 0) {
    const int64_t q = (4 * r) / den;
    const double c[4] = {1, 0, -1, 0}, s[4] = {0, -1, 0, 1};
    *re = c[q & 3];
    *im = s[q & 3];
  } else if ((8 * r) % den == 0) {
    const double h = 0.70710678118654752440084436210485;
    const int64_t q = (8 * r) / den;  // odd
    const double c[8] = {1, h, 0, -h, -1, -h, 0, h};
    const double s[8] = {0, -h, -1, -h, 0, h, 1, h};
    *re = c[q & 7];
    *im = s[q & 7];
  }
}

static int add_table(Plan& p, const std::string& key_unused, std::vector<double>&& vals) {
  (void)key_unused;
  for (size_t i = 0; i < p.tables.size(); ++i)
    if (p.tables[i].values == vals) return (int)i;
  Table t;
  t.real_bytes = p.real_type == PFFT_F32 ? 4 : 8;
  t.values = std::move(vals);
  const size_t n = t.values.size();
  t.host.resize(n * t.real_bytes);
  if (t.real_bytes == 4) {
    float* f = (float*)t.host.data();
    for (size_t i = 0; i < n; ++i) f[i] = (float)t.values[i];
  } else {
    memcpy(t.host.data(), t.values.data(), n * 8);
  }
  p.tables.push_back(std::move(t));
  return (int)p.tables.size() - 1;
}

static int twiddle_table(Plan& p, int64_t M) {
  std::vector<double> v(2 * M);
  for (int64_t k = 0; k < M; ++k) expm(k, M, &v[2 * k], &v[2 * k + 1]);
  return add_table(p, "tw", std::move(v));
}

// Tables of the Bluestein form of an M-point FFT convolved at Mb points:
// the chirp c_j = exp(-i pi j^2 / M), j < M (argument j^2 mod 2M reduced
// exactly), and the Mb-point DFT of b_m = conj(c_|m|) (|m| < M, index m mod
// Mb, zero elsewhere) divided by Mb, from a radix-2 FFT in long double.
// The kernel's forward convolution FFT runs in place, decimation in frequency
// over the radix schedule rad[0..ns): its output position
// p = q0 Mb/r0 + q1 Mb/(r0 r1) + ... holds bin q0 + r0 (q1 + r1 (q2 + ...)).
// The transform of b is stored in that order, so the pointwise product
// needs no reordering and the in-place inverse (decimation in time) returns
// natural order.
static int64_t dif_bin(int64_t pos, int64_t Mb, const int32_t* rad, int ns) {
  int64_t k = 0, w = 1, B = Mb;
  for (int s = 0; s < ns; ++s) {
    B /= rad[s];
    const int64_t q = (pos / B) % rad[s];
    k += q * w;
    w *= rad[s];
  }
  return k;
}

static void bluestein_tables(Plan& p, int64_t M, int64_t Mb, const int32_t* rad, int ns,
                             int32_t* chirp_id, int32_t* bhat_id) {
  std::vector<double> ch(2 * M);
  for (int64_t j = 0; j < M; ++j) expm((j * j) % (2 * M), 2 * M, &ch[2 * j], &ch[2 * j + 1]);
  std::vector<long double> re(Mb, 0.0L), im(Mb, 0.0L);
  for (int64_t m = 0; m < M; ++m) {
    re[m] = ch[2 * m];
    im[m] = -(long double)ch[2 * m + 1];
    if (m) { re[Mb - m] = re[m]; im[Mb - m] = im[m]; }
  }
  // in-place decimation-in-time FFT, bit-reversed input order
  for (int64_t i = 1, j = 0; i < Mb; ++i) {
    int64_t bit = Mb >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) { std::swap(re[i], re[j]); std::swap(im[i], im[j]); }
  }
  const long double tau = 6.283185307179586476925286766559005768L;
  for (int64_t len = 2; len <= Mb; len <<= 1) {
    const int64_t h = len / 2;
    for (int64_t k = 0; k < h; ++k) {
      const long double wr = cosl(tau * (long double)k / (long double)len);
      const long double wi = -sinl(tau * (long double)k / (long double)len);
      for (int64_t s = 0; s < Mb; s += len) {
        const int64_t u = s + k, v = s + k + h;
        const long double tr = re[v] * wr - im[v] * wi, ti = re[v] * wi + im[v] * wr;
        re[v] = re[u] - tr; im[v] = im[u] - ti;
        re[u] += tr; im[u] += ti;
      }
    }
  }
  std::vector<double> bh(2 * Mb);
  for (int64_t pos = 0; pos < Mb; ++pos) {
    const int64_t k = dif_bin(pos, Mb, rad, ns);
    bh[2 * pos] = (double)(re[k] / (long double)Mb);
    bh[2 * pos + 1] = (double)(im[k] / (long double)Mb);
  }
  *chirp_id = add_table(p, "bs_chirp", std::move(ch));
  *bhat_id = add_table(p, "bs_bhat", std::move(bh));
}

// LDS pitch (complex slots) of a generic-pass work line of Mw points: one
// pad slot after every 8 points, odd so that neighbouring lines start on
// other banks (the kernels' wp())
static int64_t line_pitch_slots(int64_t Mw) {
  if (Mw <= 0) return 0;
  const int64_t span = Mw + (Mw - 1) / 8;
  return span % 2 == 0 ? span + 1 : span;
}

static int64_t largest_prime_factor(int64_t m) {
  int64_t best = 1;
  for (int64_t q = 2; q * q <= m; ++q)
    while (m % q == 0) { best = q; m /= q; }
  return m > 1 ? std::max(best, m) : best;
}

// ---- radix schedule ----------------------------------------------------------
static int radix_schedule(int64_t M, int32_t* rad) {
  int ns = 0;
  int64_t m = M;
  for (int r : {8, 4, 2}) {
    while (m > 1 && m % r == 0) { rad[ns++] = r; m /= r; }
  }
  for (int r : {3, 5}) {
    while (m % r == 0) { rad[ns++] = r; m /= r; }
  }
  for (int64_t q = 7; m > 1; q += 2) {
    while (m % q == 0) {
      if (ns >= kMaxStages) return -1;
      rad[ns++] = (int32_t)q;
      m /= q;
    }
    if (q * q > m && m > 1) {
      if (ns >= kMaxStages) return -1;
      rad[ns++] = (int32_t)m;
      m = 1;
    }
  }
  return ns;
}

// ---- layouts -----------------------------------------------------------------
// Strides (in elements) of each logical dim and of the extra dim for a memory
// permutation `perm` of the N logical dims with sizes `dims`.
// Strides of the layout `perm` over `dims`; a nonzero pad0 is the padded
// length of memory axis 0 (the real view of an in-place r2c buffer holds
// 2 (n/2 + 1) reals per line, multiarrays_r2c.jl:75-86).
static void layout_strides(int N, const int32_t* perm, const int64_t* dims, int64_t* stride,
                           int64_t* stride_extra, int64_t pad0 = 0) {
  int64_t s = 1;
  for (int i = 0; i < N; ++i) {
    stride[perm[i]] = s;
    s *= (i == 0 && pad0 > dims[perm[i]]) ? pad0 : dims[perm[i]];
  }
  *stride_extra = s;
}

static void chain_perm(int N, int n, bool permute, int32_t* perm) {
  // plans.jl:503-521: stage n (0-based) reads data permuted as
  // (n, 0..n-1, n+1..N-1), the last stage fully reversed; stage 0 unpermuted.
  for (int i = 0; i < N; ++i) perm[i] = i;
  if (!permute || n == 0) return;
  if (n == N - 1) {
    for (int i = 0; i < N; ++i) perm[i] = N - 1 - i;
    return;
  }
  perm[0] = n;
  for (int i = 1; i <= n; ++i) perm[i] = i - 1;
  for (int i = n + 1; i < N; ++i) perm[i] = i;
}

// Build the pass for transforming logical dim `dim` from (perm_in, dims_in)
// into (perm_out, dims_out).
// Radix schedule, twiddles and per-kind tables of a pass's M-point line
// FFT (a.kind, a.n, a.M, a.raw_pitch set).
static int line_fft(Plan& p, PassPlan& pp) {
  PassArgs& a = pp.args;
  const int lk = a.kind;
  const int64_t M = a.M, n = a.n;
  const int64_t rs = p.real_type == PFFT_F32 ? 4 : 8;
  const int64_t cs = 2 * rs;
  // A large prime factor makes the direct stages cost M p per line; the
  // Bluestein form does the line in two FFTs of the next power of two at
  // or above 2M - 1 instead, when one line of it fits the LDS.
  int64_t Mb = 0;
  if (largest_prime_factor(M) > kBluesteinPrime) {
    Mb = 1;
    while (Mb < 2 * M - 1) Mb *= 2;
    // one work buffer (the convolution runs in place) beside the raw line
    const int64_t Tw1 = a.r2r_cplx ? 2 : 1;
    const int64_t w0 = (Tw1 * line_pitch_slots(Mb) * cs + 255) / 256 * 256;
    if (w0 + a.raw_pitch * rs > kMaxLds) Mb = 0;
  }
  const int ns = radix_schedule(Mb ? Mb : M, a.radix);
  if (ns < 0) return fail(PFFT_ERR_UNSUPPORTED, "too many radix stages");
  a.nstages = ns;
  pp.tw_id = twiddle_table(p, Mb ? Mb : M);
  if (Mb) {
    a.bs_M = Mb;
    bluestein_tables(p, M, Mb, a.radix, ns, &pp.bs_chirp_id, &pp.bs_bhat_id);
  }
  // aux tables
  std::vector<double> aux;
  if (lk == L_R2C_HALF || lk == L_C2R_HALF) {
    aux.resize(2 * (M + 1));
    for (int64_t k = 0; k <= M; ++k) expm(k, n, &aux[2 * k], &aux[2 * k + 1]);
  } else if (lk == L_DCT2 || lk == L_DST2 || lk == L_DCT3 || lk == L_DST3) {
    aux.resize(2 * n);
    for (int64_t k = 0; k < n; ++k) expm(k, 4 * n, &aux[2 * k], &aux[2 * k + 1]);
  } else if (lk == L_DCT4 || lk == L_DST4) {
    aux.resize(4 * n);
    for (int64_t k = 0; k < n; ++k) expm(k, 4 * n, &aux[2 * k], &aux[2 * k + 1]);
    for (int64_t k = 0; k < n; ++k)
      expm(2 * k + 1, 8 * n, &aux[2 * (n + k)], &aux[2 * (n + k) + 1]);
  } else if (lk == L_DCT4H || lk == L_DST4H) {
    // entries m < n/2: the pre-twiddle c_m = exp(-i pi (4m + 1) / 4n) of the
    // half z_m; entries n/2 + m: e_m = exp(i pi m / n), the post-twiddle
    // conj(e_m) of Z_m and the split factor w_n^-m folded into it
    aux.resize(2 * n);
    const int64_t h = n / 2;
    for (int64_t m = 0; m < h; ++m) {
      expm(4 * m + 1, 8 * n, &aux[2 * m], &aux[2 * m + 1]);
      expm(-m, 2 * n, &aux[2 * (h + m)], &aux[2 * (h + m) + 1]);
    }
  } else if (lk == L_C2C && M == 1024 && p.real_type == PFFT_F32) {
    // middle twiddles of the 32 x 32 MFMA form of 1024-point lines, in the
    // order its lanes read them: entry 64 i + l is w1024^(k2 n1) with
    // k2 = l mod 32 and n1 = (i mod 4) + 8 (i div 4) + 4 (l div 32), the
    // row of accumulator register i of lane l
    aux.resize(2 * 1024);
    for (int i = 0; i < 16; ++i)
      for (int l = 0; l < 64; ++l) {
        const int64_t k2 = l & 31, n1 = (i & 3) + 8 * (i >> 2) + 4 * (l >> 5);
        expm(k2 * n1, 1024, &aux[2 * (64 * i + l)], &aux[2 * (64 * i + l) + 1]);
      }
  }
  if (!aux.empty()) pp.aux_id = add_table(p, "aux", std::move(aux));
  return PFFT_OK;
}

// Line dims of a pass: length and input / output strides (elements), and the
// logical dim (N: the flattened extra dims, -1: none).
struct LineDim {
  int64_t len, si, so;
  int logical;
};

// LDS budget of one generic-pass tile (the tile size T halves until it fits)
#ifndef PFFT_GENERIC_TILE_LDS
#define PFFT_GENERIC_TILE_LDS (64 * 1024)
#endif
constexpr int64_t kGenericTileLds = PFFT_GENERIC_TILE_LDS;

// Line-dim order, tile size and LDS footprint of the generic pass (`dim`:
// the logical dim transformed).  *fits = false when one line does not fit
// the LDS.
static int line_layout(Plan& p, PassPlan& pp, std::vector<LineDim> lds, int dim, bool* fits) {
  PassArgs& a = pp.args;
  const int N = p.N;
  const int lk = a.kind;
  const int64_t M = a.M;
  const int64_t rs = p.real_type == PFFT_F32 ? 4 : 8;
  const int64_t cs = 2 * rs;
  using LD = LineDim;
  *fits = true;
  // Line dims in input-memory order; then the tile axis goes first: the
  // output's contiguous axis when the store transposes, else the input's
  // contiguous axis when the load is strided, else the line dim nearest in
  // input memory (consecutive lines are then adjacent in memory).
  std::stable_sort(lds.begin(), lds.end(), [](const LD& x, const LD& y) { return x.si < y.si; });
  int tile = -1;
  for (size_t i = 0; i < lds.size() && tile < 0; ++i)
    if (a.sk_out != 1 && lds[i].so == 1 && lds[i].len > 1) tile = (int)i;
  for (size_t i = 0; i < lds.size() && tile < 0; ++i)
    if (a.sk_in != 1 && lds[i].si == 1 && lds[i].len > 1) tile = (int)i;
  if (tile > 0) {
    const LD t = lds[tile];
    lds.erase(lds.begin() + tile);
    lds.insert(lds.begin(), t);
  }
  // drop unit dims except the first
  std::vector<LD> keep;
  for (size_t i = 0; i < lds.size(); ++i)
    if (i == 0 || lds[i].len > 1) keep.push_back(lds[i]);
  if ((int)keep.size() > kMaxDims) return fail(PFFT_ERR_UNSUPPORTED, "too many dimensions");
  a.D = (int32_t)keep.size();
  for (int i = 0; i < a.D; ++i) {
    a.ldim[i] = keep[i].len;
    a.lsin[i] = keep[i].si;
    a.lsout[i] = keep[i].so;
    a.llog[i] = (keep[i].logical >= 0 && keep[i].logical < N) ? keep[i].logical : -1;
  }
  a.pdim = dim;

  // tile size and LDS footprint of the generic pass
  const int64_t Mw = a.bs_M ? a.bs_M : M;  // points per work line
  a.Mp = line_pitch_slots(Mw);
  // c2c lines (and the real pairs of an even r2c, the Hermitian merge of an
  // even c2r) load straight into the work lines: no raw tile, no pre phase
  a.load_w0 = (lk == L_C2C || lk == L_R2C_HALF || lk == L_C2R_HALF) ? 1 : 0;
  // Two real R2R lines share one complex FFT as its real and imaginary
  // parts: DCT-I / DST-I (real transforms), DCT-II / DST-II and DHT (split
  // by Hermitian symmetry afterwards) and DCT-III / DST-III (Hermitian
  // input, real transforms)
  a.pair_real = (lk == L_DCT1 || lk == L_DST1 || lk == L_DCT2 || lk == L_DST2 ||
                 lk == L_DCT3 || lk == L_DST3 || lk == L_DHT) ? 1 : 0;
  auto footprint = [&](int64_t T, int64_t* w1off) {
    const int64_t nreal = a.r2r_cplx ? 2 * T : T;
    const int64_t Tw = a.pair_real ? (nreal + 1) / 2 : nreal;
    const int64_t w0 = lk == L_COPY ? 0 : Tw * a.Mp * cs;
    // (the Bluestein form's convolution runs in place in work buffer 0)
    const int64_t w1 = (lk == L_COPY || a.nstages == 0 || a.bs_M) ? 0 : Tw * a.Mp * cs;
    const int64_t raw = a.load_w0 ? 0 : T * a.raw_pitch * rs;
    *w1off = (w0 + 255) / 256 * 256;
    return *w1off + std::max(w1, raw);
  };
  int64_t T = 1;
  while (T < 64 && T < a.ldim[0]) T *= 2;
  int64_t off = 0;
  int64_t bytes = footprint(T, &off);
  while (bytes > kGenericTileLds && T > 1) {
    T /= 2;
    bytes = footprint(T, &off);
  }
  if (bytes > kMaxLds) {
    *fits = false;
    return PFFT_OK;
  }
  a.T = (int32_t)T;
  a.lds_w1_off = off;
  pp.lds_bytes = bytes;
  // Lines whose FFT has a direct prime stage (radix 7 and up) or runs in the
  // Bluestein form are arithmetic-heavy: twice the threads per tile hide
  // more of it (measured 13-30 % faster on such passes; the power-of-two
  // and 3/5-smooth passes measured 3-15 % slower that way)
  bool direct_prime = false;
  for (int i = 0; i < a.nstages; ++i) direct_prime = direct_prime || (a.radix[i] > 5 && a.radix[i] != 8);
  pp.threads = (a.bs_M || direct_prime) ? kBlockThreadsHeavy : kBlockThreads;
  int64_t nt = (a.ldim[0] + T - 1) / T;
  for (int i = 1; i < a.D; ++i) nt *= a.ldim[i];
  a.ntiles = nt;
  pp.grid = std::min<int64_t>(nt, 1 << 20);
  a.scale = 1.0;
  a.do_scale = 0;
  pp.gen_args = a;
  pp.gen_threads = pp.threads;
  pp.gen_lds = pp.lds_bytes;
  pp.gen_grid = pp.grid;
  return PFFT_OK;
}

// Four-step form of a c2c line too long for one LDS-resident pass: M = M1 M2
// with j = M2 j1 + j2 and k = k1 + M1 k2,
//   pass A: for each j2, the M1-point FFT over j1, times w_M^(j2 k1), written
//           to an intermediate Y[j2 + M2 (k1 + M1 r)] (r: the other line dims,
//           dense);
//   pass B: for each k1, the M2-point FFT over j2 of Y, written at k.
// Pass A's line dim 0 is j2 (its store's contiguous axis), which the kernel
// reads back as the tile's first line index for the twiddle.
static int split_pass(Plan& p, const PassPlan& whole, const std::vector<LineDim>& lds, int dim,
                      std::vector<PassPlan>& out) {
  const PassArgs& w = whole.args;
  const int64_t M = w.M;
  int64_t M2 = 0;
  for (int64_t f = 2; f <= M / 2; ++f)
    if (M % f == 0 && f * f >= M) { M2 = f; break; }
  if (!M2 || w.op || w.r2r_cplx)
    return fail(PFFT_ERR_UNSUPPORTED,
                "line of length " + std::to_string(M) + " does not fit the LDS-resident pass");
  const int64_t M1 = M / M2;
  int64_t rest = 1;
  std::vector<int64_t> R;
  for (const LineDim& d : lds) {
    R.push_back(rest);
    rest *= d.len;
  }
  PassPlan A = whole, B = whole;
  for (PassPlan* q : {&A, &B}) {
    q->tw_id = q->aux_id = q->bs_chirp_id = q->bs_bhat_id = q->fs_tw_id = -1;
    PassArgs& a = q->args;
    a.bs_M = 0;
    a.fs_M = 0;
  }
  const int64_t mid = M * rest;   // intermediate elements
  // pass A
  {
    PassArgs& a = A.args;
    a.n = a.M = a.L_in = a.L_out = M1;
    a.sk_in = M2 * w.sk_in;
    a.sk_out = M2;
    a.raw_pitch = 2 * M1 + 2;
    a.fs_M = M;
    A.fs_tw_id = twiddle_table(p, M);
    A.out_elems = mid;
    A.elt_out = whole.elt_in;
    std::vector<LineDim> la = {{M2, w.sk_in, 1, -1}};
    for (size_t i = 0; i < lds.size(); ++i) la.push_back({lds[i].len, lds[i].si, M * R[i], lds[i].logical});
    int r = line_fft(p, A);
    bool fits = true;
    if (!r) r = line_layout(p, A, la, dim, &fits);
    if (r) return r;
    if (!fits || A.args.ldim[0] != M2)
      return fail(PFFT_ERR_UNSUPPORTED,
                  "line of length " + std::to_string(M) + " does not fit the LDS-resident pass");
  }
  // pass B
  {
    PassArgs& a = B.args;
    a.n = a.M = a.L_in = a.L_out = M2;
    a.sk_in = 1;
    a.sk_out = M1 * w.sk_out;
    a.raw_pitch = 2 * M2 + 2;
    B.in_elems = mid;
    B.elt_in = whole.elt_out;
    std::vector<LineDim> lb = {{M1, M2, w.sk_out, -1}};
    for (size_t i = 0; i < lds.size(); ++i) lb.push_back({lds[i].len, M * R[i], lds[i].so, lds[i].logical});
    int r = line_fft(p, B);
    bool fits = true;
    if (!r) r = line_layout(p, B, lb, dim, &fits);
    if (r) return r;
    if (!fits)
      return fail(PFFT_ERR_UNSUPPORTED,
                  "line of length " + std::to_string(M) + " does not fit the LDS-resident pass");
  }
  A.split_part = 1;
  B.split_part = 2;
  out.push_back(A);
  out.push_back(B);
  return PFFT_OK;
}

static int build_pass(Plan& p, int dim, int32_t kind, int32_t elt_in, int32_t elt_out,
                      const int32_t* perm_in, const int64_t* dims_in, const int32_t* perm_out,
                      const int64_t* dims_out, std::vector<PassPlan>& out,
                      int64_t pad_in = 0, int64_t pad_out = 0) {
  const int N = p.N;
  PassPlan pp;
  PassArgs& a = pp.args;
  memset(&a, 0, sizeof(a));
  pp.elt_in = elt_in;
  pp.elt_out = elt_out;
  pp.tw_id = -1;
  pp.aux_id = -1;
  pp.variant = 0;

  int64_t sin_[kMaxDims], sout[kMaxDims], sin_e, sout_e;
  layout_strides(N, perm_in, dims_in, sin_, &sin_e, pad_in);
  layout_strides(N, perm_out, dims_out, sout, &sout_e, pad_out);
  // element counts of the arrays the pass touches (padding included)
  pp.in_elems = p.E * sin_e;
  pp.out_elems = p.E * sout_e;

  a.L_in = dims_in[dim];
  a.L_out = dims_out[dim];
  a.sk_in = sin_[dim];
  a.sk_out = sout[dim];
  a.in_cplx = eltype_is_complex(elt_in);
  a.out_cplx = eltype_is_complex(elt_out);

  // line kind, logical length, FFT length
  const int32_t b = kind_base(kind);
  int64_t n = a.L_in, M = 0;
  int lk = L_COPY;
  int fft_bwd = 0;
  if (b == PFFT_NOTRANSFORM) {
    lk = L_COPY;
  } else if (is_c2c(b)) {
    lk = L_C2C;
    M = n;
    fft_bwd = b == PFFT_BFFT;
  } else if (b == PFFT_RFFT) {
    if (n % 2 == 0) { lk = L_R2C_HALF; M = n / 2; }
    else { lk = L_R2C_FULL; M = n; }
  } else if (is_c2r(b)) {
    n = a.L_out;
    if (n % 2 == 0) { lk = L_C2R_HALF; M = n / 2; }
    else { lk = L_C2R_FULL; M = n; }
    fft_bwd = 1;
  } else {
    const int32_t k = b - PFFT_R2R;
    a.r2r_cplx = a.in_cplx;
    switch (k) {
      case PFFT_REDFT00:
        if (n < 2) return fail(PFFT_ERR_ARGUMENT, "REDFT00 requires a length of at least 2");
        lk = L_DCT1; M = 2 * (n - 1); break;
      case PFFT_REDFT10: lk = L_DCT2; M = n; break;
      case PFFT_REDFT01: lk = L_DCT3; M = n; fft_bwd = 1; break;
      // even lengths: one n-point FFT holds two real lines' n/2-point FFTs
      case PFFT_REDFT11:
        if (n % 2 == 0) { lk = L_DCT4H; M = n; }
        else { lk = L_DCT4; M = 2 * n; }
        break;
      case PFFT_RODFT00: lk = L_DST1; M = 2 * (n + 1); break;
      case PFFT_RODFT10: lk = L_DST2; M = n; break;
      case PFFT_RODFT01: lk = L_DST3; M = n; fft_bwd = 1; break;
      case PFFT_RODFT11:
        if (n % 2 == 0) { lk = L_DST4H; M = n; }
        else { lk = L_DST4; M = 2 * n; }
        break;
      case PFFT_DHT: lk = L_DHT; M = n; break;
      default: return fail(PFFT_ERR_ARGUMENT, "unsupported r2r kind");
    }
  }
  if (lk == L_COPY) a.r2r_cplx = 0;
  a.kind = lk;
  a.bwd = fft_bwd;
  a.n = n;
  a.M = M;

  const int comp_in = a.in_cplx ? 2 : 1;
  a.raw_pitch = comp_in * a.L_in + comp_in;
  // line dims: every logical dim but `dim`, plus the flattened extra dim
  using LD = LineDim;
  std::vector<LD> lds;
  for (int d = 0; d < N; ++d)
    if (d != dim) lds.push_back({dims_in[d], sin_[d], sout[d], d});
  if (p.E > 1) lds.push_back({p.E, sin_e, sout_e, N});
  if (lds.empty()) lds.push_back({1, 0, 0, -1});
  if (lk == L_C2C) {
    // one line of the direct form: two work buffers of M + 1 points
    const int64_t cs = p.real_type == PFFT_F32 ? 8 : 16;
    const int64_t w0 = (line_pitch_slots(M) * cs + 255) / 256 * 256;
    if (w0 + line_pitch_slots(M) * cs > kMaxLds) return split_pass(p, pp, lds, dim, out);
  }
  if (lk != L_COPY) {
    const int r = line_fft(p, pp);
    if (r) return r;
  }

  bool fits = true;
  int r = line_layout(p, pp, lds, dim, &fits);
  if (r) return r;
  if (!fits)
    return fail(PFFT_ERR_UNSUPPORTED,
                "line of length " + std::to_string(n) + " does not fit the LDS-resident pass");
  out.push_back(pp);
  return PFFT_OK;
}

// ---- spectral operators (pfft_execute_op) ------------------------------------
int op_setup(const Plan& p, int dir, const pfft_spectral_op& op, int* pass, int* side) {
  if (op.op < PFFT_OP_DERIVATIVE || op.op > PFFT_OP_DEALIAS)
    return fail(PFFT_ERR_ARGUMENT, "unknown spectral operator " + std::to_string(op.op));
  if (op.op == PFFT_OP_DERIVATIVE && (op.dim < 0 || op.dim >= p.N))
    return fail(PFFT_ERR_ARGUMENT, "derivative dimension out of range");
  for (int d = 0; d < p.N; ++d) {
    const int32_t b = kind_base(p.kinds[d]);
    if (b != PFFT_FFT && b != PFFT_BFFT && b != PFFT_RFFT)
      return fail(PFFT_ERR_UNSUPPORTED,
                  "spectral operators need wavenumbers: FFT/BFFT/RFFT dimensions only");
  }
  // mul! and the transpose of bmul! write the spectrum in the last forward
  // pass; bmul!/ldiv! and the transpose of mul! read it in the first
  // backward pass
  const bool writes = dir == PFFT_FORWARD || dir == PFFT_ADJOINT_BACKWARD;
  const std::vector<PassPlan>& list = writes ? p.fwd : p.bwd;
  *pass = writes ? (int)list.size() - 1 : 0;
  *side = writes ? 2 : 1;
  if (op.accumulate && !writes)
    return fail(PFFT_ERR_ARGUMENT,
                "accumulate applies where the spectrum is written (forward, adjoint of bmul!)");
  if (list[*pass].args.kind != L_C2C)
    return fail(PFFT_ERR_UNSUPPORTED,
                "the pass carrying a spectral operator must be a c2c one (two or more "
                "transformed dimensions)");
  if (list[*pass].split_part)
    return fail(PFFT_ERR_UNSUPPORTED,
                "a spectral operator cannot ride on a line split into two passes");
  return PFFT_OK;
}

void op_fill(const Plan& p, int dir, const pfft_spectral_op& op, int side, PassArgs& a) {
  a.op = op.op;
  a.op_side = side;
  a.op_dim = op.dim;
  a.op_conj = (dir == PFFT_ADJOINT_FORWARD || dir == PFFT_ADJOINT_BACKWARD) ? 1 : 0;
  a.op_acc = op.accumulate ? 1 : 0;
  a.op_alpha = op.alpha;
  a.kN = p.N;
  for (int d = 0; d < p.N; ++d) {
    const bool r2c = kind_base(p.kinds[d]) == PFFT_RFFT;
    a.kmode[d] = r2c ? 1 : 0;
    a.kn[d] = p.size_out[d];
    a.kmax[d] = r2c ? (double)(p.size_out[d] - 1) : (double)(p.size_out[d] / 2);
  }
}

// ---- plan construction -----------------------------------------------------
int plan_build(Plan& p, int32_t N, const int64_t* dims, const int32_t* kinds,
               int32_t real_type, int32_t permute, int32_t n_extra, const int64_t* extra,
               uint32_t flags) {
  if (N < 1 || N > kMaxDims - 1)
    return fail(PFFT_ERR_ARGUMENT, "number of transformed dimensions must be in 1.." +
                                       std::to_string(kMaxDims - 1));
  if (n_extra < 0 || n_extra > kMaxDims) return fail(PFFT_ERR_ARGUMENT, "too many extra dims");
  if (real_type != PFFT_F32 && real_type != PFFT_F64)
    return fail(PFFT_ERR_ARGUMENT, "real type must be Float32 or Float64");
  p.N = N;
  p.Ne = n_extra;
  p.real_type = real_type;
  p.permute = permute != 0;
  p.flags = flags;
  p.E = 1;
  for (int i = 0; i < n_extra; ++i) {
    if (extra[i] < 1) return fail(PFFT_ERR_ARGUMENT, "extra dims must be positive");
    p.extra[i] = extra[i];
    p.E *= extra[i];
  }
  for (int d = 0; d < N; ++d) {
    std::string why;
    if (!kind_valid(kinds[d], &why)) return fail(PFFT_ERR_ARGUMENT, why);
    if (dims[d] < 1) return fail(PFFT_ERR_ARGUMENT, "global dimensions must be positive");
    p.kinds[d] = kinds[d];
    p.size_in[d] = dims[d];
  }
  // GlobalFFTParams: in-place consistency (global_params.jl:56-60)
  bool ip0 = kind_inplace(kinds[0]);
  for (int d = 1; d < N; ++d) {
    if (kind_inplace(kinds[d]) != ip0) {
      std::string s = "cannot combine in-place and out-of-place transforms: (";
      for (int e = 0; e < N; ++e) s += (e ? ", " : "") + kind_name(kinds[e]);
      s += ")";
      return fail(PFFT_ERR_ARGUMENT, s);
    }
  }
  p.inplace = ip0;
  p.adjoint_ok = true;
  for (int d = 0; d < N; ++d)
    if (is_r2r(kind_base(kinds[d]))) p.adjoint_ok = false;
  for (int d = 0; d < N; ++d) {
    int r = length_output(kinds[d], dims[d], &p.size_out[d]);
    if (r) return r;
    if (p.size_out[d] < 1) return fail(PFFT_ERR_ARGUMENT, "transform output length is zero");
  }
  // input element type (global_params.jl:81-95)
  int32_t T = real_type;
  for (int d = 0; d < N; ++d) {
    int32_t t;
    int r = eltype_input(kinds[d], real_type, &t);
    if (r) return r;
    if (t != PFFT_NOTHING) { T = t; break; }
  }
  p.elt_in = T;

  // stages (plans.jl:406-446, 523-553)
  p.stages.clear();
  p.scale = 1.0;
  int32_t Ti = T;
  for (int n = 0; n < N; ++n) {
    Stage s;
    memset(&s, 0, sizeof(s));
    s.dim = n;
    s.kind_fwd = kinds[n];
    int r = eltype_output(kinds[n], Ti, &s.elt_out);
    if (r) return r;
    s.elt_in = Ti;
    chain_perm(N, n, p.permute, s.perm);
    for (int j = 0; j < N; ++j) {
      s.dims_before[j] = j < n ? p.size_out[j] : p.size_in[j];
      s.dims_after[j] = j <= n ? p.size_out[j] : p.size_in[j];
    }
    for (int i = 0; i < N; ++i) if (s.perm[i] == n) s.mem_axis = i;
    s.n_in = p.size_in[n];
    s.n_out = p.size_out[n];
    // binv with the input length along the transformed dimension
    r = binv(kinds[n], s.n_in, &s.kind_bwd);
    if (r) return r;
    int64_t sc;
    r = scale_factor_1d(s.kind_bwd, s.n_out, &sc);
    if (r) return r;
    p.scale *= (double)sc;
    p.stages.push_back(s);
    Ti = s.elt_out;
  }
  p.elt_out = Ti;

  // In-place plans (allocate.jl:57-83): chains that keep the element type
  // (FFT!, BFFT!, R2R!, NoTransform!) over one buffer of the largest pencil,
  // and RFFT! followed by FFT!s over the padded real layout of
  // multiarrays_r2c.jl (2 (n/2 + 1) reals per first-axis line).
  int64_t pad_real = 0;
  if (p.inplace) {
    const int32_t b0 = kind_base(kinds[0]);
    bool ok = true;
    if (b0 == PFFT_RFFT) {
      for (int d = 1; d < N; ++d) ok = ok && kind_base(kinds[d]) == PFFT_FFT;
      pad_real = 2 * (p.size_in[0] / 2 + 1);
    } else {
      for (int d = 0; d < N; ++d)
        ok = ok && kind_base(kinds[d]) != PFFT_RFFT && !is_c2r(kind_base(kinds[d]));
    }
    if (!ok)
      return fail(PFFT_ERR_UNSUPPORTED,
                  "in-place plans take FFT!/BFFT!/R2R!/NoTransform! chains, or RFFT! "
                  "followed by FFT!s (allocate.jl:57-83)");
  }
  p.pad_in0 = pad_real;

  // passes
  p.tables.clear();
  p.fwd.clear();
  p.bwd.clear();
  // Intermediate buffers are internal, so their memory order is free.  The
  // default is the reference chain (pass n writes the pencil of stage n+1);
  // when that layout would make an odd-length axis the fastest one (e.g. the
  // n/2+1 axis after an r2c), the store's runs would be misaligned, and the
  // pass keeps its input's order instead (the next pass then loads strided).
  auto choose = [&](const int32_t* perm_in, const int32_t* perm_def, const int64_t* dims,
                    int32_t* perm_out) {
    memcpy(perm_out, perm_def, sizeof(int32_t) * N);
    const int64_t fast_len = dims[perm_def[0]];
    if (fast_len > 1 && (fast_len & 1) &&
        memcmp(perm_in, perm_def, sizeof(int32_t) * N) != 0)
      memcpy(perm_out, perm_in, sizeof(int32_t) * N);
  };
  // Padded rows of intermediate buffers (the fastest dim grows by 32
  // elements).  Measured on MI355X: rows of 8 KiB and more (1024-point
  // complex lines) run 4 % faster per 1024^3 pair padded, whatever the
  // access; a buffer the next pass loads as tiles loads 6 % faster padded
  // (the 512^3 c2r pass, 4 KiB rows); a buffer of shorter rows written by
  // tile stores stores 6 % slower padded, so those keep dense rows, as do
  // rows under 2 KiB (small plans, where 32 elements would be a large
  // share).  In-place plans share one buffer between pencils and are never
  // padded.
  auto work_pad = [&](const int32_t* perm, const int64_t* dims, int32_t elt,
                      int consumer_dim) -> int64_t {
    if (p.inplace || (p.flags & PFFT_FLAG_DENSE_WORK)) return 0;
    const int64_t len = dims[perm[0]];
    const int64_t row = len * (int64_t)eltype_bytes(elt);
    const bool tile_load = perm[0] != consumer_dim;
    if (row < (tile_load ? 2048 : 8192)) return 0;
    return len + 32;
  };
  {
    int32_t cur[kMaxDims];
    int64_t cur_pad = pad_real;
    memcpy(cur, p.stages[0].perm, sizeof(cur));
    for (int n = 0; n < N; ++n) {
      const Stage& s = p.stages[n];
      int32_t outp[kMaxDims];
      if (n == N - 1) memcpy(outp, s.perm, sizeof(outp));
      else choose(cur, p.stages[n + 1].perm, s.dims_after, outp);
      const int64_t pad_out = n == N - 1 ? 0 : work_pad(outp, s.dims_after, s.elt_out, n + 1);
      int r = build_pass(p, n, s.kind_fwd, s.elt_in, s.elt_out, cur, s.dims_before, outp,
                         s.dims_after, p.fwd, cur_pad, pad_out);
      if (r) return r;
      memcpy(cur, outp, sizeof(cur));
      cur_pad = pad_out;
    }
  }
  {
    int32_t cur[kMaxDims];
    int64_t cur_pad = 0;
    memcpy(cur, p.stages[N - 1].perm, sizeof(cur));
    for (int n = N - 1; n >= 0; --n) {
      const Stage& s = p.stages[n];
      int32_t outp[kMaxDims];
      if (n == 0) memcpy(outp, s.perm, sizeof(outp));
      else choose(cur, p.stages[n - 1].perm, s.dims_before, outp);
      const int64_t pad_out = n == 0 ? pad_real : work_pad(outp, s.dims_before, s.elt_in, n - 1);
      int r = build_pass(p, n, s.kind_bwd, s.elt_out, s.elt_in, cur, s.dims_after, outp,
                         s.dims_before, p.bwd, cur_pad, pad_out);
      if (r) return r;
      memcpy(cur, outp, sizeof(cur));
      cur_pad = pad_out;
    }
  }
  size_t w[2] = {0, 0};
  auto use = [&](PassPlan& pp) {
    if (pp.dst) {
      const size_t bytes = (size_t)pp.out_elems * eltype_bytes(pp.elt_out);
      w[pp.dst - 1] = std::max(w[pp.dst - 1], bytes);
    }
  };
  if (!p.inplace) {
    // buffers: user in -> work0 -> work1 -> work0 ... -> user out
    for (auto* list : {&p.fwd, &p.bwd}) {
      const int np = (int)list->size();
      for (int i = 0; i < np; ++i) {
        PassPlan& pp = (*list)[i];
        pp.src = i == 0 ? 0 : 1 + ((i - 1) & 1);
        pp.dst = i == np - 1 ? 0 : 1 + (i & 1);
        use(pp);
      }
    }
    p.in_bytes = (size_t)p.fwd.front().in_elems * eltype_bytes(p.elt_in);
    p.out_bytes = (size_t)p.fwd.back().out_elems * eltype_bytes(p.elt_out);
  } else {
    // One user buffer U holds every pencil (ManyPencilArray, the largest
    // pencil's size).  A pass whose input and output lines occupy the same
    // bytes (contiguous on both sides, same line strides in bytes) may run in
    // place, each line read whole before it is written; any other pass moves
    // data between U and the scratch W.  The chain starts and ends in U, so
    // the moves must be even: an odd count turns one in-place-capable pass
    // into a move, and failing that the chain uses two scratch buffers.
    size_t ubytes = 0;
    for (auto* list : {&p.fwd, &p.bwd})
      for (const PassPlan& pp : *list) {
        ubytes = std::max(ubytes, (size_t)pp.in_elems * eltype_bytes(pp.elt_in));
        ubytes = std::max(ubytes, (size_t)pp.out_elems * eltype_bytes(pp.elt_out));
      }
    for (auto* list : {&p.fwd, &p.bwd}) {
      const int np = (int)list->size();
      std::vector<bool> same(np);
      int moves = 0, first_same = -1;
      for (int i = 0; i < np; ++i) {
        const PassArgs& a = (*list)[i].args;
        const int64_t bi = eltype_bytes((*list)[i].elt_in), bo = eltype_bytes((*list)[i].elt_out);
        bool ok = a.sk_in == 1 && a.sk_out == 1;
        for (int d = 0; d < a.D && ok; ++d) ok = a.lsin[d] * bi == a.lsout[d] * bo;
        same[i] = ok;
        if (!ok) ++moves;
        else if (first_same < 0) first_same = i;
      }
      if (moves % 2 == 0 || first_same >= 0) {
        int cur = 0;   // 0 = U, 1 = W
        for (int i = 0; i < np; ++i) {
          PassPlan& pp = (*list)[i];
          const bool move = !same[i] || (moves % 2 == 1 && i == first_same);
          pp.src = cur;
          pp.dst = move ? 1 - cur : cur;
          cur = pp.dst;
          use(pp);
        }
      } else {
        if (np < 2)
          return fail(PFFT_ERR_UNSUPPORTED, "in-place plan with a single transposing pass");
        for (int i = 0; i < np; ++i) {   // U -> W0 -> W1 -> W0 ... -> U
          PassPlan& pp = (*list)[i];
          pp.src = i == 0 ? 0 : 1 + ((i - 1) & 1);
          pp.dst = i == np - 1 ? 0 : 1 + (i & 1);
          use(pp);
        }
      }
    }
    p.in_bytes = p.out_bytes = ubytes;
  }
  w[0] = (w[0] + 255) / 256 * 256;
  w[1] = (w[1] + 255) / 256 * 256;
  p.work_bytes = w[0] + w[1];
  p.work0_bytes = w[0];
  return PFFT_OK;
}

std::string plan_describe(const Plan& p) {
  // GlobalFFTParams show (global_params.jl:69-75) + extra dims (plans.jl:555-562)
  std::ostringstream o;
  o << "Transforms: (";
  for (int d = 0; d < p.N; ++d) o << (d ? ", " : "") << kind_name(p.kinds[d]);
  o << (p.N == 1 ? ",)" : ")");
  o << "\nInput type: " << eltype_name(p.elt_in);
  auto tup = [&](const int64_t* v, int n) {
    o << "(";
    for (int i = 0; i < n; ++i) o << (i ? ", " : "") << v[i];
    o << (n == 1 ? ",)" : ")");
  };
  o << "\nGlobal dimensions: ";
  tup(p.size_in, p.N);
  o << " -> ";
  tup(p.size_out, p.N);
  if (p.Ne) {
    o << "\nExtra dimensions: ";
    tup(p.extra, p.Ne);
  }
  o << "\nMPI topology: 1D decomposition (1 processes), one GPU";
  return o.str();
}

}  // namespace pfft
