// Generic pass: any line length, any transform kind, float or double.
//
// One workgroup handles one tile of T lines.  The tile goes through LDS in
// four phases separated by workgroup barriers:
//   load   -- global -> raw tile in LDS (coalesced along the line when the
//             transform axis is contiguous, across the T lines otherwise)
//   pre    -- raw tile -> complex work lines (per-kind reorder / twiddle)
//   stages -- Stockham mixed-radix FFT, ping-pong between two LDS buffers
//   post   -- work lines -> output elements (per-kind twiddle / split), with
//             the store going straight to the next pencil's memory order
// Every phase is a plain function of (thread id, thread count, LDS base), so
// the same code runs in the GPU kernel and in the host emulator used by the
// CPU test-suite.
#pragma once

#include "common.hpp"

namespace pfft {

struct TileCtx {
  int64_t in_base;   // element offset of the tile's first line (input)
  int64_t out_base;  // element offset of the tile's first line (output)
  int64_t tile;      // tile index
  int64_t m0;        // index along line dim 0 of the tile's first line
  int32_t Tact;      // lines present in this tile
  int32_t Tw;        // work lines (two per line for complex R2R data)
};

// Mixed-radix digits of e = e0, e0 + step, e0 + 2 step, ... (the strided
// loops of the phases): one division per digit up front, then additions with
// carries, so the loop bodies divide by none of the run-time extents.
// Digit i < N - 1 runs over [0, rad[i]); the last one is unbounded.
template <int N>
struct Digits {
  uint32_t d[N], rad[N], s[N];
  PFFT_HD Digits(uint32_t e0, uint32_t step, const uint32_t (&radices)[N > 1 ? N - 1 : 1]) {
    for (int i = 0; i < N - 1; ++i) {
      rad[i] = radices[i];
      d[i] = e0 % rad[i];
      e0 /= rad[i];
      s[i] = step % rad[i];
      step /= rad[i];
    }
    d[N - 1] = e0;
    s[N - 1] = step;
    rad[N - 1] = 0;
  }
  PFFT_HD void next() {
    uint32_t carry = 0;
    for (int i = 0; i < N - 1; ++i) {
      d[i] += s[i] + carry;   // < 2 rad[i]: one subtraction at most
      carry = d[i] >= rad[i];
      if (carry) d[i] -= rad[i];
    }
    d[N - 1] += s[N - 1] + carry;
  }
};

// Work-line slot of point k: one padding slot after every 8 points, so
// that the butterflies' stride-RAD writes (and the stride-8 ones of the
// second stage) spread over the LDS banks instead of hitting a few of them.
// The line pitch Mp covers wp(Mw - 1) + 1 slots.
PFFT_HD uint32_t wp(uint32_t k) { return k + (k >> 3); }

// Tile index -> (block b0 of T lines along line dim 0, linear index `rest`
// over line dims 1..D-1).  Order 0 walks line dim 0 first (consecutive tiles
// are neighbouring blocks of one run); order 1 walks the other dims first, so
// that workgroups running at the same time touch addresses spread along
// those dims rather than one far-strided run.  tile_rev walks the tiles in
// descending order: a pass then first reads what the pass before it wrote
// last, while that is still in the device's last-level cache.
PFFT_HD void tile_split(const PassArgs& a, int64_t tile, int64_t* b0, int64_t* rest) {
  if (a.tile_rev) tile = a.ntiles - 1 - tile;
  const int64_t nb0 = (a.ldim[0] + a.T - 1) / a.T;
  if (a.tile_order == 0) {
    *b0 = tile % nb0;
    *rest = tile / nb0;
  } else {
    const int64_t nrest = a.ntiles / nb0;
    *b0 = tile / nrest;
    *rest = tile - *b0 * nrest;
  }
}

// Integer wavenumber of index c along logical dim d of the spectrum:
// rfftfreq along the r2c dim, fftfreq elsewhere (test/include/Grids.jl:44-59
// with a 2 pi box).  Spectrum lengths stay below 2^31.
template <typename R>
PFFT_HD R wavenumber(const PassArgs& a, int d, int c) {
  const int n = (int)a.kn[d];
  return a.kmode[d] == 1 ? R(c) : R(c < (n + 1) / 2 ? c : c - n);
}

// Wavenumber state of a spectral operator for one line: the contributions of
// the logical dims other than the pass's own, folded into (a, b):
//   derivative: a = k_dim (when dim is not the pass's);  |k|^2 forms: a = sum
//   of k_d^2;  dealiasing: b = 0 when some |k_d| is past the 2/3 limit.
template <typename R>
struct OpLine {
  R a, b;
};

template <typename R>
PFFT_HD void op_add(const PassArgs& a, int d, int c, OpLine<R>& o) {
  const int ld = a.llog[d];
  if (ld < 0) return;
  const R k = wavenumber<R>(a, ld, c);
  switch (a.op) {
    case 1: if (ld == a.op_dim) o.a = k; break;
    case 2:
    case 3: o.a += k * k; break;
    case 4: if (fabs(k) > R(2.0 / 3.0) * R(a.kmax[ld])) o.b = R(0); break;
    default: break;
  }
}

// The part shared by every line of a tile (line dims 1..D-1); *m0 receives
// the index along line dim 0 of the tile's first line.
template <typename R>
PFFT_HD OpLine<R> op_tile(const PassArgs& a, int64_t tile, int64_t* m0) {
  int64_t b0, rest;
  tile_split(a, tile, &b0, &rest);
  *m0 = b0 * a.T;
  OpLine<R> o = {R(0), R(1)};
  for (int d = 1; d < a.D; ++d) {
    op_add<R>(a, d, (int)(rest % a.ldim[d]), o);
    rest /= a.ldim[d];
  }
  return o;
}

// Line m0 + t of the tile: adds line dim 0.
template <typename R>
PFFT_HD OpLine<R> op_line(const PassArgs& a, OpLine<R> o, int64_t m) {
  op_add<R>(a, 0, (int)m, o);
  return o;
}

// alpha m(k) (conjugated for the transposes) at position j along the pass dim.
template <typename R>
PFFT_HD cpx<R> op_factor(const PassArgs& a, const OpLine<R>& o, int j) {
  const R kp = wavenumber<R>(a, a.pdim, j);
  R re = R(0), im = R(0);
  switch (a.op) {
    case 1: im = a.op_dim == a.pdim ? kp : o.a; break;
    case 2: re = -(o.a + kp * kp); break;
    case 3: { const R s2 = o.a + kp * kp; re = s2 > R(0) ? R(-1) / s2 : R(0); break; }
    case 4:
      re = (o.b != R(0) && fabs(kp) <= R(2.0 / 3.0) * R(a.kmax[a.pdim])) ? R(1) : R(0);
      break;
    default: re = R(1); break;
  }
  const R al = (R)a.op_alpha;
  return mk(re * al, (a.op_conj ? -im : im) * al);
}

PFFT_HD void tile_init(const PassArgs& a, int64_t tile, TileCtx& c) {
  int64_t b0, rest;
  tile_split(a, tile, &b0, &rest);
  int64_t ib = b0 * a.T * a.lsin[0];
  int64_t ob = b0 * a.T * a.lsout[0];
  for (int d = 1; d < a.D; ++d) {
    const int64_t m = rest % a.ldim[d];
    rest /= a.ldim[d];
    ib += m * a.lsin[d];
    ob += m * a.lsout[d];
  }
  c.in_base = ib;
  c.out_base = ob;
  const int64_t left = a.ldim[0] - b0 * a.T;
  c.Tact = (int32_t)(left < a.T ? left : a.T);
  c.tile = tile;
  c.m0 = b0 * a.T;
  // real lines of the tile: two per line for complex R2R data; paired as
  // the real and imaginary parts of one work line for DCT-I / DST-I
  const int32_t nreal = a.r2r_cplx ? 2 * c.Tact : c.Tact;
  c.Tw = a.pair_real ? (nreal + 1) / 2 : nreal;
}

// ---- phase 1: load -------------------------------------------------------
template <typename R>
PFFT_HD void phase_load(const PassArgs& a, const TileCtx& c, int tid, int nthr,
                        char* lds) {
  R* raw = (R*)(lds + a.lds_w1_off);
  const uint32_t L = (uint32_t)a.L_in;
  const uint32_t Tact = (uint32_t)c.Tact;
  const uint32_t count = Tact * L;
  const bool along = (a.sk_in == 1);
  if (a.load_w0 && a.kind == L_R2C_HALF) {
    // r2c of even length: the real pairs (x_2j, x_2j+1) straight into the
    // work lines as complex points
    cpx<R>* W = (cpx<R>*)lds;
    const uint32_t M = (uint32_t)a.M;
    const R* in = (const R*)a.in;
    Digits<2> jt(tid, nthr, {along ? M : Tact});
    for (uint32_t e = tid; e < Tact * M; e += nthr, jt.next()) {
      const uint32_t t = along ? jt.d[1] : jt.d[0];
      const uint32_t j = along ? jt.d[0] : jt.d[1];
      const int64_t g = c.in_base + (int64_t)t * a.lsin[0] + (int64_t)(2 * j) * a.sk_in;
      W[t * a.Mp + wp(j)] = mk(in[g], in[g + a.sk_in]);
    }
    return;
  }
  if (a.load_w0 && a.kind == L_C2R_HALF) {
    // c2r of even length: the Hermitian merge of X_j with X_{M-j}, both read
    // from memory (each wavefront's reads run up, or down, one contiguous
    // stretch; the second touch of a point comes from the caches)
    cpx<R>* W = (cpx<R>*)lds;
    const uint32_t M = (uint32_t)a.M;
    const cpx<R>* in = (const cpx<R>*)a.in;
    const cpx<R>* aux = (const cpx<R>*)a.aux;
    Digits<2> jt(tid, nthr, {along ? M : Tact});
    for (uint32_t e = tid; e < Tact * M; e += nthr, jt.next()) {
      const uint32_t t = along ? jt.d[1] : jt.d[0];
      const uint32_t j = along ? jt.d[0] : jt.d[1];
      const int64_t g = c.in_base + (int64_t)t * a.lsin[0];
      cpx<R> A = in[g + (int64_t)j * a.sk_in];
      cpx<R> B = conj(in[g + (int64_t)(M - j) * a.sk_in]);
      if (j == 0) { A.im = R(0); B.im = R(0); }  // Im of DC and Nyquist ignored
      else if (a.adjoint) { A = scl(A, R(0.5)); B = scl(B, R(0.5)); }  // X_k / w_k
      const cpx<R> E = A + B;
      const cpx<R> O = (A - B) * conj(aux[j]);
      W[t * a.Mp + wp(j)] = E + mul_pi(O);
    }
    return;
  }
  Digits<2> it(tid, nthr, {along ? L : Tact});
  if (a.load_w0) {
    // c2c: straight into the work lines (the pre phase has nothing to do),
    // with the spectral operator of side 1
    cpx<R>* W = (cpx<R>*)lds;
    for (uint32_t e = tid; e < count; e += nthr, it.next()) {
      const uint32_t t = along ? it.d[1] : it.d[0];
      const uint32_t j = along ? it.d[0] : it.d[1];
      const int64_t g = c.in_base + (int64_t)t * a.lsin[0] + (int64_t)j * a.sk_in;
      cpx<R> v = ((const cpx<R>*)a.in)[g];
      if (a.op && a.op_side == 1) {
        int64_t m0;
        const OpLine<R> ot = op_tile<R>(a, c.tile, &m0);
        v = v * op_factor<R>(a, op_line<R>(a, ot, m0 + t), (int)j);
      }
      W[t * a.Mp + wp(j)] = v;
    }
    return;
  }
  for (uint32_t e = tid; e < count; e += nthr, it.next()) {
    const uint32_t t = along ? it.d[1] : it.d[0];
    const uint32_t j = along ? it.d[0] : it.d[1];
    const int64_t g = c.in_base + (int64_t)t * a.lsin[0] + (int64_t)j * a.sk_in;
    if (a.in_cplx) {
      const cpx<R> v = ((const cpx<R>*)a.in)[g];
      *(cpx<R>*)(raw + t * a.raw_pitch + 2 * j) = v;
    } else {
      raw[t * a.raw_pitch + j] = ((const R*)a.in)[g];
    }
  }
}

// real scalar j of work line w (real data, or one part of complex R2R data)
template <typename R>
PFFT_HD R raw_real(const PassArgs& a, const R* raw, uint32_t w, int64_t j) {
  if (a.r2r_cplx) return raw[(w >> 1) * a.raw_pitch + 2 * j + (w & 1)];
  return raw[w * a.raw_pitch + j];
}
template <typename R>
PFFT_HD cpx<R> raw_cplx(const PassArgs& a, const R* raw, uint32_t t, int64_t j) {
  return *(const cpx<R>*)(raw + t * a.raw_pitch + 2 * j);
}

// ---- phase 2: pre-process into work buffer 0 ------------------------------
template <typename R>
PFFT_HD void phase_pre(const PassArgs& a, const TileCtx& c, int tid, int nthr,
                       char* lds) {
  if (a.kind == L_COPY || a.load_w0) return;
  const R* raw = (const R*)(lds + a.lds_w1_off);
  cpx<R>* W = (cpx<R>*)lds;
  const cpx<R>* aux = (const cpx<R>*)a.aux;
  const uint32_t M = (uint32_t)a.M;
  const int64_t n = a.n;
  const uint32_t count = (uint32_t)c.Tw * M;
  const R zero = R(0);
  Digits<2> it(tid, nthr, {M});
  for (uint32_t e = tid; e < count; e += nthr, it.next()) {
    const uint32_t w = it.d[1];
    const uint32_t j = it.d[0];
    cpx<R> v;
    switch (a.kind) {
      case L_C2C:
        v = raw_cplx(a, raw, w, j);
        if (a.op && a.op_side == 1) {
          int64_t m0;
          const OpLine<R> ot = op_tile<R>(a, c.tile, &m0);
          v = v * op_factor<R>(a, op_line<R>(a, ot, m0 + w), (int)j);
        }
        break;
      case L_R2C_HALF:
        v = mk(raw_real(a, raw, w, 2 * j), raw_real(a, raw, w, 2 * j + 1));
        break;
      case L_R2C_FULL:
        v = mk(raw_real(a, raw, w, j), zero);
        break;
      case L_DHT: {
        // real lines 2w and 2w + 1 (paired; split after the FFT)
        const uint32_t nreal = a.r2r_cplx ? 2 * (uint32_t)c.Tact : (uint32_t)c.Tact;
        v = mk(raw_real(a, raw, 2 * w, j),
               2 * w + 1 < nreal ? raw_real(a, raw, 2 * w + 1, j) : zero);
        break;
      }
      case L_C2R_HALF: {
        cpx<R> A = raw_cplx(a, raw, w, j);
        cpx<R> B = conj(raw_cplx(a, raw, w, M - j));
        if (j == 0) { A.im = zero; B.im = zero; }  // Im of DC and Nyquist ignored
        else if (a.adjoint) { A = scl(A, R(0.5)); B = scl(B, R(0.5)); }  // X_k / w_k
        const cpx<R> E = A + B;
        const cpx<R> O = (A - B) * conj(aux[j]);
        v = E + mul_pi(O);
        break;
      }
      case L_C2R_FULL: {
        if (j == 0) v = mk(raw_cplx(a, raw, w, 0).re, zero);
        else if (2 * (int64_t)j < n) v = raw_cplx(a, raw, w, j);
        else v = conj(raw_cplx(a, raw, w, n - j));
        if (a.adjoint && j != 0) v = scl(v, R(0.5));  // X_k / w_k
        break;
      }
      case L_DCT1: {
        // even extension of real lines 2w and 2w + 1 (paired)
        const int64_t src = j < n ? j : M - j;
        const uint32_t nreal = a.r2r_cplx ? 2 * (uint32_t)c.Tact : (uint32_t)c.Tact;
        v = mk(raw_real(a, raw, 2 * w, src),
               2 * w + 1 < nreal ? raw_real(a, raw, 2 * w + 1, src) : zero);
        break;
      }
      case L_DCT2:
      case L_DST2: {
        // Makhoul reorder of real lines 2w and 2w + 1 (paired): their
        // transforms are separated by the Hermitian split in the post phase
        const int64_t h = (n + 1) / 2;
        const uint32_t nreal = a.r2r_cplx ? 2 * (uint32_t)c.Tact : (uint32_t)c.Tact;
        const bool hi = 2 * w + 1 < nreal;
        const int64_t src = (int64_t)j < h ? 2 * (int64_t)j : 2 * n - 2 * (int64_t)j - 1;
        const R sg = ((int64_t)j >= h && a.kind == L_DST2) ? R(-1) : R(1);
        v = mk(sg * raw_real(a, raw, 2 * w, src), hi ? sg * raw_real(a, raw, 2 * w + 1, src) : zero);
        break;
      }
      case L_DCT3:
      case L_DST3: {
        // V_j = w_4n^-j (x_j - i x_{n-j}) of real lines 2w and 2w + 1, as
        // V_a + i V_b (V Hermitian: the transform is real, one line per part)
        const uint32_t nreal = a.r2r_cplx ? 2 * (uint32_t)c.Tact : (uint32_t)c.Tact;
        const bool hi = 2 * w + 1 < nreal;
        const bool dst = a.kind == L_DST3;
        const int64_t ik = dst ? n - 1 - (int64_t)j : (int64_t)j;
        const int64_t ink = dst ? (int64_t)j - 1 : n - (int64_t)j;
        const cpx<R> wj = conj(aux[j]);
        const cpx<R> va = wj * mk(raw_real(a, raw, 2 * w, ik),
                                  j == 0 ? zero : -raw_real(a, raw, 2 * w, ink));
        const cpx<R> vb = hi ? wj * mk(raw_real(a, raw, 2 * w + 1, ik),
                                       j == 0 ? zero : -raw_real(a, raw, 2 * w + 1, ink))
                             : mk(zero, zero);
        v = va + mul_pi(vb);
        break;
      }
      case L_DCT4:
        v = (int64_t)j < n ? scl(aux[j], raw_real(a, raw, w, j)) : mk(zero, zero);
        break;
      case L_DST4:
        v = (int64_t)j < n ? scl(aux[j], raw_real(a, raw, w, n - 1 - j))
                           : mk(zero, zero);
        break;
      case L_DCT4H:
      case L_DST4H:
        // even slots: the half z_m = (x_2m + i x_{n-1-2m}) c_m (reversed
        // line for the DST); odd slots: the second real line, none here
        if (j & 1) {
          v = mk(zero, zero);
        } else {
          const int64_t m2 = j;   // 2m
          const R xa = raw_real(a, raw, w, m2), xb = raw_real(a, raw, w, n - 1 - m2);
          v = aux[j >> 1] * (a.kind == L_DCT4H ? mk(xa, xb) : mk(xb, xa));
        }
        break;
      case L_DST1: {
        // odd extension of real lines 2w and 2w + 1 (paired)
        const int64_t jj = j;
        const uint32_t nreal = a.r2r_cplx ? 2 * (uint32_t)c.Tact : (uint32_t)c.Tact;
        const bool hi = 2 * w + 1 < nreal;
        if (jj == 0 || jj == n + 1) {
          v = mk(zero, zero);
        } else if (jj <= n) {
          v = mk(raw_real(a, raw, 2 * w, jj - 1), hi ? raw_real(a, raw, 2 * w + 1, jj - 1) : zero);
        } else {
          const int64_t src = (int64_t)M - jj - 1;
          v = mk(-raw_real(a, raw, 2 * w, src), hi ? -raw_real(a, raw, 2 * w + 1, src) : zero);
        }
        break;
      }
      default:
        v = mk(zero, zero);
    }
    W[w * a.Mp + wp(j)] = v;
  }
}

// ---- phase 3: one Stockham stage ------------------------------------------
// Radix-R stage with Ns = product of the previous radices (autosort form):
//   v_r = src[j + r M/R] * w_M^(r (j mod Ns) M/(Ns R)),  y = DFT_R(v),
//   dst[(j / Ns) Ns R + (j mod Ns) + r Ns] = y_r.
// Backward stages use DFT_bwd(v) = conj(DFT_fwd(conj(v))).
template <typename R>
PFFT_HD void dft2(cpx<R>* v) {
  const cpx<R> a = v[0], b = v[1];
  v[0] = a + b; v[1] = a - b;
}
template <typename R>
PFFT_HD void dft3(cpx<R>* v) {
  const R c = R(-0.5), s = R(0.86602540378443864676372317075294);
  const cpx<R> t = v[1] + v[2];
  const cpx<R> d = v[1] - v[2];
  const cpx<R> m = v[0] + scl(t, c);
  const cpx<R> u = mk(d.im * s, -d.re * s);  // -i s d
  v[0] = v[0] + t;
  v[1] = m + u;
  v[2] = m - u;
}
template <typename R>
PFFT_HD void dft4(cpx<R>* v) {
  const cpx<R> t0 = v[0] + v[2], t1 = v[0] - v[2];
  const cpx<R> t2 = v[1] + v[3], t3 = mul_mi(v[1] - v[3]);
  v[0] = t0 + t2; v[2] = t0 - t2; v[1] = t1 + t3; v[3] = t1 - t3;
}
template <typename R>
PFFT_HD void dft5(cpx<R>* v) {
  const R c1 = R(0.30901699437494742410229341718282);
  const R c2 = R(-0.80901699437494742410229341718282);
  const R s1 = R(0.95105651629515357211643933337938);
  const R s2 = R(0.58778525229247312916870595463907);
  const cpx<R> a0 = v[0];
  const cpx<R> sa = v[1] + v[4], da = v[1] - v[4];
  const cpx<R> sb = v[2] + v[3], db = v[2] - v[3];
  const cpx<R> p1 = a0 + scl(sa, c1) + scl(sb, c2);
  const cpx<R> p2 = a0 + scl(sa, c2) + scl(sb, c1);
  const cpx<R> q1 = scl(da, s1) + scl(db, s2);   // times -i
  const cpx<R> q2 = scl(da, s2) - scl(db, s1);   // times -i
  v[0] = a0 + sa + sb;
  v[1] = p1 + mul_mi(q1);
  v[4] = p1 - mul_mi(q1);
  v[2] = p2 + mul_mi(q2);
  v[3] = p2 - mul_mi(q2);
}
template <typename R>
PFFT_HD void dft8(cpx<R>* v) {
  const R h = R(0.70710678118654752440084436210485);
  cpx<R> b[4], c[4];
  for (int r = 0; r < 4; ++r) { b[r] = v[r] + v[r + 4]; c[r] = v[r] - v[r + 4]; }
  c[1] = mk((c[1].re + c[1].im) * h, (c[1].im - c[1].re) * h);   // * (1-i)/sqrt2
  c[2] = mul_mi(c[2]);                                           // * -i
  c[3] = mk((c[3].im - c[3].re) * h, -(c[3].re + c[3].im) * h);  // * (-1-i)/sqrt2
  dft4(b); dft4(c);
  v[0] = b[0]; v[2] = b[1]; v[4] = b[2]; v[6] = b[3];
  v[1] = c[0]; v[3] = c[1]; v[5] = c[2]; v[7] = c[3];
}

template <typename R, int RAD>
PFFT_HD void stage_fixed(const PassArgs& a, const TileCtx& c, int64_t Ns, int64_t fftM, bool bwd,
                         const cpx<R>* src, cpx<R>* dst, int tid, int nthr) {
  const cpx<R>* tw = (const cpx<R>*)a.tw;
  const uint32_t M = (uint32_t)fftM;
  const uint32_t nb = M / RAD;                 // butterflies per line
  const uint32_t step = (uint32_t)(M / (Ns * RAD));
  const uint32_t count = (uint32_t)c.Tw * nb;
  // e = (w nb + j), j = jq Ns + jj
  Digits<3> it(tid, nthr, {(uint32_t)Ns, nb / (uint32_t)Ns});
  for (uint32_t e = tid; e < count; e += nthr, it.next()) {
    const uint32_t jj = it.d[0], jq = it.d[1], w = it.d[2];
    const uint32_t j = jq * (uint32_t)Ns + jj;
    const cpx<R>* s = src + w * a.Mp;
    cpx<R> v[RAD];
    for (int r = 0; r < RAD; ++r) {
      cpx<R> x = s[wp(j + r * nb)];
      if (bwd) x = conj(x);
      v[r] = r == 0 ? x : x * tw[r * jj * step];
    }
    if (RAD == 2) dft2(v);
    else if (RAD == 3) dft3(v);
    else if (RAD == 4) dft4(v);
    else if (RAD == 5) dft5(v);
    else if (RAD == 8) dft8(v);
    cpx<R>* d = dst + w * a.Mp;
    const uint32_t ob = jq * (uint32_t)Ns * RAD + jj;
    for (int r = 0; r < RAD; ++r) d[wp(ob + r * (uint32_t)Ns)] = bwd ? conj(v[r]) : v[r];
  }
}

// Prime radix 7, 11 or 13: one butterfly per work item (the RAD inputs
// read once), the DFT in the symmetric form
//   y_0 = v_0 + sum_r A_r,  y_q, y_{RAD-q} = P_q -+ i Q_q,
//   P_q = v_0 + sum_r A_r cos(2 pi q r / RAD),  Q_q = sum_r B_r sin(2 pi q r / RAD),
// A_r = v_r + v_{RAD-r}, B_r = v_r - v_{RAD-r} (r, q = 1 .. (RAD-1)/2); the
// cosines and sines are the table's w_M^(m M / RAD) = cos - i sin.
template <typename R, int RAD>
PFFT_HD void stage_prime(const PassArgs& a, const TileCtx& c, int64_t Ns, int64_t fftM, bool bwd,
                         const cpx<R>* src, cpx<R>* dst, int tid, int nthr) {
  constexpr int H = (RAD - 1) / 2;
  const cpx<R>* tw = (const cpx<R>*)a.tw;
  const uint32_t M = (uint32_t)fftM;
  const uint32_t nb = M / RAD;
  const uint32_t step = (uint32_t)(M / (Ns * RAD));
  const uint32_t count = (uint32_t)c.Tw * nb;
  Digits<3> it(tid, nthr, {(uint32_t)Ns, nb / (uint32_t)Ns});
  for (uint32_t e = tid; e < count; e += nthr, it.next()) {
    const uint32_t jj = it.d[0], jq = it.d[1], w = it.d[2];
    const uint32_t j = jq * (uint32_t)Ns + jj;
    const cpx<R>* s = src + w * a.Mp;
    cpx<R> v[RAD];
    for (int r = 0; r < RAD; ++r) {
      cpx<R> x = s[wp(j + r * nb)];
      if (bwd) x = conj(x);
      v[r] = r == 0 ? x : x * tw[r * jj * step];
    }
    cpx<R> A[H + 1], B[H + 1];
    cpx<R> y0 = v[0];
    for (int r = 1; r <= H; ++r) {
      A[r] = v[r] + v[RAD - r];
      B[r] = v[r] - v[RAD - r];
      y0 = y0 + A[r];
    }
    cpx<R>* d = dst + w * a.Mp;
    const uint32_t ob = jq * (uint32_t)Ns * RAD + jj;
    d[wp(ob)] = bwd ? conj(y0) : y0;
    for (int q = 1; q <= H; ++q) {
      cpx<R> P = v[0], Q = mk(R(0), R(0));
      for (int r = 1; r <= H; ++r) {
        const cpx<R> wm = tw[((q * r) % RAD) * nb];   // cos - i sin
        P = P + scl(A[r], wm.re);
        Q = Q + scl(B[r], -wm.im);
      }
      const cpx<R> lo = P + mul_mi(Q), hi = P - mul_mi(Q);   // P -+ i Q
      d[wp(ob + q * (uint32_t)Ns)] = bwd ? conj(lo) : lo;
      d[wp(ob + (RAD - q) * (uint32_t)Ns)] = bwd ? conj(hi) : hi;
    }
  }
}

// Any radix (e.g. a prime factor 17, 41): one work item per output.
template <typename R>
PFFT_HD void stage_any(const PassArgs& a, const TileCtx& c, int64_t Ns, int RAD, int64_t fftM,
                       bool bwd, const cpx<R>* src, cpx<R>* dst, int tid, int nthr) {
  const cpx<R>* tw = (const cpx<R>*)a.tw;
  const uint32_t M = (uint32_t)fftM;
  const uint32_t nb = M / RAD;
  const uint32_t step = (uint32_t)(M / (Ns * RAD));
  const uint32_t count = (uint32_t)c.Tw * M;
  // e = w M + q nb + j, j = jq Ns + jj; q: output index within the butterfly
  Digits<4> it(tid, nthr, {(uint32_t)Ns, nb / (uint32_t)Ns, (uint32_t)RAD});
  for (uint32_t e = tid; e < count; e += nthr, it.next()) {
    const uint32_t jj = it.d[0], jq = it.d[1], q = it.d[2], w = it.d[3];
    const uint32_t j = jq * (uint32_t)Ns + jj;
    const cpx<R>* s = src + w * a.Mp;
    cpx<R> acc = mk(R(0), R(0));
    // twiddle of the stage times the DFT_R kernel entry w_R^(r q): the table
    // index (r jj step + ((r q) mod RAD) nb) mod M, its two terms advanced
    // by additions (jj step < M / RAD, q < RAD)
    uint32_t ta = 0, tb = 0;
    for (int r = 0; r < RAD; ++r) {
      cpx<R> x = s[wp(j + r * nb)];
      if (bwd) x = conj(x);
      uint32_t idx = ta + tb * nb;
      if (idx >= M) idx -= M;
      acc = acc + x * tw[idx];
      ta += jj * step;
      if (ta >= M) ta -= M;
      tb += q;
      if (tb >= (uint32_t)RAD) tb -= RAD;
    }
    dst[w * a.Mp + wp(jq * (uint32_t)Ns * RAD + jj + q * (uint32_t)Ns)] =
        bwd ? conj(acc) : acc;
  }
}

// Stage s of the line FFT (fftM points, radix schedule a.radix), from the
// work buffer of parity sb to the other one; bwd: backward butterflies.
// PRIME: radix 7, 11 and 13 as butterflies (stage_prime; its registers set
// the kernel's, so only the instance for lines with such stages has it),
// else as one work item per output (stage_any)
template <typename R, bool PRIME = true>
PFFT_HD void phase_stage_ex(const PassArgs& a, const TileCtx& c, int s, int sb, int64_t fftM,
                            bool bwd, int tid, int nthr, char* lds) {
  int64_t Ns = 1;
  for (int i = 0; i < s; ++i) Ns *= a.radix[i];
  const cpx<R>* src = (const cpx<R>*)(lds + ((sb & 1) ? a.lds_w1_off : 0));
  cpx<R>* dst = (cpx<R>*)(lds + ((sb & 1) ? 0 : a.lds_w1_off));
  switch (a.radix[s]) {
    case 2: stage_fixed<R, 2>(a, c, Ns, fftM, bwd, src, dst, tid, nthr); break;
    case 3: stage_fixed<R, 3>(a, c, Ns, fftM, bwd, src, dst, tid, nthr); break;
    case 4: stage_fixed<R, 4>(a, c, Ns, fftM, bwd, src, dst, tid, nthr); break;
    case 5: stage_fixed<R, 5>(a, c, Ns, fftM, bwd, src, dst, tid, nthr); break;
    case 8: stage_fixed<R, 8>(a, c, Ns, fftM, bwd, src, dst, tid, nthr); break;
    default:
      if constexpr (PRIME) {
        if (a.radix[s] == 7) { stage_prime<R, 7>(a, c, Ns, fftM, bwd, src, dst, tid, nthr); break; }
        if (a.radix[s] == 11) { stage_prime<R, 11>(a, c, Ns, fftM, bwd, src, dst, tid, nthr); break; }
        if (a.radix[s] == 13) { stage_prime<R, 13>(a, c, Ns, fftM, bwd, src, dst, tid, nthr); break; }
      }
      stage_any<R>(a, c, Ns, a.radix[s], fftM, bwd, src, dst, tid, nthr);
      break;
  }
}

// Stage s of the direct M-point line FFT.
template <typename R, bool PRIME = true>
PFFT_HD void phase_stage(const PassArgs& a, const TileCtx& c, int s, int tid,
                         int nthr, char* lds) {
  phase_stage_ex<R, PRIME>(a, c, s, s, a.M, a.bwd != 0, tid, nthr, lds);
}

// ---- Bluestein form of the line FFT ----------------------------------------
// The M-point transform of the pre-processed lines in work buffer 0, all in
// place there (one work buffer, so more workgroups fit on a CU):
//   bs_pre   a_j = x_j c_j (x conjugated first for a backward transform),
//            zero-padded to bs_M points;
//   dif      the bs_M-point forward FFT, decimation in frequency over the
//            radix schedule: bins land in digit-reversed positions;
//   bs_mul   times the transform of b, stored in those positions (its
//            1 / bs_M included);
//   dit      the bs_M-point backward FFT, decimation in time over the
//            stages in reverse: natural order again;
//   bs_post  X_k = c_k C_k (conjugated back for a backward transform), k < M.
//
// In-place stage s of radix RAD over blocks of Bs = bs_M / (r_0 ... r_{s-1})
// points: butterfly j < Bs / RAD of a block reads positions
// j + q Bs / RAD (q < RAD).  Forward: DFT_RAD, then output q times
// w_Bs^(q j).  Backward: input q times w_Bs^(-q j), then the conjugate DFT.
template <typename R, int RAD>
PFFT_HD void stage_inplace(const PassArgs& a, const TileCtx& c, int s, bool bwd, cpx<R>* W,
                           int tid, int nthr) {
  const cpx<R>* tw = (const cpx<R>*)a.tw;
  const uint32_t Mb = (uint32_t)a.bs_M;
  uint32_t Bs = Mb;
  for (int i = 0; i < s; ++i) Bs /= (uint32_t)a.radix[i];
  const uint32_t stride = Bs / RAD, tstep = Mb / Bs;
  const uint32_t nb = Mb / RAD;                      // butterflies per line
  const uint32_t count = (uint32_t)c.Tw * nb;
  // e = (w (Mb / Bs) + blk) stride + j
  Digits<3> it(tid, nthr, {stride, Mb / Bs});
  for (uint32_t e = tid; e < count; e += nthr, it.next()) {
    const uint32_t j = it.d[0], blk = it.d[1], w = it.d[2];
    cpx<R>* L = W + w * a.Mp;
    const uint32_t base = blk * Bs + j;
    cpx<R> v[RAD];
    for (int q = 0; q < RAD; ++q) {
      v[q] = L[wp(base + q * stride)];
      if (bwd) v[q] = conj(v[q]);
      if (bwd && q) v[q] = v[q] * tw[q * j * tstep];   // conj(v w^-qj)
    }
    if (RAD == 2) dft2(v);
    else if (RAD == 4) dft4(v);
    else dft8(v);
    for (int q = 0; q < RAD; ++q) {
      cpx<R> y = v[q];
      if (!bwd && q) y = y * tw[q * j * tstep];
      L[wp(base + q * stride)] = bwd ? conj(y) : y;
    }
  }
}

template <typename R>
PFFT_HD void phase_stage_inplace(const PassArgs& a, const TileCtx& c, int s, bool bwd, int tid,
                                 int nthr, char* lds) {
  cpx<R>* W = (cpx<R>*)lds;
  switch (a.radix[s]) {
    case 2: stage_inplace<R, 2>(a, c, s, bwd, W, tid, nthr); break;
    case 4: stage_inplace<R, 4>(a, c, s, bwd, W, tid, nthr); break;
    default: stage_inplace<R, 8>(a, c, s, bwd, W, tid, nthr); break;
  }
}
template <typename R>
PFFT_HD void phase_bs_pre(const PassArgs& a, const TileCtx& c, int tid, int nthr, char* lds) {
  cpx<R>* W = (cpx<R>*)lds;
  const cpx<R>* ch = (const cpx<R>*)a.bs_chirp;
  const uint32_t Mb = (uint32_t)a.bs_M, M = (uint32_t)a.M;
  const uint32_t count = (uint32_t)c.Tw * Mb;
  Digits<2> it(tid, nthr, {Mb});
  for (uint32_t e = tid; e < count; e += nthr, it.next()) {
    const uint32_t w = it.d[1], j = it.d[0];
    cpx<R>* x = W + w * a.Mp + wp(j);
    if (j < M) {
      const cpx<R> v = a.bwd ? conj(*x) : *x;
      *x = v * ch[j];
    } else {
      *x = mk(R(0), R(0));
    }
  }
}

template <typename R>
PFFT_HD void phase_bs_mul(const PassArgs& a, const TileCtx& c, int tid, int nthr, char* lds) {
  cpx<R>* W = (cpx<R>*)lds;
  const cpx<R>* bh = (const cpx<R>*)a.bs_bhat;
  const uint32_t Mb = (uint32_t)a.bs_M;
  const uint32_t count = (uint32_t)c.Tw * Mb;
  Digits<2> it(tid, nthr, {Mb});
  for (uint32_t e = tid; e < count; e += nthr, it.next()) {
    const uint32_t w = it.d[1], j = it.d[0];
    cpx<R>* x = W + w * a.Mp + wp(j);
    *x = *x * bh[j];
  }
}

template <typename R>
PFFT_HD void phase_bs_post(const PassArgs& a, const TileCtx& c, int tid, int nthr, char* lds) {
  cpx<R>* W = (cpx<R>*)lds;
  const cpx<R>* ch = (const cpx<R>*)a.bs_chirp;
  const uint32_t M = (uint32_t)a.M;
  const uint32_t count = (uint32_t)c.Tw * M;
  Digits<2> it(tid, nthr, {M});
  for (uint32_t e = tid; e < count; e += nthr, it.next()) {
    const uint32_t w = it.d[1], k = it.d[0];
    cpx<R>* x = W + w * a.Mp + wp(k);
    const cpx<R> v = *x * ch[k];
    *x = a.bwd ? conj(v) : v;
  }
}

// ---- phase 3: the line FFT, as a sequence of barrier-separated steps -------
// Direct form: the radix stages.  Bluestein form: bs_pre, the forward
// stages, bs_mul, the backward stages, bs_post.
PFFT_HD int fft_steps(const PassArgs& a) {
  return a.bs_M ? 2 * a.nstages + 3 : a.nstages;
}

template <typename R, bool PRIME = true>
PFFT_HD void phase_fft_step(const PassArgs& a, const TileCtx& c, int i, int tid, int nthr,
                            char* lds) {
  if (!a.bs_M) { phase_stage<R, PRIME>(a, c, i, tid, nthr, lds); return; }
  const int ns = a.nstages;
  if (i == 0) phase_bs_pre<R>(a, c, tid, nthr, lds);
  else if (i <= ns) phase_stage_inplace<R>(a, c, i - 1, false, tid, nthr, lds);
  else if (i == ns + 1) phase_bs_mul<R>(a, c, tid, nthr, lds);
  else if (i <= 2 * ns + 1) phase_stage_inplace<R>(a, c, 2 * ns + 1 - i, true, tid, nthr, lds);
  else phase_bs_post<R>(a, c, tid, nthr, lds);
}

// ---- phase 4: post-process and store --------------------------------------
template <typename R>
PFFT_HD R post_real(const PassArgs& a, const cpx<R>* W, uint32_t w, int64_t k) {
  const cpx<R>* aux = (const cpx<R>*)a.aux;
  const cpx<R>* L = W + w * a.Mp;
  const int64_t n = a.n;
  switch (a.kind) {
    case L_C2R_HALF: return (k & 1) ? L[wp((uint32_t)(k >> 1))].im : L[wp((uint32_t)(k >> 1))].re;
    case L_C2R_FULL: return L[wp((uint32_t)(k))].re;
    case L_DCT1: {
      // real line w is part w & 1 of work line w >> 1
      const cpx<R> z = (W + (w >> 1) * a.Mp)[wp((uint32_t)k)];
      return (w & 1) ? z.im : z.re;
    }
    case L_DCT2:
    case L_DST2: {
      // real line w is part w & 1 of work line w >> 1: its transform from
      // Z_q and conj Z_{n-q} (q = k, or n - 1 - k for the DST), then
      // y = 2 Re(w_4n^q V_q)
      const cpx<R>* Z = W + (w >> 1) * a.Mp;
      const int64_t q = a.kind == L_DST2 ? n - 1 - k : k;
      const cpx<R> Zq = Z[wp((uint32_t)q)];
      const cpx<R> Zn = conj(Z[wp((uint32_t)(q == 0 ? 0 : n - q))]);
      const cpx<R> V = (w & 1) ? scl(mul_mi(Zq - Zn), R(0.5)) : scl(Zq + Zn, R(0.5));
      return R(2) * (aux[q] * V).re;
    }
    case L_DCT3:
    case L_DST3: {
      const int64_t m = (k & 1) ? n - 1 - (k >> 1) : (k >> 1);
      const cpx<R> z = (W + (w >> 1) * a.Mp)[wp((uint32_t)m)];
      const R y = (w & 1) ? z.im : z.re;
      return (a.kind == L_DST3 && (k & 1)) ? -y : y;
    }
    case L_DCT4:
    case L_DST4: {
      const cpx<R> z = aux[n + k] * L[wp((uint32_t)(k))];
      const R y = R(2) * z.re;
      return (a.kind == L_DST4 && (k & 1)) ? -y : y;
    }
    case L_DCT4H:
    case L_DST4H: {
      // y_2m = Re(conj(e_m) (W_m + W_{m+n/2})), y_{n-1-2m} = -Im(...)
      const int64_t h = n / 2;
      const int64_t m = (k & 1) ? h - 1 - (k >> 1) : (k >> 1);
      const cpx<R> z = conj(aux[h + m]) * (L[wp((uint32_t)m)] + L[wp((uint32_t)(m + h))]);
      const R y = (k & 1) ? -z.im : z.re;
      return (a.kind == L_DST4H && (k & 1)) ? -y : y;
    }
    case L_DST1: {
      // the odd extension's transform F of (a + i b) is S_b - i S_a
      const cpx<R> z = (W + (w >> 1) * a.Mp)[wp((uint32_t)(k + 1))];
      return (w & 1) ? z.re : -z.im;
    }
    case L_DHT: {
      // the transform of real line w (part w & 1 of work line w >> 1) by the
      // Hermitian split, then Re - Im
      const cpx<R>* Z = W + (w >> 1) * a.Mp;
      const cpx<R> Zk = Z[wp((uint32_t)k)];
      const cpx<R> Zn = conj(Z[wp((uint32_t)(k == 0 ? 0 : n - k))]);
      const cpx<R> V = (w & 1) ? scl(mul_mi(Zk - Zn), R(0.5)) : scl(Zk + Zn, R(0.5));
      return V.re - V.im;
    }
    default: return R(0);
  }
}

// Hermitian weight of bin k of an r2c of length n: 1 for the bins stored
// once (DC, and Nyquist for even n), 2 for the others.  The adjoint of the
// r2c multiplies by it; the adjoint of the c2r divides by it.
PFFT_HD int herm_weight(int64_t k, int64_t n) { return (k == 0 || 2 * k == n) ? 1 : 2; }

template <typename R>
PFFT_HD cpx<R> post_cplx(const PassArgs& a, const cpx<R>* W, uint32_t w, int64_t k) {
  const cpx<R>* L = W + w * a.Mp;
  cpx<R> y;
  if (a.kind == L_R2C_HALF) {
    const cpx<R>* aux = (const cpx<R>*)a.aux;
    const int64_t M = a.M;
    const cpx<R> Z = L[wp((uint32_t)(k == M ? 0 : k))];
    const cpx<R> Zc = conj(L[wp((uint32_t)(k == 0 ? 0 : M - k))]);
    const cpx<R> E = scl(Z + Zc, R(0.5));
    const cpx<R> O = scl(mul_mi(Z - Zc), R(0.5));
    y = E + aux[k] * O;
  } else {
    y = L[wp((uint32_t)(k))];  // C2C, R2C_FULL
  }
  if (a.adjoint && (a.kind == L_R2C_HALF || a.kind == L_R2C_FULL))
    y = scl(y, R(herm_weight(k, a.n)));
  return y;
}

template <typename R>
PFFT_HD void phase_post(const PassArgs& a, const TileCtx& c, int tid, int nthr,
                        char* lds) {
  // the line FFT's output: the buffer its last stage wrote (work buffer 0
  // after the Bluestein form's two FFTs)
  const cpx<R>* W =
      (const cpx<R>*)(lds + ((a.bs_M == 0 && (a.nstages & 1)) ? a.lds_w1_off : 0));
  const R* raw = (const R*)(lds + a.lds_w1_off);
  const uint32_t L = (uint32_t)a.L_out;
  const uint32_t Tact = (uint32_t)c.Tact;
  const uint32_t count = Tact * L;
  const bool along = (a.sk_out == 1);
  const R sc = (R)a.scale;
  Digits<2> it(tid, nthr, {along ? L : Tact});
  for (uint32_t e = tid; e < count; e += nthr, it.next()) {
    const uint32_t t = along ? it.d[1] : it.d[0];
    const uint32_t k = along ? it.d[0] : it.d[1];
    const int64_t g = c.out_base + (int64_t)t * a.lsout[0] + (int64_t)k * a.sk_out;
    if (a.out_cplx) {
      cpx<R> y;
      if (a.kind == L_COPY) y = raw_cplx(a, raw, t, k);
      else if (a.r2r_cplx) y = mk(post_real(a, W, 2 * t, k), post_real(a, W, 2 * t + 1, k));
      else y = post_cplx(a, W, t, k);
      if (a.fs_M) {
        const cpx<R> w = ((const cpx<R>*)a.fs_tw)[(c.m0 + t) * k];
        y = y * (a.bwd ? conj(w) : w);
      }
      if (a.do_scale) y = scl(y, sc);
      if (a.op && a.op_side == 2) {
        int64_t m0;
        const OpLine<R> ot = op_tile<R>(a, c.tile, &m0);
        y = y * op_factor<R>(a, op_line<R>(a, ot, m0 + t), (int)k);
        if (a.op_acc) y = y + ((const cpx<R>*)a.out)[g];
      }
      ((cpx<R>*)a.out)[g] = y;
    } else {
      R y = a.kind == L_COPY ? raw[t * a.raw_pitch + k] : post_real(a, W, t, k);
      if (a.do_scale) y = y * sc;
      ((R*)a.out)[g] = y;
    }
  }
}

}  // namespace pfft
