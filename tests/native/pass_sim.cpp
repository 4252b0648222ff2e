// Host model of one column of an NTT pass kernel (csrc/ntt_pass.hpp) in its
// signed-limb arithmetic, built from the same fr29.hpp functions, with every
// bound the kernel relies on checked:
//   * each limb of a sum or difference fits a signed 32-bit integer;
//   * each product column fits a signed 64-bit accumulator (checked in
//     128-bit arithmetic beside the 64-bit product);
//   * a packed pass output has balanced digits and a top digit in [-2^23, 2^23).
// The column is the kernel's: M' = 8 A points x[rho1 + A rho2]; DFT_8 over rho2
// in registers, the inner twiddle w_M'^(rho1 k2) (k2 = 0 reduced instead),
// DFT_A over rho1, then the pass twiddle and a pbd word (intermediate pass)
// or the canonical value (last pass) at k = k2 + 8 k1.
#include <stdint.h>
#include <string.h>

#include "fr29.hpp"

using namespace pl;

namespace {

int g_bad = 0;

struct V {
  int64_t l[9];
};

V from_u(const uint32_t* b) {
  V v;
  for (int i = 0; i < 9; i++) v.l[i] = (int32_t)b[i];
  return v;
}
void to_u(const V& v, uint32_t* u) {
  for (int i = 0; i < 9; i++) u[i] = (uint32_t)v.l[i];
}
void check32(const V& v) {
  for (int i = 0; i < 9; i++)
    if (v.l[i] < INT32_MIN || v.l[i] > INT32_MAX) g_bad++;
}
V add(const V& a, const V& b) {
  V s;
  for (int i = 0; i < 9; i++) s.l[i] = a.l[i] + b.l[i];
  check32(s);
  return s;
}
V sub(const V& a, const V& b) {
  V s;
  for (int i = 0; i < 9; i++) s.l[i] = a.l[i] - b.l[i];
  check32(s);
  return s;
}

// the column sums of mont29 in 128 bits: every partial sum below 2^63
void check_columns(const uint32_t* a, const uint32_t* w, bool bm) {
  uint32_t m[9];
  __int128 acc = 0;
  const __int128 lim = ((__int128)1) << 63;
  for (int k = 0; k < 17; k++) {
    const int lo = k > 8 ? k - 8 : 0, hi = k < 8 ? k : 8;
    for (int i = lo; i <= hi; i++) {
      acc += (__int128)(int32_t)a[i] * (int32_t)w[k - i];
      if (acc >= lim || acc < -lim) g_bad++;
    }
    for (int i = lo; i <= hi; i++)
      if (i < k) {
        acc -= (__int128)(int32_t)m[i] * (int64_t)r29(k - i);
        if (acc >= lim || acc < -lim) g_bad++;
      }
    const bool bal = k <= 8 ? bm : false;  // t digits do not feed later columns
    const uint32_t dig = bal ? sext29((uint32_t)(uint64_t)acc) : ((uint32_t)(uint64_t)acc & kMask29);
    if (k <= 8) m[k] = dig;
    acc = (acc - (int32_t)dig) >> 29;
  }
}

V mul(const V& x, const uint32_t* w, bool bm, bool bt) {
  uint32_t xu[9], t[9];
  to_u(x, xu);
  check_columns(xu, w, bm);
  if (bm && bt) mont29<true, true>(xu, w, t);
  else if (bt) mont29<false, true>(xu, w, t);
  else mont29<false, false>(xu, w, t);
  return from_u(t);
}

// o -= r if o >= r (o < 2r), the kernel's fr_reduce_once
void reduce_once(uint32_t* o) {
  static const uint32_t kR[8] = {0x00000001u, 0xffffffffu, 0xfffe5bfeu, 0x53bda402u,
                                 0x09a1d805u, 0x3339d808u, 0x299d7d48u, 0x73eda753u};
  uint32_t d[8];
  int64_t b = 0;
  for (int i = 0; i < 8; i++) {
    const int64_t v = (int64_t)o[i] - kR[i] - b;
    d[i] = (uint32_t)v;
    b = v < 0;
  }
  if (!b) memcpy(o, d, sizeof d);
}

int brev(int v, int bits) {
  int r = 0;
  for (int i = 0; i < bits; i++) r |= ((v >> i) & 1) << (bits - 1 - i);
  return r;
}

// radix-2 DIF of size 2^logb on x[0..2^logb), bit-reversed output
void dft(V* x, int logb, const uint32_t* w16b /* [8][9] */) {
  const int B = 1 << logb;
  for (int st = 0; st < logb; st++) {
    const int h = B >> (st + 1);
    for (int blk = 0; blk < (1 << st); blk++) {
      const int s = blk * 2 * h;
      for (int i = 0; i < h; i++) {
        V a = x[s + i], b = x[s + i + h];
        V sum = add(a, b);
        V d = sub(a, b);
        if (i != 0) d = mul(d, w16b + 9 * ((i * 8) / h), false, false);
        x[s + i] = sum;
        x[s + i + h] = d;
      }
    }
  }
}

}  // namespace

extern "C" {

// One column; returns the number of violated bounds (0 when every bound held).
// in: M' words (caller data if ext, else pbd words); w16b: [8][9] balanced
// digits of w_16^e R' for the direction; inner: 256 pbd words of w_256^e R';
// tw: M' pbd words (intermediate pass); scale: one pbd word (last pass,
// do_scale); out: M' words.
int sim_column(int la, int last, int ext, const uint32_t* in, const uint32_t* w16b, const uint32_t* inner,
               const uint32_t* tw, int do_scale, const uint32_t* scale, uint32_t* out) {
  g_bad = 0;
  const int A = 1 << la, Mp = 8 * A, m = 3 + la;
  V Y[64][8];
  for (int rho1 = 0; rho1 < A; rho1++) {
    V z[8];
    for (int r2 = 0; r2 < 8; r2++) {
      uint32_t b[9];
      const uint32_t* src = in + 8 * (rho1 + A * r2);
      if (ext) unpack_ext(src, b);
      else pbd_unpack(src, b);
      z[r2] = from_u(b);
    }
    dft(z, 3, w16b);
    if (la > 0) {
      uint32_t u[9], b[9];
      to_u(z[0], u);
      reduce_centered(u, b);
      z[0] = from_u(b);
      for (int k2 = 1; k2 < 8; k2++) {
        uint32_t w[9];
        pbd_unpack(inner + 8 * ((rho1 * k2) << (8 - m)), w);
        z[brev(k2, 3)] = mul(z[brev(k2, 3)], w, false, true);
      }
    }
    for (int k2 = 0; k2 < 8; k2++) Y[rho1][k2] = z[brev(k2, 3)];
  }
  for (int k2 = 0; k2 < 8; k2++) {
    V yv[8];
    for (int r1 = 0; r1 < A; r1++) yv[r1] = Y[r1][k2];
    dft(yv, la, w16b);
    for (int k1 = 0; k1 < A; k1++) {
      const int k = k2 + 8 * k1;
      V v = yv[brev(k1, la)];
      uint32_t* o = out + 8 * k;
      if (last) {
        if (do_scale) {
          uint32_t w[9];
          pbd_unpack(scale, w);
          v = mul(v, w, false, false);
        }
        uint32_t u[9], t[9];
        to_u(v, u);
        reduce_below_2r(u, t);
        if ((int32_t)t[8] < 0 || t[8] >= (1u << 24)) g_bad++;
        join29(t, o);
        reduce_once(o);
      } else {
        uint32_t w[9];
        pbd_unpack(tw + 8 * k, w);
        v = mul(v, w, true, true);
        for (int i = 0; i < 8; i++)
          if (v.l[i] < -(1 << 28) || v.l[i] >= (1 << 28)) g_bad++;
        if (v.l[8] < -(1 << 23) || v.l[8] >= (1 << 23)) g_bad++;
        uint32_t u[9];
        to_u(v, u);
        pbd_pack(u, o);
      }
    }
  }
  (void)Mp;
  return g_bad;
}

// The permutation-ratio terms of one element (ops.hip perm_table_kernel and
// perm_terms_kernel) in the same arithmetic; returns the number of violated
// bounds.  b32: 32 S(beta) mod r; k29: three pbd words; w0, w1: the stored
// window values (w^i = w0 w1); the other inputs are stored-form words.
int sim_perm_terms(const uint32_t* w0, const uint32_t* w1, const uint32_t* b32, const uint32_t* gamma,
                   const uint32_t* k29, const uint32_t* w, const uint32_t* s, uint32_t* num_out,
                   uint32_t* den_out) {
  g_bad = 0;
  uint32_t B[9], b[9];
  unpack_ext(b32, B);
  unpack_ext(gamma, b);
  const V g = from_u(b);
  // the table entry of w0 per j: P(P(S(w0), beta R'), (32 k_j) R') as a pbd word
  unpack_ext(w0, b);
  const V bw = mul(from_u(b), B, false, false);
  uint32_t T[3][8];
  for (int j = 0; j < 3; j++) {
    uint32_t K[9], u[9];
    pbd_unpack(k29 + 8 * j, K);
    const V t = mul(bw, K, false, true);
    for (int i = 0; i < 8; i++)
      if (t.l[i] < -(1 << 28) || t.l[i] >= (1 << 28)) g_bad++;
    if (t.l[8] < -(1 << 23) || t.l[8] >= (1 << 23)) g_bad++;
    to_u(t, u);
    pbd_pack(u, T[j]);
  }
  unpack_ext(w1, b);
  const V x1 = from_u(b);
  V num, den;
  for (int j = 0; j < 3; j++) {
    unpack_ext(w + 8 * j, b);
    const V wg = add(from_u(b), g);
    uint32_t K[9];
    pbd_unpack(T[j], K);
    const V nj = add(wg, mul(x1, K, false, true));
    unpack_ext(s + 8 * j, b);
    const V dj = add(wg, mul(from_u(b), B, false, true));
    uint32_t u[9];
    if (j == 0) {
      num = nj;
      den = dj;
    } else {
      to_u(nj, u);
      num = mul(num, u, false, false);
      to_u(dj, u);
      den = mul(den, u, false, false);
    }
  }
  uint32_t u[9], t[9];
  to_u(num, u);
  reduce_below_2r(u, t);
  join29(t, num_out);
  reduce_once(num_out);
  to_u(den, u);
  reduce_below_2r(u, t);
  join29(t, den_out);
  reduce_once(den_out);
  return g_bad;
}

// fr29.hpp's host functions, one call each
void t_mont29(const uint32_t* a, const uint32_t* w, uint32_t* t, int bm, int bt) {
  if (bm && bt) mont29<true, true>(a, w, t);
  else if (bm) mont29<true, false>(a, w, t);
  else if (bt) mont29<false, true>(a, w, t);
  else mont29<false, false>(a, w, t);
}
void t_pbd_pack(const uint32_t* b, uint32_t* x) { pbd_pack(b, x); }
void t_pbd_unpack(const uint32_t* x, uint32_t* b) { pbd_unpack(x, b); }
void t_unpack_ext(const uint32_t* x, uint32_t* b) { unpack_ext(x, b); }
void t_to_pbd(const uint32_t* w, uint32_t* x) { to_pbd(w, x); }
void t_reduce_centered(const uint32_t* v, uint32_t* b) { reduce_centered(v, b); }
void t_reduce_below_2r(const uint32_t* v, uint32_t* t) { reduce_below_2r(v, t); }

}  // extern "C"
