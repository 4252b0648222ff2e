// Montgomery product in radix 2^29 for the NTT passes and the scans.
//
// The saturated 8 x 32-bit product (fr.hpp) spends one carry fold per
// multiply-add because a column of 32x32-bit products overflows 64 bits.  Here
// both operands are nine limbs of radix 2^29: every partial product is below
// 2^59 in magnitude and a column holds at most 17 of them plus a carry, so a
// column sums in one signed 64-bit accumulator with no carry handling: a chain
// of v_mad_i64_i32 whose carry-out masks are never read (153 multiply-adds and
// 17 column shifts per product, against 120 + 120 folds + about 50 for the
// saturated form).  `mont29` is the plain-C statement (host tests); the device
// runs the same columns as generated asm blocks (`mont29_dev`).
//
//   R' = 2^261 (nine digits), r = 1 mod 2^29: with the digit m = acc mod 2^29
//   (or its balanced form), subtracting m r clears the low 29 bits exactly.
//
// Two ways of holding the data:
//  * the scans keep elements in the stored form (x R mod r, R = 2^256, eight
//    32-bit limbs) and `fr_mul29` splits, multiplies, joins and adds r: a w / R'
//    mod r in (0, 1.031 r) for a < 2^256, w < r;
//  * the NTT passes keep elements as nine signed 32-bit limbs (`v9`) from load
//    to store: sums and differences are limb-wise, and the product takes any
//    signed limbs when the factor has balanced digits (|w_j| <= 2^28).  Factors
//    reach the passes as packed balanced digits ("pbd" words, below).
// A factor w enters as w R' mod r = (w R) 2^5 mod r, so mont29(x R, w R') =
// x w R: the data never changes form.
//
// Column bound: |a_i| < 2^31 with |w_j| <= 2^28, or a_i, w_j < 2^29: the
// signed column sum stays below 9 * 2^59 + 8 * 2^58 + 2^35 < 2^62.8.
#pragma once
#include <math.h>
#include <stdint.h>
#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#endif

namespace pl {

#if defined(__HIPCC__)
#define PL_HD29 __host__ __device__ __forceinline__
#else
#define PL_HD29 inline
#endif

struct fr9 {
  uint32_t l[9];
};

constexpr uint32_t kMask29 = (1u << 29) - 1;

// r in radix 2^29
PL_HD29 constexpr uint32_t r29(int i) {
  return i == 0 ? 0x1u : i == 1 ? 0x1ffffff8u : i == 2 ? 0x1f96ffbfu : i == 3 ? 0x1b4805ffu
       : i == 4 ? 0x1d80553bu : i == 5 ? 0x0c0404d0u : i == 6 ? 0x1520cce7u : i == 7 ? 0x0a6533afu
       : 0x0073eda7u;
}
// -r_i as a 32-bit two's complement (the signed multiply-add's operand)
PL_HD29 constexpr uint32_t nr29(int i) { return 0u - r29(i); }

// bits [s, s + 32) of the 64-bit value hi:lo
PL_HD29 uint32_t funnel_r(uint32_t hi, uint32_t lo, int s) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_alignbit(hi, lo, s);
#else
  return (uint32_t)((((uint64_t)hi << 32) | lo) >> s);
#endif
}

// eight 32-bit limbs (value < 2^256) -> nine 29-bit limbs
PL_HD29 void split29(const uint32_t* x, uint32_t* a) {
  a[0] = x[0] & kMask29;
  a[1] = funnel_r(x[1], x[0], 29) & kMask29;
  a[2] = funnel_r(x[2], x[1], 26) & kMask29;
  a[3] = funnel_r(x[3], x[2], 23) & kMask29;
  a[4] = funnel_r(x[4], x[3], 20) & kMask29;
  a[5] = funnel_r(x[5], x[4], 17) & kMask29;
  a[6] = funnel_r(x[6], x[5], 14) & kMask29;
  a[7] = funnel_r(x[7], x[6], 11) & kMask29;
  a[8] = x[7] >> 8;
}

// nine normalized 29-bit limbs (value < 2^256) -> eight 32-bit limbs
PL_HD29 void join29(const uint32_t* t, uint32_t* x) {
  x[0] = t[0] | (t[1] << 29);
  x[1] = (t[1] >> 3) | (t[2] << 26);
  x[2] = (t[2] >> 6) | (t[3] << 23);
  x[3] = (t[3] >> 9) | (t[4] << 20);
  x[4] = (t[4] >> 12) | (t[5] << 17);
  x[5] = (t[5] >> 15) | (t[6] << 14);
  x[6] = (t[6] >> 18) | (t[7] << 11);
  x[7] = (t[7] >> 21) | (t[8] << 8);
}

// the low 29 bits of v as a signed digit in [-2^28, 2^28)
PL_HD29 uint32_t sext29(uint32_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
  return (uint32_t)__builtin_amdgcn_sbfe((int)v, 0, 29);
#else
  return (uint32_t)((int32_t)(v << 3) >> 3);
#endif
}

// t = a w / 2^261 - m r / 2^261 (a, w in nine 29-bit-radix limbs taken as
// signed 32-bit integers): product scanning with a signed accumulator.
// Column k adds a_i w_{k-i} and subtracts m_i r_{k-i} (k - i >= 1); the digit
// m_k = acc mod 2^29 makes acc - m_k r_0 = acc - m_k divisible by 2^29, so the
// column closes with an arithmetic shift.  BM takes the digits m_k in
// [-2^28, 2^28) instead of [0, 2^29) (m in (-r'/2, r'/2) for r' = 2^261, so T
// is centred on a w / 2^261), BT the output limbs t_0..t_7 likewise; a
// balanced digit's column closes as (acc + 2^28) >> 29.  t_8 is the signed
// top.  T = (a w - m r) / 2^261 lies in
//   (a w / 2^261 - r, a w / 2^261]            (BM false)
//   (a w / 2^261 - r/2, a w / 2^261 + r/2)     (BM true).
// Column bound: with |a_i| < 2^31 and |w_j| <= 2^28, or |a_i| < 2^30 and
// w_j < 2^29, a column stays below 9 * 2^59 + 8 * 2^58 + 2^35 < 2^62.8.
template <bool BM = false, bool BT = false>
PL_HD29 void mont29(const uint32_t* a, const uint32_t* w, uint32_t* t) {
  uint32_t m[9];
  int64_t acc = 0;
#pragma unroll
  for (int k = 0; k < 17; k++) {
    const int lo = k > 8 ? k - 8 : 0, hi = k < 8 ? k : 8;
#pragma unroll
    for (int i = lo; i <= hi; i++) acc += (int64_t)(int32_t)a[i] * (int64_t)(int32_t)w[k - i];
#pragma unroll
    for (int i = lo; i <= hi; i++)
      if (i < k) acc -= (int64_t)(int32_t)m[i] * (int64_t)r29(k - i);
    const bool bal = k <= 8 ? BM : BT;
    const uint32_t dig = bal ? sext29((uint32_t)acc) : ((uint32_t)acc & kMask29);
    if (k <= 8)
      m[k] = dig;
    else
      t[k - 9] = dig;
    acc = (acc - (int64_t)(int32_t)dig) >> 29;  // exact
  }
  t[8] = (uint32_t)acc;
}

// x (eight 32-bit limbs, a value mod 2^256) += r
PL_HD29 void add_r256(uint32_t* x) {
  constexpr uint32_t kR[8] = {0x00000001u, 0xffffffffu, 0xfffe5bfeu, 0x53bda402u,
                              0x09a1d805u, 0x3339d808u, 0x299d7d48u, 0x73eda753u};
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) x[i] = __builtin_addc(x[i], kR[i], c, &c);
#else
  uint64_t c = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const uint64_t s = (uint64_t)x[i] + kR[i] + c;
    x[i] = (uint32_t)s;
    c = s >> 32;
  }
#endif
}

// ---- signed-digit elements inside an NTT pass ------------------------------
// Packed balanced digits ("pbd"): a 32-byte word whose bit fields
// [29 i, 29 i + 29), i < 8, hold digits b_i in [-2^28, 2^28) as 29-bit two's
// complement and whose bits [232, 256) hold b_8 in [-2^23, 2^23) as 24-bit two's
// complement; its value is sum b_i 2^(29 i).  Pass twiddle tables hold their
// factors this way, and the pass kernels hand intermediate results from pass
// to pass in it.

// the nine signed limbs of a pbd word
PL_HD29 void pbd_unpack(const uint32_t* x, uint32_t* b) {
  b[0] = sext29(x[0]);
  b[1] = sext29(funnel_r(x[1], x[0], 29));
  b[2] = sext29(funnel_r(x[2], x[1], 26));
  b[3] = sext29(funnel_r(x[3], x[2], 23));
  b[4] = sext29(funnel_r(x[4], x[3], 20));
  b[5] = sext29(funnel_r(x[5], x[4], 17));
  b[6] = sext29(funnel_r(x[6], x[5], 14));
  b[7] = sext29(funnel_r(x[7], x[6], 11));
  b[8] = (uint32_t)((int32_t)x[7] >> 8);
}

// the pbd word of digits b_0..b_7 in [-2^28, 2^28), b_8 in [-2^23, 2^23):
// word j takes the two fields it overlaps from the 64-bit window b_{i+1} :
// (b_i << 3), where b_i's field holds bit 32 j
PL_HD29 void pbd_pack(const uint32_t* b, uint32_t* x) {
  x[0] = funnel_r(b[1], b[0] << 3, 3);
  x[1] = funnel_r(b[2], b[1] << 3, 6);
  x[2] = funnel_r(b[3], b[2] << 3, 9);
  x[3] = funnel_r(b[4], b[3] << 3, 12);
  x[4] = funnel_r(b[5], b[4] << 3, 15);
  x[5] = funnel_r(b[6], b[5] << 3, 18);
  x[6] = funnel_r(b[7], b[6] << 3, 21);
  x[7] = funnel_r(b[8], b[7] << 3, 24);
}

// B' = sum_{i<8} 2^28 2^(29 i) as eight 32-bit words
PL_HD29 constexpr uint32_t kBp(int i) {
  return i == 0 ? 0x10000000u : i == 1 ? 0x02000000u : i == 2 ? 0x00400000u : i == 3 ? 0x00080000u
       : i == 4 ? 0x00010000u : i == 5 ? 0x00002000u : i == 6 ? 0x00000400u : 0x00000080u;
}

// balanced digits of any 256-bit value x (eight 32-bit limbs): with
// y = x + B', digit i < 8 is y_i - 2^28 = the 29-bit field of y XOR B' read as
// signed, and the top digit is y_8 (with y's carry out as its bit 24, < 2^25)
PL_HD29 void unpack_ext(const uint32_t* x, uint32_t* b) {
  uint32_t y[8];
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) y[i] = __builtin_addc(x[i], kBp(i), c, &c);
#else
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const uint64_t s = (uint64_t)x[i] + kBp(i) + c;
    y[i] = (uint32_t)s;
    c = (uint32_t)(s >> 32);
  }
#endif
  uint32_t z[8];
#pragma unroll
  for (int i = 0; i < 8; i++) z[i] = y[i] ^ kBp(i);
  pbd_unpack(z, b);
  b[8] = (y[7] >> 8) | (c << 24);
}

// the pbd word of a value w < 2^255 (a table factor)
PL_HD29 void to_pbd(const uint32_t* w, uint32_t* x) {
  uint32_t b[9];
  unpack_ext(w, b);
  pbd_pack(b, x);
}

// 2^232 / r, the weight of the top limb in units of r
constexpr float kTopPerR = 1.31622602e-07f;

// Balanced digits of a value congruent to v (nine signed limbs, |v_i| < 2^31)
// and within (-0.51 r, 0.51 r): v - q r with q = round(v / r) estimated from
// the top limb.  v / r = v_8 2^232 / r + e with |e| < 2^234 / r < 2^-20, and
// the float estimate errs by < 2^-14, so |v / r - q| < 0.5001.
PL_HD29 void reduce_centered(const uint32_t* v, uint32_t* b) {
  const int32_t q = (int32_t)rintf((float)(int32_t)v[8] * kTopPerR);
  int64_t acc = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    acc += (int64_t)(int32_t)v[i] - (int64_t)q * (int64_t)r29(i);
    b[i] = sext29((uint32_t)acc);
    acc = (acc - (int64_t)(int32_t)b[i]) >> 29;
  }
  b[8] = (uint32_t)(acc + (int64_t)(int32_t)v[8] - (int64_t)q * (int64_t)r29(8));
}

// The digits in [0, 2^29) (top t_8 >= 0) of a value in [0, 1.01 r) congruent
// to v (nine signed limbs, |v_i| < 2^31): v - q r with q = floor(v / r - 2^-8)
// estimated from the top limb (so 0 < v / r - q < 1.01, by the bounds above).
PL_HD29 void reduce_below_2r(const uint32_t* v, uint32_t* t) {
  const int32_t q = (int32_t)floorf((float)(int32_t)v[8] * kTopPerR - 0.00390625f);
  int64_t acc = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    acc += (int64_t)(int32_t)v[i] - (int64_t)q * (int64_t)r29(i);
    t[i] = (uint32_t)acc & kMask29;
    acc >>= 29;
  }
  t[8] = (uint32_t)(acc + (int64_t)(int32_t)v[8] - (int64_t)q * (int64_t)r29(8));
}

#undef PL_HD29

}  // namespace pl

#if defined(__HIPCC__)
#include "fr.hpp"

namespace pl {

// ---- device product: one asm block per column -----------------------------
#include "mont29_columns.inc"

// mont29 on the device; UB: w is wave-uniform (SGPRs); BM, BT as in mont29
template <bool UB, bool BM = false, bool BT = false>
__device__ __forceinline__ void mont29_dev(const uint32_t* x, const uint32_t* w, uint32_t* t) {
  uint32_t m[9];
  uint64_t acc[18];
  static_for<0, 17>([&](auto k_) {
    constexpr int k = decltype(k_)::value;
    // block k closes column k - 1, whose digit was balanced if RND
    constexpr bool RND = k >= 1 && (k - 1 <= 8 ? BM : BT);
    Mont29Col<k, UB, RND>::run(acc[k], k ? acc[k ? k - 1 : 0] : 0, x, w, m);
    constexpr bool BAL = k <= 8 ? BM : BT;
    const uint32_t lo = (uint32_t)acc[k];
    const uint32_t dig = BAL ? sext29(lo) : (lo & kMask29);
    if constexpr (k <= 8)
      m[k] = dig;
    else
      t[k - 9] = dig;
  });
  Mont29Col<17, UB, BT>::run(acc[17], acc[16], x, w, m);
  t[8] = (uint32_t)acc[17];
}

// a w / R' mod r for a < 2^256 (stored form, lazily reduced) and w given as
// nine 29-bit limbs of w R' mod r (w < r); result in [0, 1.03 r).
template <bool UB = false>
__device__ __forceinline__ fr fr_mul29(const fr& a, const uint32_t (&w)[9]) {
  uint32_t x[9], t[9];
  split29(a.l, x);
  mont29_dev<UB, false, false>(x, w, t);
  fr o;
  join29(t, o.l);
  add_r256(o.l);
  return o;
}

// the same with w read as an 8-limb table entry holding w R' mod r
__device__ __forceinline__ fr fr_mul29(const fr& a, const fr& w8) {
  uint32_t w[9];
  split29(w8.l, w);
  return fr_mul29<false>(a, w);
}

// ---- signed-limb elements inside an NTT pass ------------------------------
// value sum (int32)l_i 2^(29 i), a representative of its class mod r.  Sums
// and differences are limb-wise (no carries); a limb may use all 32 bits
// signed, and the product (balanced factor digits, |w_j| <= 2^28) takes any
// such limbs.
struct v9 {
  uint32_t l[9];
};

__device__ __forceinline__ v9 v9_add(const v9& a, const v9& b) {
  v9 s;
#pragma unroll
  for (int i = 0; i < 9; i++) s.l[i] = a.l[i] + b.l[i];
  return s;
}
__device__ __forceinline__ v9 v9_sub(const v9& a, const v9& b) {
  v9 s;
#pragma unroll
  for (int i = 0; i < 9; i++) s.l[i] = a.l[i] - b.l[i];
  return s;
}

// x w / 2^261 (mont29_dev) for a factor w in balanced digits
template <bool UB, bool BM, bool BT>
__device__ __forceinline__ v9 v9_mul(const v9& x, const uint32_t (&w)[9]) {
  v9 t;
  mont29_dev<UB, BM, BT>(x.l, w, t.l);
  return t;
}

// the canonical value in [0, r) of v (|v_i| < 2^31)
__device__ __forceinline__ fr v9_canonical(const v9& v) {
  uint32_t t[9];
  reduce_below_2r(v.l, t);
  fr o;
  join29(t, o.l);
  return fr_reduce_once(o);
}

// an ordering point for the nine limbs of v (no instruction)
__device__ __forceinline__ void pin9(v9& v) {
  asm volatile("" : "+v"(v.l[0]), "+v"(v.l[1]), "+v"(v.l[2]), "+v"(v.l[3]), "+v"(v.l[4]), "+v"(v.l[5]),
               "+v"(v.l[6]), "+v"(v.l[7]), "+v"(v.l[8]));
}

}  // namespace pl
#endif
