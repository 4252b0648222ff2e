// Montgomery product in radix 2^29 for the NTT passes.
//
// The saturated 8 x 32-bit product (fr.hpp) spends one carry fold per
// multiply-add because a column of 32x32-bit products overflows 64 bits.  Here
// both operands are split into nine 29-bit limbs: every partial product is
// below 2^58 and a column holds at most 17 of them plus a carry below 2^35, so
// a column sums in one 64-bit accumulator with no carry handling: a chain of
// v_mad_i64_i32 whose carry-out masks are never read (153 multiply-adds and
// about 70 other instructions per product, against 120 + 120 folds + about
// 50 for the saturated form).  `mont29` is the plain-C statement (host tests);
// the device runs the same columns as generated asm blocks (`mont29_dev`).
//
//   R' = 2^261 (nine digits), r = 1 mod 2^29: with the signed digit
//   m = acc mod 2^29, subtracting m r clears the low 29 bits exactly.
//   a * w / R' mod r, output in (0, 1.031 r) for a < 2^256, w < r.
//
// Data stays in the stored form (x R mod r, R = 2^256, eight 32-bit limbs,
// lazily in [0, 2r)); a factor w enters as w R' mod r = (w R) 2^5 mod r in
// nine 29-bit limbs (`fr9`), so mont29(x R, w R') = x w R.  Tables and
// constants for the passes are kept in that form.
//
// Column bound: a_i, w_j, m_i, r_j < 2^29 (a < 2^256 gives a_8 < 2^24): the
// signed column sum stays within 9 * 2^58 + 2^35 < 2^62 in magnitude.  Value:
// T = (a w - m r) / R' lies in (-r, a r / R') = (-r, 0.031 r), and the
// product returns T + r.
#pragma once
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

// t = a w / 2^261 - r (a, w in nine 29-bit limbs; a w / 2^261 taken mod r as
// below): product scanning with a signed digit.  Column k adds a_i w_{k-i}
// and subtracts m_i r_{k-i} (k - i >= 1) in a signed 64-bit accumulator; the
// digit m_k = acc mod 2^29 makes acc - m_k r_0 = acc - m_k divisible by 2^29,
// so the column closes with an arithmetic shift.  The value
// T = (a w - m r) / 2^261 lies in (-r, 0.031 r) for a < 2^256, w < r; t_0..t_7
// are its normalized limbs and t_8 its signed top (two's complement in 32
// bits).  add_r256 then gives T + r in (0, 1.031 r).
PL_HD29 void mont29(const uint32_t* a, const uint32_t* w, uint32_t* t) {
  uint32_t m[9];
  int64_t acc = 0;
#pragma unroll
  for (int k = 0; k < 17; k++) {
    const int lo = k > 8 ? k - 8 : 0, hi = k < 8 ? k : 8;
#pragma unroll
    for (int i = lo; i <= hi; i++) acc += (int64_t)a[i] * (int64_t)w[k - i];
#pragma unroll
    for (int i = lo; i <= hi; i++)
      if (i < k) acc -= (int64_t)m[i] * (int64_t)r29(k - i);
    if (k <= 8)
      m[k] = (uint32_t)acc & kMask29;
    else
      t[k - 9] = (uint32_t)acc & kMask29;
    acc >>= 29;  // arithmetic: floor division
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

#undef PL_HD29

}  // namespace pl

#if defined(__HIPCC__)
#include "fr.hpp"

namespace pl {

// ---- lazy arithmetic on stored-form elements in [0, 2r) -------------------
// 2r < 2^256, so a value below 2r still fits eight limbs; sums below 4r carry
// into a 257th bit that the trial subtraction accounts for.
__host__ __device__ constexpr uint32_t r2_limb(int i) {
  return i == 0 ? 0x00000002u : i == 1 ? 0xfffffffeu : i == 2 ? 0xfffcb7fdu : i == 3 ? 0xa77b4805u
       : i == 4 ? 0x1343b00au : i == 5 ? 0x6673b010u : i == 6 ? 0x533afa90u : 0xe7db4ea6u;
}

// a + b, reduced below 2r (a, b < 2r)
__device__ __forceinline__ fr fr_add_lazy(const fr& a, const fr& b) {
  fr s, d;
  uint32_t c = 0, bo = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) s.l[i] = __builtin_addc(a.l[i], b.l[i], c, &c);
#pragma unroll
  for (int i = 0; i < 8; i++) d.l[i] = __builtin_subc(s.l[i], r2_limb(i), bo, &bo);
  const bool ge = (c != 0) | (bo == 0);  // s >= 2r
  fr o;
#pragma unroll
  for (int i = 0; i < 8; i++) o.l[i] = ge ? d.l[i] : s.l[i];
  return o;
}

// a - b, in [0, 2r) (a, b < 2r)
__device__ __forceinline__ fr fr_sub_lazy(const fr& a, const fr& b) {
  fr d;
  uint32_t bo = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) d.l[i] = __builtin_subc(a.l[i], b.l[i], bo, &bo);
  const uint32_t mask = 0u - bo;
  fr o;
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) o.l[i] = __builtin_addc(d.l[i], r2_limb(i) & mask, c, &c);
  return o;
}

// ---- device product: one asm block per column -----------------------------
#include "mont29_columns.inc"

// mont29 on the device; UB: w is wave-uniform (SGPRs)
template <bool UB>
__device__ __forceinline__ void mont29_dev(const uint32_t* x, const uint32_t* w, uint32_t* t) {
  uint32_t m[9];
  uint64_t acc[18];
  static_for<0, 17>([&](auto k_) {
    constexpr int k = decltype(k_)::value;
    Mont29Col<k, UB>::run(acc[k], k ? acc[k ? k - 1 : 0] : 0, x, w, m);
    if constexpr (k <= 8)
      m[k] = (uint32_t)acc[k] & kMask29;
    else
      t[k - 9] = (uint32_t)acc[k] & kMask29;
  });
  Mont29Col<17, UB>::run(acc[17], acc[16], x, w, m);
  t[8] = (uint32_t)acc[17];
}

// a w / R' mod r for a < 2^256 (stored form, lazily reduced) and w given as
// nine 29-bit limbs of w R' mod r (w < r); result in [0, 1.03 r).
template <bool UB = false>
__device__ __forceinline__ fr fr_mul29(const fr& a, const uint32_t (&w)[9]) {
  uint32_t x[9], t[9];
  split29(a.l, x);
  mont29_dev<UB>(x, w, t);
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

}  // namespace pl
#endif
