// Inverse of one field element by the optimized binary GCD (T. Pornin,
// "Optimized Binary GCD for Modular Inversion", IACR ePrint 2020/972,
// Algorithm 2, with k - 1 = 31).
//
// a = y, b = r, u = 1, v = 0 keep a = u y and b = v y (mod r).  Each of the
// 17 outer steps runs 31 binary-GCD steps on 64-bit approximations of a and
// b (their low 31 bits and top 33 bits), recording the update as a 2x2
// integer matrix (f0 g0; f1 g1) with |f| + |g| <= 2^31, then applies it to the
// full values: a <- (f0 a + g0 b) / 2^31, b <- (f1 a + g1 b) / 2^31 (exact;
// negated with its row when negative), and u, v by the same combination
// divided by 2^31 modulo r (a Montgomery step: r = 1 mod 2^32).  After
// 17 * 31 >= 2 * 255 - 1 steps b = gcd = 1 and v = y^-1 mod r.  The rounds
// stop as soon as a = 0: b (odd throughout) is then gcd(y, r) = 1, and a
// further round only halves a = 0 and doubles row 1, so b' = b and
// v' = 2^31 v / 2^31 = v; random inputs converge in about 12 or 13 rounds.
//
// Plain integer code for host and device (tested on the host against
// Fermat's little theorem).  `inverse` runs on one thread; `inverse_wave`
// runs the four updates of each round on four lanes of a wave at once.
#pragma once
#include <stdint.h>
#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#endif

namespace pl {
namespace inv {

#if defined(__HIPCC__)
#define PL_HD __host__ __device__ __forceinline__
#else
#define PL_HD inline
#endif

constexpr uint32_t kR[8] = {0x00000001u, 0xffffffffu, 0xfffe5bfeu, 0x53bda402u,
                            0x09a1d805u, 0x3339d808u, 0x299d7d48u, 0x73eda753u};

// 10-limb two's-complement accumulator (320 bits)
struct S10 {
  uint32_t w[10];
};

PL_HD int bitlen8(const uint32_t* x) {
  for (int i = 7; i >= 0; i--)
    if (x[i]) return 32 * i + 32 - __builtin_clz(x[i]);
  return 0;
}

// bits [s, s + 33) of x (s >= 31), as an integer
PL_HD uint64_t bits33(const uint32_t* x, int s) {
  const int w = s >> 5, sh = s & 31;
  // limbs w, w + 1, w + 2 (0 past the top) by selects, not an index, so that
  // on a GPU the limbs stay in (scalar) registers
  uint64_t x0 = 0, x1 = 0, x2 = 0;
  for (int i = 0; i < 8; i++) {
    x0 = i == w ? x[i] : x0;
    x1 = i == w + 1 ? x[i] : x1;
    x2 = i == w + 2 ? x[i] : x2;
  }
  uint64_t lo = x0 | (x1 << 32);
  uint64_t v = sh ? ((lo >> sh) | (x2 << (64 - sh))) : lo;
  return v & ((1ull << 33) - 1);
}

// p = x * a as 10 limbs (x: 8 limbs unsigned, a <= 2^31; limb 9 zero)
PL_HD void mul_u32(const uint32_t* x, uint64_t a, uint32_t* p) {
  uint64_t c = 0;
  for (int i = 0; i < 8; i++) {
    const uint64_t t = (uint64_t)x[i] * a + c;  // a < 2^32: no overflow
    p[i] = (uint32_t)t;
    c = t >> 32;
  }
  p[8] = (uint32_t)c;
  p[9] = 0;
}

PL_HD bool s10_neg(const S10& a) { return (a.w[9] >> 31) != 0; }

// x (nine limbs, x < 5 r) <- x mod r.  qe = floor(X / (r_hi + 1)) from the
// top 64 bits X = x >> 224 (r_hi = r >> 224) is floor(x / r) or one less:
// X / (r_hi + 1) < x / r < (X + 1) / r_hi, and the two ends differ by
// (X + r_hi + 1) / (r_hi (r_hi + 1)) < 6 / r_hi.  Subtract qe r, then r once
// more if x is still >= r.
PL_HD void reduce_below_5r(uint32_t* x) {
  const uint64_t X = ((uint64_t)x[8] << 32) | x[7], rh1 = (uint64_t)kR[7] + 1;
  const uint32_t qe = (uint32_t)(X >= rh1) + (uint32_t)(X >= 2 * rh1) + (uint32_t)(X >= 3 * rh1) +
                      (uint32_t)(X >= 4 * rh1);
  {
    int64_t bw = 0;
    uint64_t c = 0;
    for (int i = 0; i < 9; i++) {
      const uint64_t qr = (i < 8 ? (uint64_t)qe * kR[i] : 0) + c;  // qe r, limb i, with its carry
      c = qr >> 32;
      int64_t s = (int64_t)x[i] - (int64_t)(uint32_t)qr + bw;
      x[i] = (uint32_t)s;
      bw = s >> 32;  // 0 or -1
    }
  }
  uint32_t d[9];
  int64_t bw = 0;
  for (int i = 0; i < 9; i++) {
    int64_t s = (int64_t)x[i] - (i < 8 ? kR[i] : 0) + bw;
    d[i] = (uint32_t)s;
    bw = s >> 32;  // 0 or -1
  }
  if (bw == 0)
    for (int i = 0; i < 9; i++) x[i] = d[i];
}

// x <- -x mod r for x in [0, r)
PL_HD void negate_mod_r(uint32_t* x) {
  uint32_t nz = 0;
  for (int i = 0; i < 8; i++) nz |= x[i];
  int64_t bw = 0;
  for (int i = 0; i < 8; i++) {
    int64_t s = (int64_t)(nz ? kR[i] : 0) - x[i] + bw;
    x[i] = (uint32_t)s;
    bw = s >> 32;  // 0 or -1
  }
}

// the 31 binary-GCD steps of one outer round on the 64-bit approximations of
// a and b; m = (f0, g0, f1, g1)
PL_HD void inner_steps(const uint32_t* a, const uint32_t* b, int64_t* m) {
  uint32_t ab[8];
  for (int i = 0; i < 8; i++) ab[i] = a[i] | b[i];
  int n = bitlen8(ab);
  if (n < 64) n = 64;
  uint64_t xa = (uint64_t)(a[0] & 0x7fffffffu) | (bits33(a, n - 33) << 31);
  uint64_t xb = (uint64_t)(b[0] & 0x7fffffffu) | (bits33(b, n - 33) << 31);
  // One step: if a is odd, swap the rows when a < b, then row 0 -= row 1;
  // a halves, row 1 doubles.  The rows (f0, g0) and (f1, g1) run packed as
  // f + 2^32 g (mod 2^64): every step is linear in the rows, so the packed
  // words follow them exactly, and after j steps |f0|, |g0| <= 2^j - 1 and
  // |f1|, |g1| <= 2^j, so after 30 steps every factor is the sign-extended
  // half it sits in.  The 31st step, whose row 1 can reach 2^31, runs on the
  // unpacked factors.  a < b is the borrow of a - b from bit 63: logic on
  // scalar registers on a GPU, where a 64-bit compare would take a vector one.
  uint64_t p0 = 1, p1 = (uint64_t)1 << 32;
  for (int j = 0; j < 30; j++) {
    if (xa & 1) {
      const uint64_t d = xa - xb;
      // -1 when xa < xb, else 0, applied as a mask (a branch on it would
      // become a 64-bit compare)
      const uint64_t lt = (uint64_t)((int64_t)((~xa & xb) | (~(xa ^ xb) & d)) >> 63);
      xb ^= lt & (xa ^ xb);  // b <- a on a swap
      xa = (d ^ lt) - lt;    // a - b after the swap: -d on a swap, else d
      const uint64_t tp = lt & (p0 ^ p1);
      p0 ^= tp;
      p1 ^= tp;
      p0 -= p1;
    }
    xa >>= 1;
    p1 <<= 1;
  }
  int64_t f0 = (int32_t)(uint32_t)p0, f1 = (int32_t)(uint32_t)p1;
  int64_t g0 = (int64_t)(p0 - (uint64_t)f0) >> 32, g1 = (int64_t)(p1 - (uint64_t)f1) >> 32;
  if (xa & 1) {
    if (xa < xb) {
      int64_t tf = f0; f0 = f1; f1 = tf;
      int64_t tg = g0; g0 = g1; g1 = tg;
    }
    f0 -= f1;
    g0 -= g1;
  }
  f1 *= 2;
  g1 *= 2;
  m[0] = f0;
  m[1] = g0;
  m[2] = f1;
  m[3] = g1;
}

// One of the four full-width updates of an outer round, by role:
//   0: |f0 a + g0 b| / 2^31 -> res, *neg = its sign was negative
//   1: |f1 a + g1 b| / 2^31 -> res, *neg likewise
//   2: (f0 u + g0 v) / 2^31 mod r -> res
//   3: (f1 u + g1 v) / 2^31 mod r -> res
// Role 2 (3) then takes role 0's (1's) sign: res <- -res mod r if that row
// was negated (fix by negate_mod_r).  The four are independent, so the wave
// form runs them on four lanes at once.
// All four roles run one code path (in the wave form, no divergence but the
// final reduction), in two 10-limb carry passes: t = f x + g y with (x, y) =
// (a, b) or (u, v), as (x |f| ^ m_f) + (y |g| ^ m_g) + [f < 0] + [g < 0]
// (two's complement); then, when t < 0, the integer rows negate it
// (~t + 1) and the modular rows add 2^32 r, in the same pass as the
// modular rows' q r, q = -t mod 2^31 (a Montgomery step; adding 2^32 r
// leaves the low word, and q, as they are; q = 0 for the integer rows);
// both shift right by 31, and the modular rows reduce below r.
PL_HD void round_role(int role, const uint32_t* a, const uint32_t* b, const uint32_t* u, const uint32_t* v,
                      const int64_t* m, uint32_t* res, bool* neg) {
  const bool modr = role >= 2;
  // selects, not an index: the role differs per lane in the wave form
  const int64_t f = (role & 1) ? m[2] : m[0], g = (role & 1) ? m[3] : m[1];
  uint32_t x[8], y[8];
  for (int i = 0; i < 8; i++) {
    x[i] = modr ? u[i] : a[i];
    y[i] = modr ? v[i] : b[i];
  }
  const bool nf = f < 0, ng = g < 0;
  uint32_t px[10], py[10];
  mul_u32(x, nf ? (uint64_t)(-f) : (uint64_t)f, px);
  mul_u32(y, ng ? (uint64_t)(-g) : (uint64_t)g, py);
  S10 t;
  {
    const uint32_t mf = nf ? 0xffffffffu : 0u, mg = ng ? 0xffffffffu : 0u;
    uint64_t c = (uint64_t)nf + (uint64_t)ng;
    for (int i = 0; i < 10; i++) {
      const uint64_t s = (uint64_t)(px[i] ^ mf) + (py[i] ^ mg) + c;
      t.w[i] = (uint32_t)s;
      c = s >> 32;
    }
  }
  const bool tneg = s10_neg(t);
  *neg = !modr && tneg;
  {
    const uint32_t flip = *neg ? 0xffffffffu : 0u;
    const bool addr = modr && tneg;  // |t| < 2^32 r: t + 2^32 r >= 0
    // q = -t mod 2^31 makes t + q r divisible by 2^31 (r^-1 = 1 mod 2^31)
    const uint32_t q = modr ? (0u - t.w[0]) & 0x7fffffffu : 0u;
    uint64_t c = *neg ? 1 : 0;
    for (int i = 0; i < 10; i++) {
      // below 2^32 + 2^63 + 2^32 + 2^32: no overflow
      const uint64_t rr = (addr && i >= 1 && i <= 8) ? kR[i - 1] : 0;
      const uint64_t s = (uint64_t)(t.w[i] ^ flip) + (i < 8 ? (uint64_t)q * kR[i] : 0) + rr + c;
      t.w[i] = (uint32_t)s;
      c = s >> 32;
    }
  }
  uint32_t z[9];
  for (int i = 0; i < 8; i++) z[i] = (t.w[i] >> 31) | (t.w[i + 1] << 1);
  z[8] = (t.w[8] >> 31) | (t.w[9] << 1);
  if (modr) reduce_below_5r(z);  // z < 5 r
  for (int i = 0; i < 8; i++) res[i] = z[i];
}

// u' = (u f + v g) / 2^31 mod r, with u, v < r: the modular row of a round
PL_HD void mod_update(uint32_t* out, const uint32_t* u, int64_t f, const uint32_t* v, int64_t g) {
  const int64_t m[4] = {f, g, 0, 0};
  bool neg;
  round_role(2, u, v, u, v, m, out, &neg);
}

// y^-1 mod r for 0 < y < r (y as 8 little-endian limbs); 0 for y = 0
PL_HD void inverse(const uint32_t* y, uint32_t* out) {
  uint32_t a[8], b[8], u[8], v[8];
  bool zero = true;
  for (int i = 0; i < 8; i++) {
    a[i] = y[i];
    b[i] = kR[i];
    u[i] = 0;
    v[i] = 0;
    zero = zero && y[i] == 0;
  }
  if (zero) {
    for (int i = 0; i < 8; i++) out[i] = 0;
    return;
  }
  u[0] = 1;
  for (int outer = 0; outer < 17; outer++) {
    int64_t m[4];
    inner_steps(a, b, m);
    uint32_t res[4][8];
    bool neg[4];
    for (int role = 0; role < 4; role++) round_role(role, a, b, u, v, m, res[role], &neg[role]);
    if (neg[0]) negate_mod_r(res[2]);
    if (neg[1]) negate_mod_r(res[3]);
    uint32_t az = 0;
    for (int i = 0; i < 8; i++) {
      a[i] = res[0][i];
      b[i] = res[1][i];
      u[i] = res[2][i];
      v[i] = res[3][i];
      az |= a[i];
    }
    if (!az) break;  // converged: later rounds leave b = 1 and v as they are
  }
  for (int i = 0; i < 8; i++) out[i] = v[i];
}

#if defined(__HIPCC__)
// The same inverse by one wave: every lane runs the (uniform) inner steps,
// lane l runs role l & 3 of each round's update, and the four results are
// shared by readlane.  Call from all lanes of a wave, converged; every lane
// returns y^-1.
__device__ __forceinline__ void inverse_wave(const uint32_t* y, uint32_t* out) {
  const int role = (int)(__lane_id() & 3);
  uint32_t a[8], b[8], u[8], v[8];
  uint32_t nz = 0;
  for (int i = 0; i < 8; i++) {
    a[i] = (uint32_t)__builtin_amdgcn_readfirstlane((int)y[i]);
    b[i] = kR[i];
    u[i] = 0;
    v[i] = 0;
    nz |= a[i];
  }
  if (!nz) {
    for (int i = 0; i < 8; i++) out[i] = 0;
    return;
  }
  u[0] = 1;
  for (int outer = 0; outer < 17; outer++) {
    int64_t m[4];
    inner_steps(a, b, m);
    uint32_t res[8];
    bool neg;
    round_role(role, a, b, u, v, m, res, &neg);
    const int n0 = __builtin_amdgcn_readlane((int)neg, 0), n1 = __builtin_amdgcn_readlane((int)neg, 1);
    if ((role == 2 && n0) || (role == 3 && n1)) negate_mod_r(res);
    uint32_t az = 0;
    for (int i = 0; i < 8; i++) {
      a[i] = (uint32_t)__builtin_amdgcn_readlane((int)res[i], 0);
      b[i] = (uint32_t)__builtin_amdgcn_readlane((int)res[i], 1);
      u[i] = (uint32_t)__builtin_amdgcn_readlane((int)res[i], 2);
      v[i] = (uint32_t)__builtin_amdgcn_readlane((int)res[i], 3);
      az |= a[i];
    }
    if (!az) break;  // converged (a uniform branch: a is the same on every lane)
  }
  for (int i = 0; i < 8; i++) out[i] = v[i];
}
#endif

#undef PL_HD

}  // namespace inv
}  // namespace pl
