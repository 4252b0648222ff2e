// Host-side Fr arithmetic for building constants (roots of unity, coset
// shifts, 2^-k) at domain-creation time.  4 x 64-bit limbs with
// unsigned __int128; the bytes are identical to the device's 8 x 32-bit
// limbs (little-endian).  Not on any hot path.
#pragma once
#include <stdint.h>
#include <string.h>

namespace pl {
namespace host {

struct fr64 {
  uint64_t l[4];
};

static const uint64_t kP[4] = {0xffffffff00000001ull, 0x53bda402fffe5bfeull,
                               0x3339d80809a1d805ull, 0x73eda753299d7d48ull};
static const uint64_t kN0 = 0xfffffffeffffffffull;  // -r^-1 mod 2^64
static const uint64_t kRR[4] = {0xc999e990f3f29c6dull, 0x2b6cedcb87925c23ull,
                                0x05d314967254398full, 0x0748d9d99f59ff11ull};
static const uint64_t kOne[4] = {0x00000001fffffffeull, 0x5884b7fa00034802ull,
                                 0x998c4fefecbc4ff5ull, 0x1824b159acc5056full};

inline bool geq_p(const uint64_t* a) {
  for (int i = 3; i >= 0; i--) {
    if (a[i] != kP[i]) return a[i] > kP[i];
  }
  return true;
}

inline void sub_p(uint64_t* a) {
  unsigned __int128 br = 0;
  for (int i = 0; i < 4; i++) {
    unsigned __int128 d = (unsigned __int128)a[i] - kP[i] - (uint64_t)br;
    a[i] = (uint64_t)d;
    br = (d >> 64) & 1;
  }
}

inline fr64 mul(const fr64& a, const fr64& b) {
  uint64_t t[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < 4; i++) {
    unsigned __int128 c = 0;
    for (int j = 0; j < 4; j++) {
      c = (unsigned __int128)a.l[j] * b.l[i] + t[j] + (uint64_t)c;
      t[j] = (uint64_t)c;
      c >>= 64;
    }
    unsigned __int128 s = (unsigned __int128)t[4] + (uint64_t)c;
    t[4] = (uint64_t)s;
    t[5] = (uint64_t)(s >> 64);
    uint64_t m = t[0] * kN0;
    c = ((unsigned __int128)m * kP[0] + t[0]) >> 64;
    for (int j = 1; j < 4; j++) {
      c = (unsigned __int128)m * kP[j] + t[j] + (uint64_t)c;
      t[j - 1] = (uint64_t)c;
      c >>= 64;
    }
    s = (unsigned __int128)t[4] + (uint64_t)c;
    t[3] = (uint64_t)s;
    t[4] = t[5] + (uint64_t)(s >> 64);
  }
  fr64 r;
  memcpy(r.l, t, 32);
  if (t[4] || geq_p(r.l)) sub_p(r.l);
  return r;
}

inline fr64 add(const fr64& a, const fr64& b) {
  fr64 r;
  unsigned __int128 c = 0;
  for (int i = 0; i < 4; i++) {
    c = (unsigned __int128)a.l[i] + b.l[i] + (uint64_t)c;
    r.l[i] = (uint64_t)c;
    c >>= 64;
  }
  if (c || geq_p(r.l)) sub_p(r.l);
  return r;
}

inline fr64 sub(const fr64& a, const fr64& b) {
  fr64 r;
  unsigned __int128 br = 0;
  for (int i = 0; i < 4; i++) {
    unsigned __int128 d = (unsigned __int128)a.l[i] - b.l[i] - (uint64_t)br;
    r.l[i] = (uint64_t)d;
    br = (d >> 64) & 1;
  }
  if (br) {
    unsigned __int128 c = 0;
    for (int i = 0; i < 4; i++) {
      c = (unsigned __int128)r.l[i] + kP[i] + (uint64_t)c;
      r.l[i] = (uint64_t)c;
      c >>= 64;
    }
  }
  return r;
}

inline bool is_zero(const fr64& a) { return (a.l[0] | a.l[1] | a.l[2] | a.l[3]) == 0; }

inline fr64 one() {
  fr64 r;
  memcpy(r.l, kOne, 32);
  return r;
}

inline fr64 from_u64(uint64_t v) {
  fr64 a = {{v, 0, 0, 0}}, rr;
  memcpy(rr.l, kRR, 32);
  return mul(a, rr);
}

// a^e, e a 256-bit little-endian exponent
inline fr64 pow(const fr64& a, const uint64_t* e) {
  fr64 acc = one(), base = a;
  for (int i = 0; i < 256; i++) {
    if ((e[i >> 6] >> (i & 63)) & 1) acc = mul(acc, base);
    base = mul(base, base);
  }
  return acc;
}

inline fr64 pow_u64(const fr64& a, uint64_t e) {
  uint64_t ee[4] = {e, 0, 0, 0};
  return pow(a, ee);
}

inline fr64 inv(const fr64& a) {
  uint64_t e[4] = {kP[0] - 2, kP[1], kP[2], kP[3]};
  return pow(a, e);
}

// primitive 2^lg-th root of unity: 7^((r-1)/2^32) squared (32-lg) times
inline fr64 root_of_unity(int lg) {
  uint64_t e[4];
  uint64_t rm1[4] = {kP[0] - 1, kP[1], kP[2], kP[3]};
  for (int i = 0; i < 4; i++) e[i] = (rm1[i] >> 32) | (i < 3 ? rm1[i + 1] << 32 : 0);
  fr64 w = pow(from_u64(7), e);
  for (int i = lg; i < 32; i++) w = mul(w, w);
  return w;
}

// primitive 3*2^lg-th root of unity: 7^((r-1)/(3*2^32)) squared (32-lg)
// times, so its cube is root_of_unity(lg) (3 divides r - 1)
inline fr64 root_of_unity_3n(int lg) {
  uint64_t e[4];
  uint64_t rm1[4] = {kP[0] - 1, kP[1], kP[2], kP[3]};
  for (int i = 0; i < 4; i++) e[i] = (rm1[i] >> 32) | (i < 3 ? rm1[i + 1] << 32 : 0);
  unsigned __int128 rem = 0;
  for (int i = 3; i >= 0; i--) {  // e /= 3, exact
    const unsigned __int128 cur = (rem << 64) | e[i];
    e[i] = (uint64_t)(cur / 3);
    rem = cur % 3;
  }
  fr64 w = pow(from_u64(7), e);
  for (int i = lg; i < 32; i++) w = mul(w, w);
  return w;
}

inline fr64 domain_size_inverse(int lg) {
  uint64_t t[4] = {0, 0, 0, 0};
  t[lg >> 6] = 1ull << (lg & 63);
  fr64 a = {{t[0], t[1], t[2], t[3]}}, rr;
  memcpy(rr.l, kRR, 32);
  return inv(mul(a, rr));
}

}  // namespace host
}  // namespace pl
