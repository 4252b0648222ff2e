// Harness over the reference's own CPU field (vendored blst, compiled from
// /root/reference/PLONK/blst/src/assembly.S and the header
// PLONK/utils/mont/cpu/ff/bls12-381.hpp).  TEST INFRASTRUCTURE ONLY: it pins
// the oracle's field arithmetic against the reference's arithmetic and is the
// timed "reference" CPU baseline for the serial grand product and division.
// Built by oracle/Makefile into oracle/_ref/ (never committed).
#include <cstdint>
#include <cstring>
#include "ff/bls12-381.hpp"

using rfr = cpu::BLS12_381_Fr_G1;

static inline rfr ld(const uint64_t* p) { rfr x; std::memcpy(&x, p, 32); return x; }
static inline void st(uint64_t* p, const rfr& x) { std::memcpy(p, &x, 32); }

extern "C" {
void ref_mul(uint64_t* o, const uint64_t* a, const uint64_t* b) { rfr x = ld(a), y = ld(b); st(o, x * y); }
void ref_add(uint64_t* o, const uint64_t* a, const uint64_t* b) { rfr x = ld(a), y = ld(b); st(o, x + y); }
void ref_sub(uint64_t* o, const uint64_t* a, const uint64_t* b) { rfr x = ld(a), y = ld(b); st(o, x - y); }
void ref_inv(uint64_t* o, const uint64_t* a) { rfr x = ld(a); st(o, 1 / x); }

// Loop of naive_gp (miss_ops.cu:34-47) over the reference's field type:
// out[i] = prod_{j<=i} in[j] for i <= n-2, out[n-1] = 0.
void ref_naive_gp(const uint64_t* in, uint64_t* out, uint64_t n) {
  if (n == 0) return;
  std::memset(out, 0, 32 * n);
  static const uint64_t one[4] = {8589934590UL, 6378425256633387010UL,
                                  11064306276430008309UL, 1739710354780652911UL};
  rfr state = ld(one);
  st(out, state);
  const rfr* inp = reinterpret_cast<const rfr*>(in);
  for (uint64_t i = 0; i + 1 < n; ++i) {
    state *= inp[i];
    st(out + 4 * i, state);
  }
}

// Loop of naive_div (miss_ops.cu:49-87): long division by [c, d1], n-1
// quotient coefficients.  Stops at an empty remainder where the reference
// would read remainder_[-1].
void ref_naive_div(const uint64_t* in, uint64_t n, const uint64_t* c, const uint64_t* d1, uint64_t* q) {
  if (n < 2) return;
  rfr* rem = new rfr[n];
  std::memcpy(rem, in, 32 * n);
  std::memset(q, 0, 32 * (n - 1));
  rfr div[2] = {ld(c), ld(d1)};
  rfr lead_inv = 1 / div[1];
  rfr* qq = reinterpret_cast<rfr*>(q);
  uint64_t len = n;
  while (len >= 2) {
    rfr cur = rem[len - 1] * lead_inv;
    uint64_t deg = len - 2;
    qq[deg] = cur;
    for (int i = 0; i < 2; ++i) {
      rfr t = cur * div[i];
      rem[deg + i] -= t;
    }
    while (len > 0 && rem[len - 1].is_zero()) len--;
  }
  delete[] rem;
}
}

extern "C" {
// Horner evaluation sum c_i x^i over the reference's field type (the
// reference has no CPU evaluation; its GPU evaluate is function.cu:170-181):
// the one-core CPU baseline of the evaluation.
void ref_eval(const uint64_t* c, uint64_t n, const uint64_t* x, uint64_t* y) {
  rfr acc;
  std::memset(&acc, 0, 32);
  const rfr xx = ld(x);
  for (uint64_t i = n; i-- > 0;) acc = acc * xx + ld(c + 4 * i);
  st(y, acc);
}

// The PLONK permutation grand product of the miss_ops permutation workload over
// the reference's field type, element by element as the reference's
// element-wise operators compose it (function.cu:41-152; the quotient term
// shape of zksnark_compute_permutation_constraint.cu:13-14, div_mod as a
// reciprocal per element, x/0 = 0), then the exclusive product z:
//   p_i = prod_j (w_j + beta k_j omega^i + gamma) / prod_j (w_j + beta s_j + gamma)
//   z_0 = 1, z_{i+1} = z_i p_i.
void ref_perm_grand_product(const uint64_t* const* w, const uint64_t* const* s, const uint64_t* beta,
                            const uint64_t* gamma, const uint64_t* k, const uint64_t* omega, uint64_t n,
                            uint64_t* z) {
  static const uint64_t one[4] = {8589934590UL, 6378425256633387010UL, 11064306276430008309UL,
                                  1739710354780652911UL};
  const rfr b = ld(beta), g = ld(gamma), om = ld(omega);
  rfr bk[3];
  for (int j = 0; j < 3; j++) bk[j] = b * ld(k + 4 * j);
  rfr wi = ld(one), acc = ld(one);
  for (uint64_t i = 0; i < n; i++) {
    st(z + 4 * i, acc);
    rfr num = ld(one), den = ld(one);
    for (int j = 0; j < 3; j++) {
      const rfr wj = ld(w[j] + 4 * i);
      num = num * (wj + bk[j] * wi + g);
      den = den * (wj + b * ld(s[j] + 4 * i) + g);
    }
    acc = acc * (num * (1 / den));
    wi = wi * om;
  }
}
}
