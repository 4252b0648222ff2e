// BLS12-381 scalar field Fr in Montgomery form, for gfx950 (CDNA4) VALU.
//
// An element is 8 little-endian 32-bit limbs (32 bytes), always canonical:
// value in [0, r), stored as x*R mod r with R = 2^256.  The byte layout is the
// one of the reference's `cuda::BLS12_381_Fr_G1` (8 x u32) and
// `cpu::BLS12_381_Fr_G1` (4 x u64): PLONK/utils/mont/cuda/ff/bls12-381.hpp:7-45,
// PLONK/utils/mont/cpu/ff/bls12-381.hpp:6-32.
//
// Arithmetic (all results canonical, bit-identical to any other exact
// implementation):
//   add  : 8-limb add chain, trial subtraction of r, select      (reference mont_t.cuh:215-222)
//   sub  : 8-limb sub chain, masked add-back of r                (reference mont_t.cuh:267-290)
//   mul  : Montgomery product a*b*R^-1 mod r, finely integrated
//          product scanning (FIPS) over 32-bit limbs.  Each limb product
//          is one v_mad_u64_u32 (64-bit accumulate, carry-out to an SGPR
//          lane mask) plus one v_addc_co_u32 folding that carry into a
//          third accumulator word, issued in blocks of three products so
//          that every carry is read two instructions after it is written.
//          r = 1 (mod 2^32), so the Montgomery digit is m = -t (mod 2^32)
//          and m*r0 = m needs no multiply.                        (reference mont_t.cuh:396-419)
// v_mad_u64_u32 runs at the full VALU rate on gfx950, so the carry handling
// is as expensive as the multiplies; FIPS keeps it to one op per product.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace pl {

// Compile-time loop: f(std::integral_constant<int, I>) for I in [B, E).
// Used wherever a loop indexes register arrays with field ops in its body;
// the unroller's size threshold is too small for those bodies.
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

struct __align__(16) fr {
  uint32_t l[8];
};

// r = 0x73eda753299d7d483339d80809a1d80553bda402fffe5bfeffffffff00000001
#define PL_R0 0x00000001u
#define PL_R1 0xffffffffu
#define PL_R2 0xfffe5bfeu
#define PL_R3 0x53bda402u
#define PL_R4 0x09a1d805u
#define PL_R5 0x3339d808u
#define PL_R6 0x299d7d48u
#define PL_R7 0x73eda753u

__host__ __device__ constexpr uint32_t r_limb(int i) {
  return i == 0 ? PL_R0 : i == 1 ? PL_R1 : i == 2 ? PL_R2 : i == 3 ? PL_R3
       : i == 4 ? PL_R4 : i == 5 ? PL_R5 : i == 6 ? PL_R6 : PL_R7;
}

// R mod r (Montgomery one)
__host__ __device__ constexpr uint32_t one_limb(int i) {
  return i == 0 ? 0xfffffffeu : i == 1 ? 0x00000001u : i == 2 ? 0x00034802u
       : i == 3 ? 0x5884b7fau : i == 4 ? 0xecbc4ff5u : i == 5 ? 0x998c4fefu
       : i == 6 ? 0xacc5056fu : 0x1824b159u;
}

__device__ __forceinline__ fr fr_zero() {
  fr z;
#pragma unroll
  for (int i = 0; i < 8; i++) z.l[i] = 0;
  return z;
}

__device__ __forceinline__ fr fr_one() {
  fr z;
#pragma unroll
  for (int i = 0; i < 8; i++) z.l[i] = one_limb(i);
  return z;
}

__device__ __forceinline__ bool fr_is_zero(const fr& a) {
  uint32_t o = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) o |= a.l[i];
  return o == 0;
}

// ---- carry-chain primitives ----------------------------------------------
// Add/sub chains use the compiler's carry builtins (it picks VOP2/VOP3
// encodings within the constant-bus limit and keeps r's limbs in VGPRs).
// Only the multiply-accumulate needs inline asm: v_mad_u64_u32's carry-out
// of the 64-bit sum is not reachable from C.  A carry there is a wave64 lane
// mask in an SGPR pair.
typedef uint64_t lanemask;

// ---- multiply-accumulate columns ------------------------------------------
// The column accumulator of the product is (acc: 64 bits, ext: 32 bits).
// Each limb product is one v_mad_u64_u32 into acc whose carry-out (a lane
// mask) is folded into ext.  A carry must be read at least two VALU
// instructions after it is written; each column is one asm statement,
// software-pipelined over three carry registers so that every fold lands
// two or more instructions after its mad with no filler, and the wait state
// the compiler places after every asm statement is paid once per column.
// MontCol<K, UB>::run adds column K's products (m_i r_{K-i}, then a_i b_{K-i})
// and opens the column (ext = first carry); UB: b is wave-uniform (SGPRs).
#include "mont_columns.inc"

// ---- field ops -----------------------------------------------------------

// a - r if a >= r (a < 2r, a fits 256 bits)
__device__ __forceinline__ fr fr_reduce_once(const fr& a) {
  fr d;
  uint32_t b = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) d.l[i] = __builtin_subc(a.l[i], r_limb(i), b, &b);
  fr o;
#pragma unroll
  for (int i = 0; i < 8; i++) o.l[i] = b ? a.l[i] : d.l[i];  // borrow: a < r, keep a
  return o;
}

__device__ __forceinline__ fr fr_add(const fr& a, const fr& b) {
  fr s;
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) s.l[i] = __builtin_addc(a.l[i], b.l[i], c, &c);
  // a + b < 2r < 2^256: no carry out of limb 7
  return fr_reduce_once(s);
}

__device__ __forceinline__ fr fr_sub(const fr& a, const fr& b) {
  fr d;
  uint32_t bo = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) d.l[i] = __builtin_subc(a.l[i], b.l[i], bo, &bo);
  // borrow: add r back under a mask
  const uint32_t mask = 0u - bo;
  fr o;
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) o.l[i] = __builtin_addc(d.l[i], r_limb(i) & mask, c, &c);
  return o;
}

__device__ __forceinline__ fr fr_neg(const fr& a) {
  return fr_sub(fr_zero(), a);
}

// Montgomery multiplication, FIPS.  acc is a 96-bit column accumulator
// (lo64 = acc, top word = ext).
// UB = true: b is wave-uniform (compile-time constant or kernel argument) and
// is read from SGPRs, saving VGPRs and v_mov traffic.
template <bool UB>
__device__ __forceinline__ fr fr_mul_t(const fr& a, const fr& b) {
  uint32_t m[8];
  uint64_t acc = 0;
  uint32_t ext = 0;
  static_for<0, 8>([&](auto k_) {
    constexpr int k = decltype(k_)::value;
    MontCol<k, UB>::run(acc, ext, m, a.l, b.l);
    // Montgomery digit: m_k = -lo (mod 2^32) since -r^-1 = -1 (mod 2^32);
    // lo + m_k*r0 = lo + m_k = 2^32 * (lo != 0), so the next column starts
    // at (ext:hi) + (lo != 0) -- a 64-bit add, no carry flags
    const uint32_t lo = (uint32_t)acc;
    m[k] = 0u - lo;
    acc = ((((uint64_t)ext) << 32) | (acc >> 32)) + __builtin_elementwise_min(lo, 1u);
  });
  fr t;
  static_for<8, 15>([&](auto k_) {
    constexpr int k = decltype(k_)::value;
    MontCol<k, UB>::run(acc, ext, m, a.l, b.l);
    t.l[k - 8] = (uint32_t)acc;
    acc = ((uint64_t)ext << 32) | (uint32_t)(acc >> 32);
  });
  t.l[7] = (uint32_t)acc;
  // result < 2r < 2^256, so the bits above limb 7 are zero here
  return fr_reduce_once(t);
}

#ifdef PL_MUL_NOINLINE
__device__ __noinline__ fr fr_mul_call(fr a, fr b) { return fr_mul_t<false>(a, b); }
__device__ __forceinline__ fr fr_mul(const fr& a, const fr& b) { return fr_mul_call(a, b); }
__device__ __forceinline__ fr fr_mul_u(const fr& a, const fr& b) { return fr_mul_call(a, b); }
#else
__device__ __forceinline__ fr fr_mul(const fr& a, const fr& b) { return fr_mul_t<false>(a, b); }
__device__ __forceinline__ fr fr_mul_u(const fr& a, const fr& b) { return fr_mul_t<true>(a, b); }
#endif
__device__ __forceinline__ fr fr_sqr(const fr& a) { return fr_mul(a, a); }

// a^e, e a 64-bit exponent (square and multiply, MSB first)
__device__ __forceinline__ fr fr_pow_u64(const fr& a, uint64_t e) {
  fr r = fr_one();
  if (e == 0) return r;
  int top = 63 - __builtin_clzll(e);
  r = a;
  for (int i = top - 1; i >= 0; i--) {
    r = fr_sqr(r);
    if ((e >> i) & 1) r = fr_mul(r, a);
  }
  return r;
}

// ---- memory helpers: one element = two 16-byte vector accesses -----------
__device__ __forceinline__ fr fr_load(const fr* p) {
  const uint4* q = reinterpret_cast<const uint4*>(p);
  uint4 x = q[0], y = q[1];
  fr r;
  r.l[0] = x.x; r.l[1] = x.y; r.l[2] = x.z; r.l[3] = x.w;
  r.l[4] = y.x; r.l[5] = y.y; r.l[6] = y.z; r.l[7] = y.w;
  return r;
}
__device__ __forceinline__ void fr_store(fr* p, const fr& a) {
  uint4* q = reinterpret_cast<uint4*>(p);
  q[0] = make_uint4(a.l[0], a.l[1], a.l[2], a.l[3]);
  q[1] = make_uint4(a.l[4], a.l[5], a.l[6], a.l[7]);
}

}  // namespace pl
