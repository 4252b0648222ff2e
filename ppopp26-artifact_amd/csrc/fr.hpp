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

__device__ __forceinline__ bool fr_eq(const fr& a, const fr& b) {
  uint32_t o = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) o |= a.l[i] ^ b.l[i];
  return o == 0;
}

// ---- carry-chain primitives ----------------------------------------------
// Add/sub chains use the compiler's carry builtins (it picks VOP2/VOP3
// encodings within the constant-bus limit and keeps r's limbs in VGPRs).
// Only the multiply-accumulate needs inline asm: v_mad_u64_u32's carry-out
// of the 64-bit sum is not reachable from C.  A carry there is a wave64 lane
// mask in an SGPR pair.
typedef uint64_t lanemask;

// ---- multiply-accumulate blocks --------------------------------------------
// The column accumulator of the product is (acc: 64 bits, ext: 32 bits).
// Each limb product is one v_mad_u64_u32 into acc whose carry-out (a lane
// mask) is folded into ext.  A carry read must come two VALU instructions
// after its write, so products are issued in blocks of three: three mads
// (carries to three SGPR pairs), then three folds, each fold exactly two
// instructions after its mad.  A block is one asm statement, so the wait
// state the compiler adds after every asm statement is paid once per three
// products.  Blocks of two and one put s_nop fillers where a third product
// would be.  FIRST: the block opens a column (ext = first carry).  SY: the y
// operands are wave-uniform (SGPRs).
#define PL_MAD3 "v_mad_u64_u32 %0, %2, %6, %7, %0\n\tv_mad_u64_u32 %0, %3, %8, %9, %0\n\tv_mad_u64_u32 %0, %4, %10, %11, %0\n\t"
#define PL_FOLD3_FIRST "v_cndmask_b32_e64 %1, 0, 1, %2\n\tv_addc_co_u32_e64 %1, %5, %1, 0, %3\n\tv_addc_co_u32_e64 %1, %5, %1, 0, %4"
#define PL_FOLD3 "v_addc_co_u32_e64 %1, %5, %1, 0, %2\n\tv_addc_co_u32_e64 %1, %5, %1, 0, %3\n\tv_addc_co_u32_e64 %1, %5, %1, 0, %4"
#define PL_MAC3(FOLD, EXT, K0, K1, K2)                                                          \
  asm(PL_MAD3 FOLD : "+v"(acc), EXT(ext), "=&s"(c0), "=&s"(c1), "=&s"(c2), "=&s"(d)                \
      : "v"(x0), K0(y0), "v"(x1), K1(y1), "v"(x2), K2(y2))
#define PL_MAC3_ALL(FOLD, EXT)                                              \
  if constexpr (S0 && S1 && S2) PL_MAC3(FOLD, EXT, "s", "s", "s");           \
  else if constexpr (S0 && S1) PL_MAC3(FOLD, EXT, "s", "s", "v");            \
  else if constexpr (S0 && S2) PL_MAC3(FOLD, EXT, "s", "v", "s");            \
  else if constexpr (S0) PL_MAC3(FOLD, EXT, "s", "v", "v");                  \
  else if constexpr (S1 && S2) PL_MAC3(FOLD, EXT, "v", "s", "s");            \
  else if constexpr (S1) PL_MAC3(FOLD, EXT, "v", "s", "v");                  \
  else if constexpr (S2) PL_MAC3(FOLD, EXT, "v", "v", "s");                  \
  else PL_MAC3(FOLD, EXT, "v", "v", "v");
// S0..S2: y operand i is wave-uniform (SGPR)
template <bool FIRST, bool S0, bool S1, bool S2>
__device__ __forceinline__ void mac3(uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1, uint32_t x2, uint32_t y2,
                                     uint64_t& acc, uint32_t& ext) {
  lanemask c0, c1, c2, d;
  if constexpr (FIRST) {
    PL_MAC3_ALL(PL_FOLD3_FIRST, "=&v")
  } else {
    PL_MAC3_ALL(PL_FOLD3, "+v")
  }
}
#define PL_MAD2 "v_mad_u64_u32 %0, %2, %5, %6, %0\n\tv_mad_u64_u32 %0, %3, %7, %8, %0\n\ts_nop 0\n\t"
#define PL_FOLD2_FIRST "v_cndmask_b32_e64 %1, 0, 1, %2\n\tv_addc_co_u32_e64 %1, %4, %1, 0, %3"
#define PL_FOLD2 "v_addc_co_u32_e64 %1, %4, %1, 0, %2\n\tv_addc_co_u32_e64 %1, %4, %1, 0, %3"
#define PL_MAC2(FOLD, EXT, K0, K1) \
  asm(PL_MAD2 FOLD : "+v"(acc), EXT(ext), "=&s"(c0), "=&s"(c1), "=&s"(d) : "v"(x0), K0(y0), "v"(x1), K1(y1))
#define PL_MAC2_ALL(FOLD, EXT)                                   \
  if constexpr (S0 && S1) PL_MAC2(FOLD, EXT, "s", "s");          \
  else if constexpr (S0) PL_MAC2(FOLD, EXT, "s", "v");           \
  else if constexpr (S1) PL_MAC2(FOLD, EXT, "v", "s");           \
  else PL_MAC2(FOLD, EXT, "v", "v");
template <bool FIRST, bool S0, bool S1>
__device__ __forceinline__ void mac2(uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1, uint64_t& acc,
                                     uint32_t& ext) {
  lanemask c0, c1, d;
  if constexpr (FIRST) {
    PL_MAC2_ALL(PL_FOLD2_FIRST, "=&v")
  } else {
    PL_MAC2_ALL(PL_FOLD2, "+v")
  }
}
#define PL_MAD1 "v_mad_u64_u32 %0, %2, %4, %5, %0\n\ts_nop 1\n\t"
#define PL_MAC1(FOLD, EXT, K0) asm(PL_MAD1 FOLD : "+v"(acc), EXT(ext), "=&s"(c0), "=&s"(d) : "v"(x0), K0(y0))
template <bool FIRST, bool S0>
__device__ __forceinline__ void mac1(uint32_t x0, uint32_t y0, uint64_t& acc, uint32_t& ext) {
  lanemask c0, d;
  if constexpr (FIRST && S0) PL_MAC1("v_cndmask_b32_e64 %1, 0, 1, %2", "=&v", "s");
  else if constexpr (FIRST) PL_MAC1("v_cndmask_b32_e64 %1, 0, 1, %2", "=&v", "v");
  else if constexpr (S0) PL_MAC1("v_addc_co_u32_e64 %1, %3, %1, 0, %2", "+v", "s");
  else PL_MAC1("v_addc_co_u32_e64 %1, %3, %1, 0, %2", "+v", "v");
}
#undef PL_MAD3
#undef PL_FOLD3_FIRST
#undef PL_FOLD3
#undef PL_MAC3
#undef PL_MAC3_ALL
#undef PL_MAD2
#undef PL_FOLD2_FIRST
#undef PL_FOLD2
#undef PL_MAC2
#undef PL_MAC2_ALL
#undef PL_MAD1
#undef PL_MAC1

// L products x(i) * y(i), i < L, opening a column, into (acc, ext) in
// blocks of three; sy(i) tells whether y(i) is wave-uniform
template <int L, typename FX, typename FY, typename SYF>
__device__ __forceinline__ void mac_column(FX&& fx, FY&& fy, SYF sy, uint64_t& acc, uint32_t& ext) {
  using std::integral_constant;
  static_for<0, L / 3>([&](auto b_) {
    constexpr int i = 3 * decltype(b_)::value;
    mac3<i == 0, sy(i), sy(i + 1), sy(i + 2)>(fx(integral_constant<int, i>{}), fy(integral_constant<int, i>{}),
                                               fx(integral_constant<int, i + 1>{}), fy(integral_constant<int, i + 1>{}),
                                               fx(integral_constant<int, i + 2>{}), fy(integral_constant<int, i + 2>{}),
                                               acc, ext);
  });
  constexpr int R = L % 3, i = L - R;
  if constexpr (R == 2)
    mac2<i == 0, sy(i), sy(i + 1)>(fx(integral_constant<int, i>{}), fy(integral_constant<int, i>{}),
                                   fx(integral_constant<int, i + 1>{}), fy(integral_constant<int, i + 1>{}), acc, ext);
  else if constexpr (R == 1)
    mac1<i == 0, sy(i)>(fx(integral_constant<int, i>{}), fy(integral_constant<int, i>{}), acc, ext);
}

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
    // column k: m_i * r_{k-i} (i < k, r index >= 1), then a_i * b_{k-i} (i <= k)
    struct SY {
      constexpr bool operator()(int j) const { return j < k ? true : UB; }
    };
    mac_column<2 * k + 1>(
        [&](auto j_) {
          constexpr int j = decltype(j_)::value;
          if constexpr (j < k) return m[j];
          else return a.l[j - k];
        },
        [&](auto j_) {
          constexpr int j = decltype(j_)::value;
          if constexpr (j < k) return r_limb(k - j);
          else return b.l[2 * k - j];
        },
        SY{}, acc, ext);
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
    constexpr int i0 = k - 7, L = 15 - k;  // i in [k-7, 8)
    struct SY {
      constexpr bool operator()(int j) const { return j < L ? true : UB; }
    };
    mac_column<2 * L>(
        [&](auto j_) {
          constexpr int j = decltype(j_)::value;
          if constexpr (j < L) return m[i0 + j];
          else return a.l[i0 + j - L];
        },
        [&](auto j_) {
          constexpr int j = decltype(j_)::value;
          if constexpr (j < L) return r_limb(k - i0 - j);
          else return b.l[k - i0 - (j - L)];
        },
        SY{}, acc, ext);
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
