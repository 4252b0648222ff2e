// Element-wise Fr operators, batch inversion and the PLONK permutation ratio.
//
//  element-wise   add/sub/mul/div, scalar forms, neg, inv, to/from Montgomery
//                 [function.cuh:14-48; mont_arithmetic.cuh:17-69]
//  batch inverse  out[i] = num[i] / in[i] with x/0 = 0 (the reference's
//                 reciprocal: mont_t.cuh:1117-1126).  Montgomery's trick per
//                 2048-element block, then once more over the block totals,
//                 so one call performs a single field inversion.
//  perm ratio     p_i = prod_j (w_j + beta k_j w^i + gamma) / prod_j (w_j + beta s_j + gamma)
//                 (term shape of zksnark_compute_permutation_constraint.cu:13-14;
//                 the reference builds it from element-wise calls)
//
// The single inversion uses Pornin's optimized binary GCD (fr_inv.hpp) on the
// canonical residue and a Montgomery product with R^3 mod r to return to
// Montgomery form.
#include "pl_internal.hpp"
#include "fr29.hpp"
#include "fr_inv.hpp"

namespace pl {

constexpr int kOpThreads = 256;
// batch inversion: 256 threads x 8 elements per block (the down-sweep keeps
// each thread's eight local prefixes in registers)
constexpr int kOpLgPer = 3;
constexpr int kOpPer = 1 << kOpLgPer;
constexpr int kOpSeg = kOpThreads * kOpPer;

// R^2 and R^3 mod r (R = 2^256)
__host__ __device__ constexpr uint32_t rr_limb(int i) {
  return i == 0 ? 0xf3f29c6du : i == 1 ? 0xc999e990u : i == 2 ? 0x87925c23u : i == 3 ? 0x2b6cedcbu
       : i == 4 ? 0x7254398fu : i == 5 ? 0x05d31496u : i == 6 ? 0x9f59ff11u : 0x0748d9d9u;
}
__host__ __device__ constexpr uint32_t rrr_limb(int i) {
  return i == 0 ? 0x439b73afu : i == 1 ? 0xc62c1807u : i == 2 ? 0x8cf06990u : i == 3 ? 0x1b3e0d18u
       : i == 4 ? 0xc7b5f418u : i == 5 ? 0x73d13c71u : i == 6 ? 0xc8db33e9u : 0x6e2a5bb9u;
}

__device__ __forceinline__ fr const_fr(uint32_t (*f)(int)) {
  fr r;
#pragma unroll
  for (int i = 0; i < 8; i++) r.l[i] = f(i);
  return r;
}

// ---- single inversion ----------------------------------------------------
// Montgomery-form inverse of a (a = x R -> x^-1 R); zero maps to zero.  The
// integer inverse of the canonical value x R comes from the optimized binary
// GCD in fr_inv.hpp (at most 17 rounds of 31 steps),
// run by a whole wave: the four full-width updates of each round go to four
// lanes at once, one code path for the four, stopping once converged
// (about 12 of the 17 rounds for a random input; 70 -> ~34 us against one
// lane).  Call from all lanes of a wave, converged, with a the same on
// every lane (lane 0's is used); every lane returns the inverse.
__device__ fr fr_inv_wave(const fr& a_in) {
  const fr a = fr_reduce_once(a_in);
  fr r;
  inv::inverse_wave(a.l, r.l);  // zero maps to zero
  // r = (xR)^-1 = x^-1 R^-1 ; times R^3 under Montgomery: x^-1 R
  return fr_mul(r, const_fr(rrr_limb));
}

// ---- element-wise ----------------------------------------------------------
enum { OP_ADD = 0, OP_SUB, OP_MUL, OP_DIV, OP_ADD_S, OP_SUB_S, OP_MUL_S, OP_DIV_S, OP_NEG, OP_INV, OP_TO_MONT,
       OP_FROM_MONT };

__global__ void __launch_bounds__(256) vec_op_kernel(int op, const fr* a, const fr* b, fr* out, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  fr x = fr_load(a + i);
  fr r;
  switch (op) {
    case OP_ADD: r = fr_add(x, fr_load(b + i)); break;
    case OP_SUB: r = fr_sub(x, fr_load(b + i)); break;
    case OP_MUL: r = fr_mul(x, fr_load(b + i)); break;
    case OP_ADD_S: r = fr_add(x, fr_load(b)); break;
    case OP_SUB_S: r = fr_sub(x, fr_load(b)); break;
    case OP_MUL_S: r = fr_mul(x, fr_load(b)); break;
    case OP_NEG: r = fr_neg(x); break;
    case OP_TO_MONT: r = fr_mul(x, const_fr(rr_limb)); break;
    default: {  // OP_FROM_MONT: x * 1 (raw)
      fr one_raw = fr_zero();
      one_raw.l[0] = 1;
      r = fr_mul(x, one_raw);
    }
  }
  fr_store(out + i, r);
}

// ---- batch inversion ---------------------------------------------------------
__device__ __forceinline__ void lds_put2(uint4 (*l)[kOpThreads], int t, const fr& v) {
  l[0][t] = make_uint4(v.l[0], v.l[1], v.l[2], v.l[3]);
  l[1][t] = make_uint4(v.l[4], v.l[5], v.l[6], v.l[7]);
}
__device__ __forceinline__ fr lds_get2(uint4 (*l)[kOpThreads], int t) {
  uint4 a = l[0][t], b = l[1][t];
  fr v;
  v.l[0] = a.x; v.l[1] = a.y; v.l[2] = a.z; v.l[3] = a.w;
  v.l[4] = b.x; v.l[5] = b.y; v.l[6] = b.z; v.l[7] = b.w;
  return v;
}

// The batch inversion multiplies in signed radix-2^29 limbs (fr29.hpp) and
// reads every word W as the value W / R' (R' = 2^261, "R'-view"): the product
// P(W1, W2) = W1 W2 / R' is then closed (P of the words of a and b is the word
// of a b), and the R'-view one is the word R' mod r.  Only the single inversion
// and the final words relate to the stored form: a stored word S(x) = x R is
// the R'-view word of x / 32, so a level-0 output P(S(num), inverse of S(x)) =
// 32 S(num / x), and without num the inverse alone is 1024 S(1 / x).  Every
// output is linear in the one field inversion at the top, which absorbs that
// factor (inv_single_kernel).

// R' mod r: the R'-view one
__device__ __forceinline__ fr rp_one() {
  fr r;
  r.l[0] = 0xffffffbau; r.l[1] = 0x00000045u; r.l[2] = 0x0072d846u; r.l[3] = 0x1a25272eu;
  r.l[4] = 0x5dbeee8bu; r.l[5] = 0xfe2eedcdu; r.l[6] = 0x9eefbe41u; r.l[7] = 0x4d043f42u;
  return r;
}

// P(a, b) of two words, canonical
__device__ __forceinline__ fr mul_rp(const fr& a, const fr& b) {
  v9 x, y;
  unpack_ext(a.l, x.l);
  unpack_ext(b.l, y.l);
  return v9_canonical(v9_mul<false, false, false>(x, y.l));
}

// P of two pbd words (values within 1.07 r), as a pbd word: the product lies
// in (-1.02 r, 0.02 r) and its balanced digits fit
__device__ __forceinline__ fr mul_pbd(const fr& a, const fr& b) {
  v9 x;
  uint32_t y[9];
  pbd_unpack(a.l, x.l);
  pbd_unpack(b.l, y);
  const v9 t = v9_mul<false, false, true>(x, y);
  fr o;
  pbd_pack(t.l, o.l);
  return o;
}

// Exclusive product over the 256 thread values (pbd words) of the threads
// before t (suffix: after t), as a pbd word.  Work-efficient (up-sweep, then
// down-sweep): 2 x 255 products per block instead of 8 per thread.  one: the
// pbd word of the R'-view one.
__device__ fr block_scan_excl(const fr& v, uint4 (*l)[kOpThreads], bool suffix, const fr& one) {
  const int t = threadIdx.x;
  const int slot = suffix ? kOpThreads - 1 - t : t;
  lds_put2(l, slot, v);
  for (int d = 1; d < kOpThreads; d <<= 1) {
    __syncthreads();
    const int i = (t + 1) * 2 * d - 1;
    if (i < kOpThreads) lds_put2(l, i, mul_pbd(lds_get2(l, i - d), lds_get2(l, i)));
  }
  __syncthreads();
  if (t == 0) lds_put2(l, kOpThreads - 1, one);
  for (int d = kOpThreads / 2; d >= 1; d >>= 1) {
    __syncthreads();
    const int i = (t + 1) * 2 * d - 1;
    if (i < kOpThreads) {
      const fr left = lds_get2(l, i - d), right = lds_get2(l, i);
      lds_put2(l, i - d, right);
      lds_put2(l, i, mul_pbd(left, right));
    }
  }
  __syncthreads();
  const fr r = lds_get2(l, slot);
  __syncthreads();
  return r;
}

// The same exclusive product by Hillis-Steele steps: 8 levels of products
// instead of 16 (more products in all, a shorter chain), for a single-block
// top level where the latency is the cost.
__device__ fr block_scan_excl_short(fr v, uint4 (*l)[kOpThreads], bool suffix, const fr one) {
  const int t = threadIdx.x;
  lds_put2(l, t, v);
  __syncthreads();
  for (int s = 1; s < kOpThreads; s <<= 1) {
    const int o = suffix ? t + s : t - s;
    const bool has = suffix ? o < kOpThreads : o >= 0;
    const fr other = has ? lds_get2(l, o) : one;
    __syncthreads();
    if (has) v = mul_pbd(v, other);
    lds_put2(l, t, v);
    __syncthreads();
  }
  const int e = suffix ? t + 1 : t - 1;
  const fr ex = (suffix ? e < kOpThreads : e >= 0) ? lds_get2(l, e) : one;
  __syncthreads();
  return ex;
}

// The exclusive prefix product of a and the exclusive suffix product of b
// over the 256 thread values (pbd words), by Hillis-Steele steps with the two
// scans interleaved: 8 levels of two independent products, every thread busy
// (a missing partner multiplies by one).
__device__ void block_scan2_short(fr& a, fr& b, uint4 (*la)[kOpThreads], uint4 (*lb)[kOpThreads], const fr one) {
  const int t = threadIdx.x;
  lds_put2(la, t, a);
  lds_put2(lb, t, b);
  __syncthreads();
  for (int s = 1; s < kOpThreads; s <<= 1) {
    const fr oa = t >= s ? lds_get2(la, t - s) : one;
    const fr ob = t + s < kOpThreads ? lds_get2(lb, t + s) : one;
    __syncthreads();
    a = mul_pbd(a, oa);
    b = mul_pbd(b, ob);
    lds_put2(la, t, a);
    lds_put2(lb, t, b);
    __syncthreads();
  }
  const fr ea = t > 0 ? lds_get2(la, t - 1) : one;
  const fr eb = t + 1 < kOpThreads ? lds_get2(lb, t + 1) : one;
  __syncthreads();
  a = ea;
  b = eb;
}

// Phase 1: total[b] = product over block b of in[i] (zeros taken as one).
// Blocks of 256 x 2^LG elements: LG = kOpLgPer in bulk levels; a single-block
// top level takes the smallest LG that covers it.
template <int LG>
__global__ void __launch_bounds__(kOpThreads) inv_reduce_kernel(const fr* in, uint64_t n, fr* total) {
  constexpr int kOpPer = 1 << LG, kOpLgPer = LG, kOpLgSeg = 8 + LG;
  __shared__ uint4 l[2][kOpThreads];
  // order-free product: thread t takes elements b*kOpSeg + 256 k + t (coalesced)
  const uint64_t b0 = (uint64_t)blockIdx.x << kOpLgSeg;
  fr v[kOpPer];
  static_for<0, kOpPer>([&](auto k_) {
    constexpr int k = decltype(k_)::value;
    const uint64_t i = b0 + (uint64_t)k * kOpThreads + threadIdx.x;
    fr x = i < n ? fr_load(in + i) : rp_one();
    v[k] = fr_is_zero(x) ? rp_one() : x;
  });
  // pairwise tree in registers: eight products of unpacked words, then the
  // products of products (all limbs normalized digits, so any pair is a valid
  // operand pair)
  fr acc;
  if constexpr (kOpPer == 1) {
    acc = v[0];
  } else {
    v9 p[kOpPer / 2];
    static_for<0, kOpPer / 2>([&](auto k_) {
      constexpr int k = decltype(k_)::value;
      v9 x, y;
      unpack_ext(v[k].l, x.l);
      unpack_ext(v[k + kOpPer / 2].l, y.l);
      p[k] = v9_mul<false, false, false>(x, y.l);
    });
    static_for<0, kOpLgPer - 1>([&](auto s_) {
      constexpr int h = kOpPer / 2 >> (decltype(s_)::value + 1);
      static_for<0, h>([&](auto k_) {
        constexpr int k = decltype(k_)::value;
        p[k] = v9_mul<false, false, false>(p[k], p[k + h].l);
      });
    });
    acc = v9_canonical(p[0]);
  }
  lds_put2(l, threadIdx.x, acc);
  __syncthreads();
  for (int s = kOpThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      acc = mul_rp(acc, lds_get2(l, threadIdx.x + s));
      lds_put2(l, threadIdx.x, acc);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) fr_store(total + blockIdx.x, acc);
}

// Phase 3: with binv[b] the inverse of block b's product, out[i] = num[i] /
// in[i] (num == nullptr: one); zero inputs give zero.  out may equal in or num.
// Local prefixes L_k are kept as pbd words (products with balanced digits; a
// product of two values below 1.3 r lies in (-r, 0.03 r), inside a pbd word).
// SHORT: Hillis-Steele thread scans (single-block top level)
template <int LG, bool SHORT>
__global__ void __launch_bounds__(kOpThreads) inv_down_kernel(const fr* in, const fr* num, fr* out, uint64_t n,
                                                               const fr* binv) {
  constexpr int kOpPer = 1 << LG, kOpLgSeg = 8 + LG;
  __shared__ uint4 l[2][kOpThreads];
  const int t = threadIdx.x;
  const uint64_t base = ((uint64_t)blockIdx.x << kOpLgSeg) + (uint64_t)t * kOpPer;
  fr e[kOpPer];
  uint32_t zmask = 0;
  // local prefix L_k = x_0 ... x_k (pbd words in e[]); the thread product is
  // L_(kOpPer-1)
  v9 L;
  static_for<0, kOpPer>([&](auto k_) {
    constexpr int k = decltype(k_)::value;
    fr v = base + k < n ? fr_load(in + base + k) : rp_one();
    if (fr_is_zero(v)) {
      zmask |= 1u << k;
      v = rp_one();
    }
    v9 x;
    unpack_ext(v.l, x.l);
    if constexpr (k == 0) {
      L = x;
      e[0] = v;  // L_0 = x_0 as loaded (below 2^256, beyond a pbd word's range)
    } else {
      // |L_k| < 1.07 r: the product of a value below 1.07 r (or x_0 < 2^256) and
      // x_k < 2^256 lies in (-1.07 r, 0.07 r)
      L = v9_mul<false, false, true>(L, x.l);
      pbd_pack(L.l, e[k].l);
    }
  });
  fr tp;
  pbd_pack(L.l, tp.l);
  fr one;
  {
    const fr o = rp_one();
    to_pbd(o.l, one.l);
  }
  const fr suf = SHORT ? block_scan_excl_short(tp, l, true, one) : block_scan_excl(tp, l, true, one);
  const fr pre = SHORT ? block_scan_excl_short(tp, l, false, one) : block_scan_excl(tp, l, false, one);
  // inv(L_last) = binv * suffix(> t) * prefix(< t): the block product over
  // everything outside this thread's run, times the block's inverse
  v9 cur;
  {
    v9 a, b, c;
    unpack_ext(fr_load(binv + blockIdx.x).l, a.l);
    pbd_unpack(suf.l, b.l);
    pbd_unpack(pre.l, c.l);
    cur = v9_mul<false, false, false>(v9_mul<false, false, false>(a, b.l), c.l);
  }
  // walk back: inv(x_k) = inv(L_k) L_{k-1} (inv(x_0) = inv(L_0));
  // inv(L_{k-1}) = inv(L_k) x_k
  static_for<0, kOpPer>([&](auto k_) {
    constexpr int k = kOpPer - 1 - decltype(k_)::value;
    v9 inv_k = cur;
    if constexpr (k > 0) {
      uint32_t Lp[9];
      if constexpr (k - 1 == 0) unpack_ext(e[0].l, Lp);
      else pbd_unpack(e[k > 1 ? k - 1 : 1].l, Lp);
      inv_k = v9_mul<false, false, false>(cur, Lp);
      // x_k from the original input (one again for zeros)
      fr xk = base + k < n ? fr_load(in + base + k) : rp_one();
      if ((zmask >> k) & 1) xk = rp_one();
      uint32_t X[9];
      unpack_ext(xk.l, X);
      cur = v9_mul<false, false, false>(cur, X);
    }
    if (base + k < n) {
      fr r;
      if ((zmask >> k) & 1) {
        r = fr_zero();
      } else {
        if (num) {
          uint32_t N[9];
          unpack_ext(fr_load(num + base + k).l, N);
          inv_k = v9_mul<false, false, false>(inv_k, N);
        }
        r = v9_canonical(inv_k);
      }
      fr_store(out + base + k, r);
    }
  });
}

// single inversion at the top of the batch: *p <- the R'-view inverse of *p
// times 2^-5 (level-0 outputs with num) or 2^-10 (without), i.e.
// fr_inv_wave(W) times 32 or 1 (see the batch-inversion notes above); one
// block of one wave
__global__ void inv_single_kernel(fr* p, int with_num) {
  if (threadIdx.x < 64 && blockIdx.x == 0) {
    fr z = fr_inv_wave(fr_load(p));
    if (with_num) {
#pragma unroll
      for (int d = 0; d < 5; d++) z = fr_add(z, z);
    }
    if (threadIdx.x == 0) fr_store(p, z);
  }
}

static uint64_t inv_scratch_elems(uint64_t n) {
  uint64_t tot = 1;
  while (n > 1) {
    n = (n + kOpSeg - 1) / kOpSeg;
    tot += 2 * n;
  }
  return tot;
}

// out[i] = num[i] / in[i] (num == nullptr: 1/in[i]); x/0 = 0.  with_num: whether
// the level-0 call has num (sets the factor the single inversion absorbs).
static hipError_t batch_inverse_level(const fr* in, const fr* num, fr* out, uint64_t n, fr* scratch, int with_num,
                                      hipStream_t s) {
  const uint64_t nb = (n + kOpSeg - 1) / kOpSeg;
  fr* tot = scratch;
  fr* tinv = scratch + nb;
  if (nb == 1) {
    // one block: 256 x 2^LG with the smallest LG that covers n, short scans
    const dim3 g(1), b(kOpThreads);
    if (n <= 256) {
      hipLaunchKernelGGL(inv_reduce_kernel<0>, g, b, 0, s, in, n, tot);
      hipLaunchKernelGGL(inv_single_kernel, dim3(1), dim3(64), 0, s, tot, with_num);
      hipLaunchKernelGGL((inv_down_kernel<0, true>), g, b, 0, s, in, num, out, n, (const fr*)tot);
    } else if (n <= 512) {
      hipLaunchKernelGGL(inv_reduce_kernel<1>, g, b, 0, s, in, n, tot);
      hipLaunchKernelGGL(inv_single_kernel, dim3(1), dim3(64), 0, s, tot, with_num);
      hipLaunchKernelGGL((inv_down_kernel<1, true>), g, b, 0, s, in, num, out, n, (const fr*)tot);
    } else if (n <= 1024) {
      hipLaunchKernelGGL(inv_reduce_kernel<2>, g, b, 0, s, in, n, tot);
      hipLaunchKernelGGL(inv_single_kernel, dim3(1), dim3(64), 0, s, tot, with_num);
      hipLaunchKernelGGL((inv_down_kernel<2, true>), g, b, 0, s, in, num, out, n, (const fr*)tot);
    } else {
      hipLaunchKernelGGL(inv_reduce_kernel<kOpLgPer>, g, b, 0, s, in, n, tot);
      hipLaunchKernelGGL(inv_single_kernel, dim3(1), dim3(64), 0, s, tot, with_num);
      hipLaunchKernelGGL((inv_down_kernel<kOpLgPer, true>), g, b, 0, s, in, num, out, n, (const fr*)tot);
    }
    return hipGetLastError();
  }
  hipLaunchKernelGGL(inv_reduce_kernel<kOpLgPer>, dim3((unsigned)nb), dim3(kOpThreads), 0, s, in, n, tot);
  // block totals are never zero (zeros were skipped): invert them as a batch
  hipError_t e = batch_inverse_level(tot, nullptr, tinv, nb, scratch + 2 * nb, with_num, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((inv_down_kernel<kOpLgPer, false>), dim3((unsigned)nb), dim3(kOpThreads), 0, s, in, num, out, n,
                     (const fr*)tinv);
  return hipGetLastError();
}

static hipError_t batch_inverse(const fr* in, const fr* num, fr* out, uint64_t n, fr* scratch, hipStream_t s) {
  return batch_inverse_level(in, num, out, n, scratch, num ? 1 : 0, s);
}

// ---- permutation ratio ------------------------------------------------------
struct RatioArgs {
  const fr* w[3];
  const fr* s[3];
  const fr* beta;
  const fr* gamma;
  const fr* w0;  // w^i = w0[i & 16383] * w1[i >> 14]
  const fr* w1;
  fr* tab;       // [3][tstride] pbd words of (beta k_j w0[lo]) R' (perm_table_kernel)
  uint32_t tstride;  // min(n, 16384): the low-window entries in use
  fr k[3];
  fr k29[3];  // pbd words of (32 k_j) R' = 1024 k_j R mod r
  fr* num;
  fr* den;
  uint64_t n;
};

// Terms in signed radix-2^29 limbs (fr29.hpp).  With S(a) = a R (stored form)
// and the product P(x, y) = x y / 2^261:
//   P(S(w0), beta R') = S(beta w0)                (beta R' = 32 S(beta))
//   P(., (32 k) R')   = (beta k w0) R'            (host constant, pbd word)
//     -- once per call for the 16384 low-window entries (perm_table_kernel) --
//   P((beta k w0) R', S(w1)) = S(beta k w^i)      (w^i = w0[lo] w1[hi])
//   P(S(s_j), beta R') = S(beta s_j)
// so n_j = S(w_j + gamma + beta k_j w^i) and d_j = S(w_j + gamma + beta s_j)
// exactly, and the two products of three terms each carry 2^-10:
// num = S(n_0 n_1 n_2 / 1024), den = S(d_0 d_1 d_2 / 1024), whose ratio is the
// permutation ratio.  Limbs: the inputs are balanced digits (<= 2^28); the
// three-term sums <= 3 2^28 (products with balanced outputs), so a product of
// two of them stays below 81 2^56 + 2^61 < 2^62.8 per column.
// per-thread constants of the terms: B = beta R' (= 32 S(beta)), g = gamma
struct TermConsts {
  uint32_t B[9];
  v9 g;
};
__device__ __forceinline__ TermConsts term_consts(const RatioArgs& a) {
  TermConsts c;
  fr b32 = fr_load(a.beta);
#pragma unroll
  for (int d = 0; d < 5; d++) b32 = fr_add(b32, b32);
  unpack_ext(b32.l, c.B);
  unpack_ext(fr_load(a.gamma).l, c.g.l);
  return c;
}

constexpr int kWin = 16384;  // entries of the low window table
static uint64_t perm_table_elems(uint64_t n) { return 3 * std::min<uint64_t>(n, kWin); }

// tab[j][lo] = P(P(S(w0[lo]), beta R'), (32 k_j) R') = (beta k_j w0[lo]) R', so
// that P(tab[j][lo], S(w1[hi])) = S(beta k_j w^i)
__global__ void __launch_bounds__(256) perm_table_kernel(RatioArgs a, uint32_t count) {
  const uint32_t lo = blockIdx.x * blockDim.x + threadIdx.x;
  if (lo >= count) return;
  const TermConsts c = term_consts(a);
  v9 w;
  unpack_ext(fr_load(a.w0 + lo).l, w.l);
  const v9 bw = v9_mul<false, false, false>(w, c.B);
#pragma unroll
  for (int j = 0; j < 3; j++) {
    uint32_t K[9];
    pbd_unpack(a.k29[j].l, K);
    const v9 t = v9_mul<false, false, true>(bw, K);  // within (-1.02 r, 0.02 r): a pbd word
    fr o;
    pbd_pack(t.l, o.l);
    fr_store(a.tab + j * a.tstride + lo, o);
  }
}

// v (nine signed limbs, |v_i| < 2^31) carried to balanced digits, the same
// integer: digits 0..7 in [-2^28, 2^28), the carries into the top limb
__device__ __forceinline__ void v9_carry_balanced(v9& v) {
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const uint32_t d = sext29(v.l[i]);
    v.l[i + 1] += (uint32_t)(((int32_t)v.l[i] - (int32_t)d) >> 29);  // exact
    v.l[i] = d;
  }
}

// num and den of element i (before reduction).  SK: k = (1, 7, 13), the
// reference's coset constants: t = S(beta w^i) is one product and the other
// two terms are 7 t and 13 t by limb shifts.  t has balanced digits
// (|t_i| <= 2^28), so 7 t_i fits a signed limb; carried back to balanced
// digits, 7 t plus 6 t fits again (13 t = 7 t + 6 t); each carried value is
// the same integer, so n_j = w_j + gamma + k_j t keeps the three-term bound.
template <bool SK>
__device__ __forceinline__ void perm_terms_at(const RatioArgs& a, const TermConsts& c, uint64_t i, v9& num,
                                              v9& den) {
  v9 w1;
  unpack_ext(fr_load(a.w1 + (i >> 14)).l, w1.l);
  v9 t1, t7;  // SK: beta w^i, then 7 beta w^i (balanced digits)
#pragma unroll
  for (int j = 0; j < 3; j++) {
    v9 wj, sj;
    unpack_ext(fr_load(a.w[j] + i).l, wj.l);
    const v9 wg = v9_add(wj, c.g);
    v9 kt;
    if (!SK || j == 0) {
      uint32_t K[9];
      pbd_unpack(fr_load(a.tab + j * a.tstride + (i & (kWin - 1))).l, K);
      kt = v9_mul<false, false, true>(w1, K);
      if (SK) t1 = kt;
    } else if (j == 1) {
#pragma unroll
      for (int l = 0; l < 9; l++) t7.l[l] = (t1.l[l] << 3) - t1.l[l];
      v9_carry_balanced(t7);
      kt = t7;
    } else {
#pragma unroll
      for (int l = 0; l < 9; l++) kt.l[l] = t7.l[l] + (t1.l[l] << 2) + (t1.l[l] << 1);
      v9_carry_balanced(kt);
    }
    const v9 nj = v9_add(wg, kt);
    unpack_ext(fr_load(a.s[j] + i).l, sj.l);
    const v9 dj = v9_add(wg, v9_mul<false, false, true>(sj, c.B));
    num = j == 0 ? nj : v9_mul<false, false, false>(num, nj.l);  // the first factor needs no product
    den = j == 0 ? dj : v9_mul<false, false, false>(den, dj.l);
  }
}

// ZERO_DEN (grand product): a zero den is stored as num = 0, den = R' mod r,
// and the products of num and den over each run of 8 consecutive elements go
// to rn/rd: three butterfly steps across the run's lanes, each one product
// per lane (step 1 pairs num on even lanes with den on odd lanes; after steps
// 2 and 3, lane 8g holds the run's num product and lane 8g + 1 its den
// product).  Lanes past n carry R' mod r.
template <bool ZERO_DEN, bool SK>
__global__ void __launch_bounds__(256) perm_terms_kernel(RatioArgs a, fr* rn, fr* rd) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if constexpr (!ZERO_DEN) {
    if (i >= a.n) return;
    const TermConsts c = term_consts(a);
    v9 num, den;
    perm_terms_at<SK>(a, c, i, num, den);
    fr_store(a.num + i, v9_canonical(num));
    fr_store(a.den + i, v9_canonical(den));
  } else {
    fr nw = rp_one(), dw = rp_one();
    if (i < a.n) {
      const TermConsts c = term_consts(a);
      v9 num, den;
      perm_terms_at<SK>(a, c, i, num, den);
      nw = v9_canonical(num);
      dw = v9_canonical(den);
      if (fr_is_zero(dw)) {
        nw = fr_zero();
        dw = rp_one();
      }
      fr_store(a.num + i, nw);
      fr_store(a.den + i, dw);
    }
    const int lane = threadIdx.x & 7;
    v9 p;
    {
      const fr send = (lane & 1) ? nw : dw;
      v9 mine;
      unpack_ext(((lane & 1) ? dw : nw).l, mine.l);
      fr o;
#pragma unroll
      for (int j = 0; j < 8; j++) o.l[j] = (uint32_t)__shfl_xor((int)send.l[j], 1);
      uint32_t other[9];
      unpack_ext(o.l, other);
      p = v9_mul<false, false, false>(mine, other);
    }
#pragma unroll
    for (int m = 2; m <= 4; m <<= 1) {
      uint32_t q[9];
#pragma unroll
      for (int j = 0; j < 9; j++) q[j] = (uint32_t)__shfl_xor((int)p.l[j], m);
      p = v9_mul<false, false, false>(p, q);
    }
    if ((i & ~(uint64_t)7) < a.n && lane < 2) fr_store((lane ? rd : rn) + (i >> 3), v9_canonical(p));
  }
}

// ---- permutation grand product without a ratio vector -----------------------
// z_i = prod_{j<i} num_j / den_j (exclusive, z_0 = 1) as
//   z_i = [prod_{j<i} num_j] [prod_{j>=i} den_j] / prod_j den_j,
// so the one field inversion multiplies carries only and no element needs an
// inverse of its own.  Up: per block of 2048 elements (256 threads x 8
// consecutive) the thread products of num and den (kept for the down-sweep)
// and the block products, recursively over the block products until one
// block holds them.  Top (one block): the thread scans, the inversion of the
// total den product, and the carries c_b = [prod_{b'<b} num] [prod_{b'>b} den]
// / total for the level below.  Down: each level rescans its blocks from their
// carries; level 0 takes its den suffix inclusive, upper levels exclusive.
// A zero den_j (ratio 0 under x/0 = 0) is stored as num_j = 0, den_j = 1,
// which gives the same z.
// Words are read in the R'-view (the batch-inversion notes above): num_j and
// den_j are stored words S(x), i.e. R'-view words of x / 32; the 32s cancel
// between the two products and the total but for one, which the inverse
// absorbs (32 fr_inv_wave(W) is the R'-view word of 1 / (32 W/R')), so level
// 0 yields the stored words S(z_i).

// Up-sweep over one level: thread products of num and den over each
// thread's 8 consecutive elements (tn, td: one per thread of every block, for
// the down-sweep) and the block products (bn, bd).
__global__ void __launch_bounds__(kOpThreads) gp_up_kernel(const fr* nv, const fr* dv, uint64_t n, fr* tn, fr* td,
                                                            fr* bn, fr* bd) {
  __shared__ uint4 ln[2][kOpThreads], ld[2][kOpThreads];
  const int t = threadIdx.x;
  const uint64_t tid = (uint64_t)blockIdx.x * kOpThreads + t;
  const uint64_t base = tid * kOpPer;
  v9 pn, pd;
  static_for<0, kOpPer>([&](auto k_) {
    constexpr int k = decltype(k_)::value;
    const bool live = base + k < n;
    const fr nw = live ? fr_load(nv + base + k) : rp_one();
    const fr dw = live ? fr_load(dv + base + k) : rp_one();
    v9 x, y;
    unpack_ext(nw.l, x.l);
    unpack_ext(dw.l, y.l);
    pn = k == 0 ? x : v9_mul<false, false, false>(pn, x.l);
    pd = k == 0 ? y : v9_mul<false, false, false>(pd, y.l);
  });
  fr cn = v9_canonical(pn), cd = v9_canonical(pd);
  fr_store(tn + tid, cn);
  fr_store(td + tid, cd);
  lds_put2(ln, t, cn);
  lds_put2(ld, t, cd);
  __syncthreads();
  for (int s = kOpThreads / 2; s > 0; s >>= 1) {
    if (t < s) {
      cn = mul_rp(cn, lds_get2(ln, t + s));
      cd = mul_rp(cd, lds_get2(ld, t + s));
      lds_put2(ln, t, cn);
      lds_put2(ld, t, cd);
    }
    __syncthreads();
  }
  if (t == 0) {
    fr_store(bn + blockIdx.x, cn);
    fr_store(bd + blockIdx.x, cd);
  }
}

// 32 fr_inv_wave(total): the R'-view word of 1 / (32 total / R'); by a whole
// wave, as fr_inv_wave
__device__ __forceinline__ fr gp_inverse(const fr& total) {
  fr z = fr_inv_wave(total);
#pragma unroll
  for (int d = 0; d < 5; d++) z = fr_add(z, z);
  return z;
}

// Down-sweep over one level (in place: out may be nv).  Thread t of block b
// owns elements base = (256 b + t) 2^LG + k, k < 2^LG, and writes
//   out = c_b pre_t suf_t prod_{k'<k} num [prod_{k'>=k} den (INCL) or k' > k]
// with pre_t / suf_t the products over the block's threads before / after t.
// TOP: a single block; the thread products are formed here and c =
// gp_inverse(total den), or c = one with the total stored to *tot when a
// later launch inverts it.  Otherwise c_b = carry[b] and the thread products
// come from the up-sweep (tn, td).
template <int LG, bool INCL, bool TOP>
__device__ __forceinline__ void gp_down_body(const fr* nv, const fr* dv, const fr* tn, const fr* td, uint64_t n,
                                             const fr* carry, fr* out, fr* tot,
                                             uint32_t blk) {
  constexpr int kPer = 1 << LG;
  __shared__ uint4 l[2][kOpThreads], l2[2][kOpThreads];
  __shared__ uint32_t cinv[8];
  const int t = threadIdx.x;
  const uint64_t tid = (uint64_t)blk * kOpThreads + t;
  const uint64_t base = tid * kPer;
  fr one;
  {
    const fr o = rp_one();
    to_pbd(o.l, one.l);
  }
  // local den suffixes S[k] (pbd words; a product of a value below 1.07 r and
  // one below r lies in (-1.07 r, 0.07 r))
  fr S[kPer];
  v9 acc;
  static_for<0, kPer>([&](auto k_) {
    constexpr int k = kPer - 1 - decltype(k_)::value;
    const fr dw = base + k < n ? fr_load(dv + base + k) : rp_one();
    v9 y;
    unpack_ext(dw.l, y.l);
    if constexpr (INCL) {
      acc = k == kPer - 1 ? y : v9_mul<false, false, true>(acc, y.l);
      pbd_pack(acc.l, S[k].l);
    } else {
      if constexpr (k == kPer - 1) {
        S[k] = one;
        acc = y;
      } else {
        pbd_pack(acc.l, S[k].l);
        acc = v9_mul<false, false, true>(acc, y.l);
      }
    }
  });
  // acc: the thread's den product
  fr tnw, tdw;
  if constexpr (TOP) {
    v9 pn;
    static_for<0, kPer>([&](auto k_) {
      constexpr int k = decltype(k_)::value;
      const fr nw = base + k < n ? fr_load(nv + base + k) : rp_one();
      v9 x;
      unpack_ext(nw.l, x.l);
      pn = k == 0 ? x : v9_mul<false, false, false>(pn, x.l);
    });
    pbd_pack(acc.l, tdw.l);
    if constexpr (kPer == 1) {
      pbd_pack(pn.l, tnw.l);  // balanced digits of a canonical value
    } else {
      uint32_t b[9];
      reduce_centered(pn.l, b);
      pbd_pack(b, tnw.l);
    }
  } else {
    to_pbd(fr_load(tn + tid).l, tnw.l);
    to_pbd(fr_load(td + tid).l, tdw.l);
  }
  fr pre = tnw, suf = tdw;
  block_scan2_short(pre, suf, l, l2, one);
  fr cw;
  if constexpr (TOP) {
    if (t < 64) {
      // total den product = suf_0 x (thread 0's product): lane 0's value
      // (wave 0 runs it whole for the wave-wide inversion)
      v9 a;
      uint32_t b[9];
      pbd_unpack(suf.l, a.l);
      pbd_unpack(tdw.l, b);
      const fr total = v9_canonical(v9_mul<false, false, false>(a, b));
      fr z;
      if (tot) {
        if (t == 0) fr_store(tot, total);  // inverted by gp_inverse beside the next sweep
        z = rp_one();
      } else {
        z = gp_inverse(total);
      }
      if (t == 0) {
#pragma unroll
        for (int i = 0; i < 8; i++) cinv[i] = z.l[i];
      }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 8; i++) cw.l[i] = cinv[i];
  } else {
    cw = fr_load(carry + blk);
  }
  v9 cur;
  {
    v9 a;
    uint32_t b[9], c[9];
    unpack_ext(cw.l, a.l);
    pbd_unpack(pre.l, b);
    pbd_unpack(suf.l, c);
    cur = v9_mul<false, false, false>(v9_mul<false, false, false>(a, b), c);
  }
  static_for<0, kPer>([&](auto k_) {
    constexpr int k = decltype(k_)::value;
    const bool live = base + k < n;
    const fr nw = live ? fr_load(nv + base + k) : rp_one();  // read before out[] is written
    if (live) {
      uint32_t sk[9];
      pbd_unpack(S[k].l, sk);
      fr_store(out + base + k, v9_canonical(v9_mul<false, false, false>(cur, sk)));
    }
    if constexpr (k < kPer - 1) {
      uint32_t x[9];
      unpack_ext(nw.l, x);
      cur = v9_mul<false, false, false>(cur, x);
    }
  });
}

// bulk levels: three waves per SIMD (the local den suffixes stay in registers)
// inv != nullptr: one more block than the level needs; block 0 (dispatched
// first) has its wave 0 write gp_inverse(*tot) to inv while the other blocks
// sweep (the inverse then multiplies the run carries, gp_run_down_kernel)
template <bool INCL>
__global__ void __launch_bounds__(kOpThreads) __attribute__((amdgpu_waves_per_eu(3)))
gp_down_kernel(const fr* nv, const fr* dv, const fr* tn, const fr* td, uint64_t n, const fr* carry, fr* out,
               const fr* tot, fr* inv) {
  uint32_t blk = blockIdx.x;
  if (inv) {
    if (blk == 0) {
      if (threadIdx.x < 64) {
        const fr z = gp_inverse(fr_load(tot));
        if (threadIdx.x == 0) fr_store(inv, z);
      }
      return;
    }
    blk--;
  }
  gp_down_body<kOpLgPer, INCL, false>(nv, dv, tn, td, n, carry, out, nullptr, blk);
}

// the single-block top level
// tot == nullptr: the inversion runs here (no level below the top but the
// runs); otherwise the total is stored for gp_down_kernel's extra block
template <int LG, bool INCL>
__global__ void __launch_bounds__(kOpThreads) gp_top_kernel(const fr* nv, const fr* dv, uint64_t n, fr* out,
                                                             fr* tot) {
  gp_down_body<LG, INCL, true>(nv, dv, nullptr, nullptr, n, nullptr, out, tot, 0);
}

// Level 0: thread = run of 8 consecutive elements with carry c (the num
// products of the runs before it, the den products of those after it, the
// inverse): z_k = c prod_{k'<k} num prod_{k'>=k} den.  No scans.
__global__ void __launch_bounds__(256) gp_run_down_kernel(const fr* nv, const fr* dv, uint64_t n, const fr* carry,
                                                          const fr* inv, fr* out) {
  const uint64_t run = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t base = run * kOpPer;
  if (base >= n) return;
  // local den suffixes (pbd words; products of a value below 1.07 r and one
  // below r lie in (-1.07 r, 0.07 r))
  fr S[kOpPer];
  v9 acc;
  static_for<0, kOpPer>([&](auto k_) {
    constexpr int k = kOpPer - 1 - decltype(k_)::value;
    const fr dw = base + k < n ? fr_load(dv + base + k) : rp_one();
    v9 y;
    unpack_ext(dw.l, y.l);
    acc = k == kOpPer - 1 ? y : v9_mul<false, false, true>(acc, y.l);
    pbd_pack(acc.l, S[k].l);
  });
  v9 cur;
  unpack_ext(fr_load(carry + run).l, cur.l);
  if (inv) {
    uint32_t w[9];
    unpack_ext(fr_load(inv).l, w);
    cur = v9_mul<false, false, false>(cur, w);
  }
  // four quarters: each quarter's num values are loaded before its outputs
  // are written (out may equal nv), so a thread waits for four loads, not
  // eight, with few values live
  constexpr int Q = kOpPer / 4;
  static_for<0, 4>([&](auto g_) {
    constexpr int k0 = decltype(g_)::value * Q;
    fr nw[Q];
    static_for<0, Q>([&](auto j_) {
      constexpr int j = decltype(j_)::value;
      nw[j] = base + k0 + j < n ? fr_load(nv + base + k0 + j) : rp_one();
    });
    static_for<0, Q>([&](auto j_) {
      constexpr int j = decltype(j_)::value, k = k0 + j;
      if (base + k < n) {
        uint32_t sk[9];
        pbd_unpack(S[k].l, sk);
        fr_store(out + base + k, v9_canonical(v9_mul<false, false, false>(cur, sk)));
      }
      if constexpr (k < kOpPer - 1) {
        uint32_t x[9];
        unpack_ext(nw[j].l, x);
        cur = v9_mul<false, false, false>(cur, x);
      }
    });
  });
}

// level sizes of the grand product: the runs of 8 elements (n_1 = ceil(n / 8)),
// then n_{L+1} = ceil(n_L / 2048) while n_L > 2048; scratch holds den (n), the
// run products (2 n_1), then per up-swept level the thread products (2 x 256
// per block) and the next level's num and den, then the total and its inverse
static uint64_t gp_scratch_elems(uint64_t n) {
  uint64_t m = (n + kOpPer - 1) / kOpPer;
  uint64_t tot = n + 2 * m + 2;  // + the total den product and its inverse
  while (m > (uint64_t)kOpSeg) {
    const uint64_t nb = (m + kOpSeg - 1) / kOpSeg;
    tot += 2 * nb * kOpThreads + 2 * nb;
    m = nb;
  }
  return tot;
}

template <bool INCL>
static void gp_top(const fr* nv, const fr* dv, uint64_t n, fr* out, fr* tot, hipStream_t s) {
  const dim3 g(1), b(kOpThreads);
  if (n <= 256) hipLaunchKernelGGL((gp_top_kernel<0, INCL>), g, b, 0, s, nv, dv, n, out, tot);
  else if (n <= 512) hipLaunchKernelGGL((gp_top_kernel<1, INCL>), g, b, 0, s, nv, dv, n, out, tot);
  else if (n <= 1024) hipLaunchKernelGGL((gp_top_kernel<2, INCL>), g, b, 0, s, nv, dv, n, out, tot);
  else hipLaunchKernelGGL((gp_top_kernel<kOpLgPer, INCL>), g, b, 0, s, nv, dv, n, out, tot);
}

uint64_t ops_scratch_elems(uint64_t n) {
  return perm_table_elems(n) + std::max(n + inv_scratch_elems(n), gp_scratch_elems(n));
}

hipError_t vec_op(int op, const fr* a, const fr* b, fr* out, uint64_t n, fr* scratch, hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (op == OP_INV) return batch_inverse(a, nullptr, out, n, scratch, s);
  if (op == OP_DIV) return batch_inverse(b, a, out, n, scratch, s);
  if (op == OP_DIV_S) {
    // out = a * (1/b): invert the scalar into scratch, then multiply
    fr* binv = scratch;
    hipError_t e = batch_inverse(b, nullptr, binv, 1, scratch + 1, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(vec_op_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (int)OP_MUL_S, a,
                       (const fr*)binv, out, n);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(vec_op_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, op, a, b, out, n);
  return hipGetLastError();
}

// ---- powers and broadcasts ------------------------------------------------------
// exp_mod: out[i] = a[i]^e, left-to-right square and multiply with the
// exponent uniform over the launch; e = 0 gives one and e = 1 a copy
// (mont_arithmetic.cu:258-273, cpu/mont_arithmetic.cpp:66-81).
__global__ void __launch_bounds__(256) exp_kernel(const fr* a, fr* out, uint64_t n, uint32_t e) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  fr_store(out + i, e == 0 ? fr_one() : fr_pow_u64(fr_load(a + i), e));
}

// gen_sequence: out[i] = x^i (poly_eval_kernel, mont_arithmetic.cu:104-114,
// raises x to each i alone).  Here the head out[0..min(n, 2^14)) is raised
// directly (at most 14 squarings), and every later tile of 2^14 elements is
// the head times one factor: out[(q << 14) + j] = out[j] * x^(q << 14), with
// x^(q << 14) = out[q]^(2^14) (q < 2^14 for n <= 2^28), one product per element.
constexpr uint32_t kSeqLgTile = 14;
__global__ void __launch_bounds__(256) seq_head_kernel(const fr* x, fr* out, uint64_t n0) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n0) return;
  fr_store(out + i, fr_pow_u64(fr_load(x), i));
}
// block b: tile q = b + 1; reads out[q] and out[0..2^14), writes tile q only
__global__ void __launch_bounds__(256) seq_tail_kernel(fr* out, uint64_t n) {
  __shared__ uint4 hq[2];
  const uint64_t q = (uint64_t)blockIdx.x + 1;
  if (threadIdx.x == 0) {
    fr h = fr_load(out + q);
    for (uint32_t k = 0; k < kSeqLgTile; k++) h = fr_sqr(h);
    hq[0] = make_uint4(h.l[0], h.l[1], h.l[2], h.l[3]);
    hq[1] = make_uint4(h.l[4], h.l[5], h.l[6], h.l[7]);
  }
  __syncthreads();
  fr h;
  h.l[0] = hq[0].x; h.l[1] = hq[0].y; h.l[2] = hq[0].z; h.l[3] = hq[0].w;
  h.l[4] = hq[1].x; h.l[5] = hq[1].y; h.l[6] = hq[1].z; h.l[7] = hq[1].w;
  const uint64_t base = q << kSeqLgTile;
  for (uint32_t j = threadIdx.x; j < (1u << kSeqLgTile) && base + j < n; j += blockDim.x)
    fr_store(out + base + j, fr_mul(fr_load(out + j), h));
}

// repeat_to_poly: out[i] = *x (repeat_kernel, mont_arithmetic.cu:214-220)
__global__ void __launch_bounds__(256) repeat_kernel(const fr* x, fr* out, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  fr_store(out + i, fr_load(x));
}

hipError_t exp_mod(const fr* a, fr* out, uint64_t n, uint32_t e, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(exp_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, out, n, e);
  return hipGetLastError();
}

hipError_t gen_sequence(const fr* x, fr* out, uint64_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const uint64_t n0 = std::min<uint64_t>(n, 1ull << kSeqLgTile);
  hipLaunchKernelGGL(seq_head_kernel, dim3((unsigned)((n0 + 255) / 256)), dim3(256), 0, s, x, out, n0);
  if (n > n0) {
    const uint64_t tiles = (n + (1ull << kSeqLgTile) - 1) >> kSeqLgTile;
    hipLaunchKernelGGL(seq_tail_kernel, dim3((unsigned)(tiles - 1)), dim3(256), 0, s, out, n);
  }
  return hipGetLastError();
}

hipError_t repeat_to_poly(const fr* x, fr* out, uint64_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(repeat_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, out, n);
  return hipGetLastError();
}

static RatioArgs ratio_args(const fr* const w[3], const fr* const sg[3], const fr* beta, const fr* gamma,
                            const fr* w0, const fr* w1, const host::fr64 k[3], uint64_t n) {
  RatioArgs a;
  for (int j = 0; j < 3; j++) {
    a.w[j] = w[j];
    a.s[j] = sg[j];
    memcpy(a.k[j].l, k[j].l, 32);
    // (32 k_j) R' = 1024 k_j R: the Montgomery product of k_j R and 1024 R
    const host::fr64 k1024 = host::mul(k[j], host::from_u64(1024));
    fr kk;
    memcpy(kk.l, k1024.l, 32);
    to_pbd(kk.l, a.k29[j].l);
  }
  a.beta = beta;
  a.gamma = gamma;
  a.w0 = w0;
  a.w1 = w1;
  a.tab = nullptr;
  a.tstride = 0;
  a.num = nullptr;
  a.den = nullptr;
  a.n = n;
  return a;
}

// k = (1, 7, 13) in Montgomery form, the coset constants of the miss_ops
// permutation workload (terms of the shape of
// zksnark_compute_permutation_constraint.cu:13-14): the terms kernel then
// forms k_1 t and k_2 t by limb shifts
static bool reference_ks(const host::fr64 k[3]) {
  const uint64_t v[3] = {1, 7, 13};
  for (int j = 0; j < 3; j++) {
    const host::fr64 m = host::from_u64(v[j]);
    if (memcmp(m.l, k[j].l, sizeof(m.l)) != 0) return false;
  }
  return true;
}

// the window tables at the front of scratch (perm_table_elems(n) elements)
static void perm_tables(RatioArgs& a, fr*& scratch, hipStream_t s) {
  const uint32_t count = (uint32_t)std::min<uint64_t>(a.n, kWin);
  a.tab = scratch;
  a.tstride = count;
  scratch += perm_table_elems(a.n);
  hipLaunchKernelGGL(perm_table_kernel, dim3((count + 255) / 256), dim3(256), 0, s, a, count);
}

hipError_t perm_ratio(const fr* const w[3], const fr* const sg[3], const fr* beta, const fr* gamma,
                      const fr* w0, const fr* w1, const host::fr64 k[3], fr* ratio, uint64_t n, fr* scratch,
                      hipStream_t s) {
  RatioArgs a = ratio_args(w, sg, beta, gamma, w0, w1, k, n);
  perm_tables(a, scratch, s);
  a.num = ratio;
  a.den = scratch;
  if (reference_ks(k))
    hipLaunchKernelGGL((perm_terms_kernel<false, true>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a,
                       nullptr, nullptr);
  else
    hipLaunchKernelGGL((perm_terms_kernel<false, false>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a,
                       nullptr, nullptr);
  // ratio = num / den in place over num
  return batch_inverse(scratch, ratio, ratio, n, scratch + n, s);
}

hipError_t perm_grand_product(const fr* const w[3], const fr* const sg[3], const fr* beta, const fr* gamma,
                              const fr* w0, const fr* w1, const host::fr64 k[3], fr* z, uint64_t n, fr* scratch,
                              hipStream_t s) {
  RatioArgs a = ratio_args(w, sg, beta, gamma, w0, w1, k, n);
  perm_tables(a, scratch, s);
  a.num = z;
  a.den = scratch;
  const uint64_t nr = (n + kOpPer - 1) / kOpPer;
  fr* rn = scratch + n;
  fr* rd = rn + nr;
  if (reference_ks(k))
    hipLaunchKernelGGL((perm_terms_kernel<true, true>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, rn,
                       rd);
  else
    hipLaunchKernelGGL((perm_terms_kernel<true, false>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, rn,
                       rd);
  // levels above the elements: the runs, then block products
  struct Level {
    fr *nv, *dv, *tn, *td;
    uint64_t n;
  } lv[8];
  int top = 0;
  lv[0] = {rn, rd, nullptr, nullptr, nr};
  fr* p = rd + nr;
  while (lv[top].n > (uint64_t)kOpSeg) {
    Level& c = lv[top];
    const uint64_t nb = (c.n + kOpSeg - 1) / kOpSeg;
    c.tn = p;
    c.td = p + nb * kOpThreads;
    p += 2 * nb * kOpThreads;
    fr* bn = p;
    fr* bd = p + nb;
    p += 2 * nb;
    hipLaunchKernelGGL(gp_up_kernel, dim3((unsigned)nb), dim3(kOpThreads), 0, s, (const fr*)c.nv, (const fr*)c.dv,
                       c.n, c.tn, c.td, bn, bd);
    lv[++top] = {bn, bd, nullptr, nullptr, nb};
  }
  // carries in place over each level's num (the runs' carries over rn).  With
  // a level between the top and the runs, the top stores the total den product
  // and the first sweep below it inverts it in an extra block.
  fr* tot = top > 0 ? p : nullptr;
  fr* inv = top > 0 ? p + 1 : nullptr;
  gp_top<false>(lv[top].nv, lv[top].dv, lv[top].n, lv[top].nv, tot, s);
  for (int L = top - 1; L >= 0; L--) {
    const Level& c = lv[L];
    const bool first = L == top - 1;
    const unsigned nb = (unsigned)((c.n + kOpSeg - 1) / kOpSeg) + (first ? 1 : 0);
    hipLaunchKernelGGL(gp_down_kernel<false>, dim3(nb), dim3(kOpThreads), 0, s, c.nv, c.dv, c.tn, c.td, c.n,
                       (const fr*)lv[L + 1].nv, c.nv, (const fr*)tot, first ? inv : nullptr);
  }
  hipLaunchKernelGGL(gp_run_down_kernel, dim3((unsigned)((nr + 255) / 256)), dim3(256), 0, s, (const fr*)z,
                     (const fr*)scratch, n, (const fr*)rn, (const fr*)inv, z);
  return hipGetLastError();
}

}  // namespace pl
