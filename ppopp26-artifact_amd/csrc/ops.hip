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
// GCD in fr_inv.hpp (17 rounds of 31 steps, no data-dependent trip count).
__device__ fr fr_inv_single(const fr& a_in) {
  const fr a = fr_reduce_once(a_in);
  if (fr_is_zero(a)) return fr_zero();
  fr r;
  inv::inverse(a.l, r.l);
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
// fr_inv_single(W) times 32 or 1 (see the batch-inversion notes above)
__global__ void inv_single_kernel(fr* p, int with_num) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    fr z = fr_inv_single(fr_load(p));
    if (with_num) {
#pragma unroll
      for (int d = 0; d < 5; d++) z = fr_add(z, z);
    }
    fr_store(p, z);
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
  fr k[3];
  fr k29[3];  // pbd words of (32 k_j) R' = 1024 k_j R mod r
  fr* num;
  fr* den;
  uint64_t n;
};

// Terms in signed radix-2^29 limbs (fr29.hpp).  With S(a) = a R (stored form)
// and the product P(x, y) = x y / 2^261:
//   P(S(w0), S(w1)) = S(w^i / 32)               (two stored-form windows)
//   P(., beta R')   = S(beta w^i / 32)          (beta R' = 32 S(beta))
//   P(., (32 k) R') = S(beta k w^i)             (host constant, pbd word)
//   P(S(s_j), beta R') = S(beta s_j)
// so n_j = S(w_j + gamma + beta k_j w^i) and d_j = S(w_j + gamma + beta s_j)
// exactly, and the two products of three terms each carry 2^-10:
// num = S(n_0 n_1 n_2 / 1024), den = S(d_0 d_1 d_2 / 1024), whose ratio is the
// permutation ratio.  Limbs: the inputs are balanced digits (<= 2^28); the
// three-term sums <= 3 2^28 (products with balanced outputs), so a product of
// two of them stays below 81 2^56 + 2^61 < 2^62.8 per column.
__global__ void __launch_bounds__(256) perm_terms_kernel(RatioArgs a) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  fr b32 = fr_load(a.beta);
#pragma unroll
  for (int d = 0; d < 5; d++) b32 = fr_add(b32, b32);  // beta R' = 32 S(beta)
  uint32_t B[9];
  unpack_ext(b32.l, B);
  v9 g, w0, w1;
  unpack_ext(fr_load(a.gamma).l, g.l);
  unpack_ext(fr_load(a.w0 + (i & 16383)).l, w0.l);
  unpack_ext(fr_load(a.w1 + (i >> 14)).l, w1.l);
  const v9 wi = v9_mul<false, false, false>(w0, w1.l);
  const v9 bx = v9_mul<false, false, false>(wi, B);
  v9 num, den;
#pragma unroll
  for (int j = 0; j < 3; j++) {
    v9 wj, sj;
    unpack_ext(fr_load(a.w[j] + i).l, wj.l);
    const v9 wg = v9_add(wj, g);
    uint32_t K[9];
    pbd_unpack(a.k29[j].l, K);
    const v9 nj = v9_add(wg, v9_mul<false, false, true>(bx, K));
    unpack_ext(fr_load(a.s[j] + i).l, sj.l);
    const v9 dj = v9_add(wg, v9_mul<false, false, true>(sj, B));
    num = j == 0 ? nj : v9_mul<false, false, false>(num, nj.l);  // the first factor needs no product
    den = j == 0 ? dj : v9_mul<false, false, false>(den, dj.l);
  }
  fr_store(a.num + i, v9_canonical(num));
  fr_store(a.den + i, v9_canonical(den));
}

uint64_t ops_scratch_elems(uint64_t n) { return n + inv_scratch_elems(n); }

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

hipError_t perm_ratio(const fr* const w[3], const fr* const sg[3], const fr* beta, const fr* gamma,
                      const fr* w0, const fr* w1, const host::fr64 k[3], fr* ratio, uint64_t n, fr* scratch,
                      hipStream_t s) {
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
  a.num = ratio;
  a.den = scratch;
  a.n = n;
  hipLaunchKernelGGL(perm_terms_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
  // ratio = num / den in place over num
  return batch_inverse(scratch, ratio, ratio, n, scratch + n, s);
}

}  // namespace pl
