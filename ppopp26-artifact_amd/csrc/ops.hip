// Element-wise Fr operators, batch inversion and the PLONK permutation ratio.
//
//  element-wise   add/sub/mul/div, scalar forms, neg, inv, to/from Montgomery
//                 [function.cuh:14-48; mont_arithmetic.cuh:17-69]
//  batch inverse  out[i] = num[i] / in[i] with x/0 = 0 (the reference's
//                 reciprocal: mont_t.cuh:1117-1126).  Montgomery's trick per
//                 4096-element block, then once more over the block totals,
//                 so one call performs a single field inversion.
//  perm ratio     p_i = prod_j (w_j + beta k_j w^i + gamma) / prod_j (w_j + beta s_j + gamma)
//                 (term shape of zksnark_compute_permutation_constraint.cu:13-14;
//                 the reference builds it from element-wise calls)
//
// The single inversion uses the binary extended Euclidean algorithm on the
// canonical residue (variable time) and a Montgomery product with R^3 mod r
// to return to Montgomery form.
#include "pl_internal.hpp"
#include "fr_inv.hpp"

namespace pl {

constexpr int kOpThreads = 256;
constexpr int kOpPer = 16;
constexpr int kOpSeg = kOpThreads * kOpPer;
constexpr int kOpLgSeg = 12;

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

// Inclusive prefix (dir = +1) or suffix (dir = -1) products of the 256 thread
// values; returns the exclusive value for this thread (one at the edge).
__device__ fr block_scan_prod(fr v, uint4 (*l)[kOpThreads], bool suffix) {
  const int t = threadIdx.x;
  lds_put2(l, t, v);
  __syncthreads();
  for (int s = 1; s < kOpThreads; s <<= 1) {
    const int o = suffix ? t + s : t - s;
    const bool has = suffix ? o < kOpThreads : o >= 0;
    fr other = has ? lds_get2(l, o) : fr_one();
    __syncthreads();
    if (has) v = fr_mul(v, other);
    lds_put2(l, t, v);
    __syncthreads();
  }
  const int e = suffix ? t + 1 : t - 1;
  fr ex = (suffix ? e < kOpThreads : e >= 0) ? lds_get2(l, e) : fr_one();
  __syncthreads();
  return ex;
}

// Phase 1: total[b] = product over block b of in[i] (zeros taken as one)
__global__ void __launch_bounds__(kOpThreads) inv_reduce_kernel(const fr* in, uint64_t n, fr* total) {
  __shared__ uint4 l[2][kOpThreads];
  // order-free product: thread t takes elements b*4096 + 256 k + t (coalesced)
  const uint64_t b0 = (uint64_t)blockIdx.x << kOpLgSeg;
  fr v[kOpPer];
  static_for<0, kOpPer>([&](auto k_) {
    constexpr int k = decltype(k_)::value;
    const uint64_t i = b0 + (uint64_t)k * kOpThreads + threadIdx.x;
    fr x = i < n ? fr_load(in + i) : fr_one();
    v[k] = fr_is_zero(x) ? fr_one() : x;
  });
  static_for<0, 4>([&](auto s_) {
    constexpr int h = kOpPer >> (decltype(s_)::value + 1);
    static_for<0, h>([&](auto k_) {
      constexpr int k = decltype(k_)::value;
      v[k] = fr_mul(v[k], v[k + h]);
    });
  });
  fr acc = v[0];
  lds_put2(l, threadIdx.x, acc);
  __syncthreads();
  for (int s = kOpThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      acc = fr_mul(acc, lds_get2(l, threadIdx.x + s));
      lds_put2(l, threadIdx.x, acc);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) fr_store(total + blockIdx.x, acc);
}

// Phase 3: with binv[b] = (block product)^-1, out[i] = num[i] / in[i]
// (num == nullptr: one); zero inputs give zero.  out may equal in or num.
__global__ void __launch_bounds__(kOpThreads) inv_down_kernel(const fr* in, const fr* num, fr* out, uint64_t n,
                                                               const fr* binv) {
  __shared__ uint4 l[2][kOpThreads];
  const int t = threadIdx.x;
  const uint64_t base = ((uint64_t)blockIdx.x << kOpLgSeg) + (uint64_t)t * kOpPer;
  fr e[kOpPer];
  uint32_t zmask = 0;
  static_for<0, kOpPer>([&](auto k_) {
    constexpr int k = decltype(k_)::value;
    fr v = base + k < n ? fr_load(in + base + k) : fr_one();
    if (fr_is_zero(v)) {
      zmask |= 1u << k;
      v = fr_one();
    }
    e[k] = v;
  });
  // local prefix L_k = e_0 ... e_k in e[]; the thread product is L_15
  static_for<1, kOpPer>([&](auto k_) {
    constexpr int k = decltype(k_)::value;
    e[k] = fr_mul(e[k - 1], e[k]);
  });
  const fr tp = e[kOpPer - 1];
  fr suf = block_scan_prod(tp, l, true);   // product of threads > t
  fr pre = block_scan_prod(tp, l, false);  // product of threads < t
  // inv(L_15) = binv * suffix(> t) * prefix(< t): the block product over
  // everything outside this thread's run, times the block's inverse
  fr cur = fr_mul(fr_mul(fr_load(binv + blockIdx.x), suf), pre);
  // walk back: inv(x_k) = inv(L_k) L_{k-1} (inv(x_0) = inv(L_0));
  // inv(L_{k-1}) = inv(L_k) x_k
  static_for<0, kOpPer>([&](auto k_) {
    constexpr int k = kOpPer - 1 - decltype(k_)::value;
    fr inv_k = cur;
    if (k > 0) {
      inv_k = fr_mul(cur, e[k > 0 ? k - 1 : 0]);
      // x_k from the original input (one again for zeros)
      fr xk = base + k < n ? fr_load(in + base + k) : fr_one();
      if ((zmask >> k) & 1) xk = fr_one();
      cur = fr_mul(cur, xk);
    }
    if (base + k < n) {
      fr r = ((zmask >> k) & 1) ? fr_zero() : inv_k;
      if (num) r = fr_mul(r, fr_load(num + base + k));
      fr_store(out + base + k, r);
    }
  });
}

// single-thread inversion of one element in place
__global__ void inv_single_kernel(fr* p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) fr_store(p, fr_inv_single(fr_load(p)));
}

static uint64_t inv_scratch_elems(uint64_t n) {
  uint64_t tot = 1;
  while (n > 1) {
    n = (n + kOpSeg - 1) / kOpSeg;
    tot += 2 * n;
  }
  return tot;
}

// out[i] = num[i] / in[i] (num == nullptr: 1/in[i]); x/0 = 0
static hipError_t batch_inverse(const fr* in, const fr* num, fr* out, uint64_t n, fr* scratch, hipStream_t s) {
  const uint64_t nb = (n + kOpSeg - 1) / kOpSeg;
  fr* tot = scratch;
  fr* tinv = scratch + nb;
  hipLaunchKernelGGL(inv_reduce_kernel, dim3((unsigned)nb), dim3(kOpThreads), 0, s, in, n, tot);
  if (nb == 1) {
    hipLaunchKernelGGL(inv_single_kernel, dim3(1), dim3(64), 0, s, tot);
    hipLaunchKernelGGL(inv_down_kernel, dim3((unsigned)nb), dim3(kOpThreads), 0, s, in, num, out, n, (const fr*)tot);
    return hipGetLastError();
  }
  // block totals are never zero (zeros were skipped): invert them as a batch
  hipError_t e = batch_inverse(tot, nullptr, tinv, nb, scratch + 2 * nb, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(inv_down_kernel, dim3((unsigned)nb), dim3(kOpThreads), 0, s, in, num, out, n, (const fr*)tinv);
  return hipGetLastError();
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
  fr* num;
  fr* den;
  uint64_t n;
};

__global__ void __launch_bounds__(256) perm_terms_kernel(RatioArgs a) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const fr beta = fr_load(a.beta), gamma = fr_load(a.gamma);
  const fr wi = fr_mul(fr_load(a.w0 + (i & 16383)), fr_load(a.w1 + (i >> 14)));
  const fr bx = fr_mul(beta, wi);  // beta * w^i
  fr num, den;
#pragma unroll
  for (int j = 0; j < 3; j++) {
    const fr wj = fr_add(fr_load(a.w[j] + i), gamma);
    const fr nj = fr_add(wj, fr_mul_u(bx, a.k[j]));
    const fr dj = fr_add(wj, fr_mul(beta, fr_load(a.s[j] + i)));
    num = j == 0 ? nj : fr_mul(num, nj);  // the first factor needs no product
    den = j == 0 ? dj : fr_mul(den, dj);
  }
  fr_store(a.num + i, num);
  fr_store(a.den + i, den);
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
