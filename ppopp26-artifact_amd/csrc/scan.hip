// Scans over Fr: grand product, division by (X + c), evaluation.
//
//  grand product  z[0] = 1, z[i+1] = z[i] * p[i]     (exclusive prefix product)
//                 [accumulate_mul_poly, mont_arithmetic.cu:181-201, 333-361]
//  division       s_i = a_i + z * s_{i+1}, s_n = 0, z = -c; q[i] = s_{i+1},
//                 q[n-1] = 0, remainder a(z) = s_0     (affine suffix scan)
//                 [poly_div_poly, mont_arithmetic.cu:165-179, 203-212, 304-331]
//  evaluation     y = sum c_i x^i = s_0 of the same scan with z = x
//                 [evaluate, mont_arithmetic.cu:104-163, 283-302]
//
// The reference runs log2(n) Hillis-Steele passes over the whole array (one
// mulmod per element per pass).  Here every scan is work-efficient: a block
// owns 4096 consecutive elements (256 threads x 16), a reduce kernel folds
// each block to one value, the block values are scanned by the same code one
// level up (recursively), and a downsweep kernel rescans each block from its
// carry.  About 3 mulmods per element in total, three memory sweeps.
#include "fr29.hpp"
#include "pl_internal.hpp"

namespace pl {

constexpr int kScanThreads = 256;
#ifndef PL_SCAN_LGPER
#define PL_SCAN_LGPER 4
#endif
constexpr int kLgPer = PL_SCAN_LGPER;        // elements per thread: 2^kLgPer
constexpr int kPer = 1 << kLgPer;
constexpr int kSeg = kScanThreads * kPer;    // elements per block
constexpr int kLgSeg = 8 + kLgPer;

static inline unsigned sgrid(uint64_t n) { return (unsigned)((n + kSeg - 1) / kSeg); }

__device__ __forceinline__ void lds_put(uint4 (*l)[kScanThreads], int t, const fr& v) {
  l[0][t] = make_uint4(v.l[0], v.l[1], v.l[2], v.l[3]);
  l[1][t] = make_uint4(v.l[4], v.l[5], v.l[6], v.l[7]);
}
__device__ __forceinline__ fr lds_get(uint4 (*l)[kScanThreads], int t) {
  uint4 a = l[0][t], b = l[1][t];
  fr v;
  v.l[0] = a.x; v.l[1] = a.y; v.l[2] = a.z; v.l[3] = a.w;
  v.l[4] = b.x; v.l[5] = b.y; v.l[6] = b.z; v.l[7] = b.w;
  return v;
}

// ---------------------------------------------------------------------------
// multiplicative prefix scan
// ---------------------------------------------------------------------------

// thread-level exclusive prefix across the block (Hillis-Steele over 256
// thread totals, multiplicative); returns the product of threads < t
__device__ fr block_excl_prod(fr v, uint4 (*l)[kScanThreads], fr* block_total) {
  const int t = threadIdx.x;
  fr incl = v;
  lds_put(l, t, incl);
  __syncthreads();
  for (int s = 1; s < kScanThreads; s <<= 1) {
    fr other = t >= s ? lds_get(l, t - s) : fr_one();
    __syncthreads();
    if (t >= s) incl = fr_mul(incl, other);
    lds_put(l, t, incl);
    __syncthreads();
  }
  fr excl = t > 0 ? lds_get(l, t - 1) : fr_one();
  if (block_total) *block_total = lds_get(l, kScanThreads - 1);
  __syncthreads();
  return excl;
}

// Block geometry of a scan kernel: 256 threads x 2^LG elements.  LG = kLgPer
// for the bulk; LG = 0 for a top level of at most 256 values, where the
// latency of the per-thread chains (2^LG dependent products) dominates.
template <int LG>
struct ScanGeom {
  static constexpr int kPer = 1 << LG;
  static constexpr int kLgSeg = 8 + LG;
};

// partial[b] = product of block b's elements.  The product is order-free,
// so thread t takes elements b*kSeg + 256 k + t: every load is coalesced.
template <int LG>
__global__ void __launch_bounds__(kScanThreads) pfx_reduce_kernel(const fr* in, uint64_t n, fr* partial) {
  constexpr int kPer = ScanGeom<LG>::kPer, kLgSeg = ScanGeom<LG>::kLgSeg, kLgPer = LG;
  __shared__ uint4 l[2][kScanThreads];
  const uint64_t b0 = (uint64_t)blockIdx.x << kLgSeg;
  fr v[kPer];
  static_for<0, kPer>([&](auto k_) {
    constexpr int k = decltype(k_)::value;
    const uint64_t i = b0 + (uint64_t)k * kScanThreads + threadIdx.x;
    v[k] = i < n ? fr_load(in + i) : fr_one();
  });
  // pairwise tree in registers (independent products)
  static_for<0, kLgPer>([&](auto s_) {
    constexpr int h = kPer >> (decltype(s_)::value + 1);
    static_for<0, h>([&](auto k_) {
      constexpr int k = decltype(k_)::value;
      v[k] = fr_mul(v[k], v[k + h]);
    });
  });
  fr acc = v[0];
  // tree product over the block
  lds_put(l, threadIdx.x, acc);
  __syncthreads();
  for (int s = kScanThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      acc = fr_mul(acc, lds_get(l, threadIdx.x + s));
      lds_put(l, threadIdx.x, acc);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) fr_store(partial + blockIdx.x, acc);
}

// out_excl[i] = carry_b * prod_{block start <= j < i} in[j]; optional inclusive
// output and block total (single-block use).  carry == nullptr means one.
template <int LG>
__global__ void __launch_bounds__(kScanThreads) pfx_down_kernel(const fr* in, fr* out_excl, fr* out_incl, uint64_t n,
                                                                 const fr* carry, fr* total) {
  constexpr int kPer = ScanGeom<LG>::kPer, kLgSeg = ScanGeom<LG>::kLgSeg;
  __shared__ uint4 l[2][kScanThreads];
  const uint64_t base = ((uint64_t)blockIdx.x << kLgSeg) + (uint64_t)threadIdx.x * kPer;
  fr e[kPer];
  static_for<0, kPer>([&](auto k_) {
    constexpr int k = decltype(k_)::value;
    e[k] = base + k < n ? fr_load(in + base + k) : fr_one();
  });
  // local inclusive products in place: e[k] = in[base] * ... * in[base+k]
  static_for<1, kPer>([&](auto k_) {
    constexpr int k = decltype(k_)::value;
    e[k] = fr_mul(e[k - 1], e[k]);
  });
  fr btot;
  fr pre = block_excl_prod(e[kPer - 1], l, &btot);
  if (carry) {
    fr c = fr_load(carry + blockIdx.x);
    pre = fr_mul(c, pre);
    btot = fr_mul(c, btot);
  }
  if (total && blockIdx.x == 0 && threadIdx.x == 0) fr_store(total, btot);
  static_for<0, kPer>([&](auto k_) {
    constexpr int k = decltype(k_)::value;
    if (base + k < n) {
      fr ex = k == 0 ? pre : fr_mul(pre, e[k > 0 ? k - 1 : 0]);
      if (out_incl) fr_store(out_incl + base + k, fr_mul(pre, e[k]));
      fr_store(out_excl + base + k, ex);
    }
  });
}

// Exclusive prefix product of n elements; scratch needs pfx_scratch_elems(n).
static uint64_t pfx_scratch_elems(uint64_t n) {
  uint64_t tot = 0;
  while (n > (uint64_t)kSeg) {
    n = (n + kSeg - 1) / kSeg;
    tot += 2 * n;
  }
  return tot + 1;
}

static hipError_t prefix_product(const fr* in, fr* out_excl, fr* out_incl, uint64_t n, fr* total, fr* scratch,
                                 hipStream_t s) {
  if (n <= (uint64_t)kScanThreads) {
    hipLaunchKernelGGL(pfx_down_kernel<0>, dim3(1), dim3(kScanThreads), 0, s, in, out_excl, out_incl, n,
                       (const fr*)nullptr, total);
    return hipGetLastError();
  }
  if (n <= (uint64_t)kSeg) {
    hipLaunchKernelGGL(pfx_down_kernel<kLgPer>, dim3(1), dim3(kScanThreads), 0, s, in, out_excl, out_incl, n,
                       (const fr*)nullptr, total);
    return hipGetLastError();
  }
  const uint64_t nb = (n + kSeg - 1) / kSeg;
  fr* part = scratch;
  fr* part_excl = scratch + nb;
  hipLaunchKernelGGL(pfx_reduce_kernel<kLgPer>, dim3(sgrid(n)), dim3(kScanThreads), 0, s, in, n, part);
  hipError_t e = prefix_product(part, part_excl, nullptr, nb, total, scratch + 2 * nb, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(pfx_down_kernel<kLgPer>, dim3(sgrid(n)), dim3(kScanThreads), 0, s, in, out_excl, out_incl, n,
                     (const fr*)part_excl, (fr*)nullptr);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// affine suffix scan  s_i = a_i + m * s_{i+1}
// pw: device table of m^(2^j) R' mod r, j = 0..kLgSeg (m^kSeg last), R' = 2^261:
// every product of the scan multiplies a value by a power of m, so it runs as
// the radix-2^29 product (fr29.hpp) with the power in that form: the Horner
// steps in signed limbs (horner29), the block combines on canonical values.
// ---------------------------------------------------------------------------

// a * w R' / R' = a w (stored form), canonical
__device__ __forceinline__ fr mulp(const fr& a, const fr& w) { return fr_reduce_once(fr_mul29(a, w)); }
// R' mod r: the R'-form of one
__host__ __device__ constexpr uint32_t rp_one_limb(int i) {
  return i == 0 ? 0xffffffbau : i == 1 ? 0x00000045u : i == 2 ? 0x0072d846u : i == 3 ? 0x1a25272eu
       : i == 4 ? 0x5dbeee8bu : i == 5 ? 0xfe2eedcdu : i == 6 ? 0x9eefbe41u : 0x4d043f42u;
}

// block suffix combine: V_t <- V_t + m^(2^LG * 2^s) V_{t+2^s}; returns the
// suffix value starting at thread t (the block's carry folded in by the caller)
template <int LG>
__device__ fr block_suffix(fr v, uint4 (*l)[kScanThreads], const fr* pw) {
  constexpr int kLgPer = LG;
  const int t = threadIdx.x;
  lds_put(l, t, v);
  __syncthreads();
  for (int s = 0; (1 << s) < kScanThreads; s++) {
    const int o = 1 << s;
    fr other = t + o < kScanThreads ? lds_get(l, t + o) : fr_zero();
    __syncthreads();
    if (t + o < kScanThreads) v = fr_add(v, mulp(other, fr_load(pw + kLgPer + s)));  // m^(kPer * 2^s)
    lds_put(l, t, v);
    __syncthreads();
  }
  return v;
}

// sum_k v_k w^k by Horner in signed limbs (fr29.hpp): h <- P(h, w R') + v_k,
// limb-wise sums of a product (digits in [0, 2^29)) and balanced digits, so
// every limb stays below 2^30 and the value within (-r, 1.02 r); canonical
// once at the end.  w9: balanced digits of w R'.
template <int K>
__device__ __forceinline__ fr horner29(const fr (&v)[K], const uint32_t (&w9)[9]) {
  v9 h;
  unpack_ext(v[K - 1].l, h.l);
  static_for<0, K - 1>([&](auto k_) {
    constexpr int k = K - 2 - decltype(k_)::value;
    v9 x;
    unpack_ext(v[k].l, x.l);
    h = v9_add(v9_mul<false, false, false>(h, w9), x);
  });
  return v9_canonical(h);
}

// The powers table beyond level 0: pw[from + j] = pw[from + j - 1]^2 for
// j < count - from.  A multi-level scan's first reduce launch runs it in one
// extra block (block 0, dispatched first) beside the reduction; only the
// levels above read those entries.
struct PwExt {
  fr* pw;
  int from, count;  // count == 0: no extension
};
__device__ void pw_extend(const PwExt& e) {
  fr m = fr_load(e.pw + e.from - 1);
  for (int j = e.from; j < e.count; j++) {
    m = mulp(m, m);
    fr_store(e.pw + j, m);
  }
}

// partial[b] = sum over block b of a_k m^(k - start).  The sum is
// order-free: thread t takes elements start + 256 j + t (coalesced loads),
// H_t = sum_j a m^(256 j) by Horner in m^256, then m^t H_t; a plain sum over
// the block finishes it.
template <int LG>
__global__ void __launch_bounds__(kScanThreads) sfx_reduce_kernel(const fr* a, uint64_t n, const fr* pw, fr* partial,
                                                                   PwExt ext, const fr* mtab) {
  constexpr int kPer = ScanGeom<LG>::kPer, kLgSeg = ScanGeom<LG>::kLgSeg;
  __shared__ uint4 l[2][kScanThreads];
  const int t = threadIdx.x;
  uint32_t blk = blockIdx.x;
  if (ext.count) {
    if (blk == 0) {
      if (t == 0) pw_extend(ext);
      return;
    }
    blk--;
  }
  const uint64_t b0 = (uint64_t)blk << kLgSeg;
  uint32_t M[9];
  unpack_ext(fr_load(pw + 8).l, M);  // m^256, balanced digits
  fr v[kPer];
  static_for<0, kPer>([&](auto k_) {
    constexpr int k = decltype(k_)::value;
    const uint64_t i = b0 + (uint64_t)k * kScanThreads + t;
    v[k] = i < n ? fr_load(a + i) : fr_zero();
  });
  fr h = horner29<kPer>(v, M);
  // m^t R': from the table when the caller built one (sfx_powers_kernel),
  // else from the powers m^(2^bit) R'
  fr mt;
  if (mtab) {
    mt = fr_load(mtab + t);
  } else {
#pragma unroll
    for (int i = 0; i < 8; i++) mt.l[i] = rp_one_limb(i);
    static_for<0, 8>([&](auto b_) {
      constexpr int b = decltype(b_)::value;
      if ((t >> b) & 1) mt = mulp(mt, fr_load(pw + b));
    });
  }
  h = mulp(h, mt);
  lds_put(l, t, h);
  __syncthreads();
  for (int s = kScanThreads / 2; s > 0; s >>= 1) {
    if (t < s) {
      h = fr_add(h, lds_get(l, t + s));
      lds_put(l, t, h);
    }
    __syncthreads();
  }
  if (t == 0) fr_store(partial + blk, h);
}

// partial[b] as sfx_reduce_kernel computes it, from contiguous runs (thread t
// owns the kPer elements the down-sweep gives it): the local Horner value L_t
// goes to lt for the down-sweep (which then skips its own), then the block
// suffix over the L_t gives the block's value at thread 0.
__global__ void __launch_bounds__(kScanThreads) sfx_run_reduce_kernel(const fr* a, uint64_t n, const fr* pw,
                                                                      fr* partial, fr* lt, PwExt ext) {
  __shared__ uint4 l[2][kScanThreads];
  const int t = threadIdx.x;
  uint32_t blk = blockIdx.x;
  if (ext.count) {
    if (blk == 0) {
      if (t == 0) pw_extend(ext);
      return;
    }
    blk--;
  }
  const uint64_t tid = (uint64_t)blk * kScanThreads + t;
  const uint64_t base = tid * kPer;
  uint32_t m[9];
  unpack_ext(fr_load(pw).l, m);  // balanced digits
  fr e[kPer];
  static_for<0, kPer>([&](auto k_) {
    constexpr int k = decltype(k_)::value;
    e[k] = base + k < n ? fr_load(a + base + k) : fr_zero();
  });
  const fr h = horner29<kPer>(e, m);
  fr_store(lt + tid, h);
  const fr v = block_suffix<kLgPer>(h, l, pw);
  if (t == 0) fr_store(partial + blk, v);
}

// out[i] = s_{i+1} within each block given carry[b] = s at the block's end
// (nullptr: zero); total (block 0) = s_0.  out may equal a.
// LT: the local Horner values come from sfx_run_reduce_kernel (lt); the
// elements are then loaded only for the rescan, in two halves, so that fewer
// values are live (more waves per SIMD).  Otherwise computed here.
template <int LG, bool LT>
__global__ void __launch_bounds__(kScanThreads) sfx_down_kernel(const fr* a, fr* out, uint64_t n, const fr* pw,
                                                                 const fr* carry, fr* total, const fr* lt) {
  constexpr int kPer = ScanGeom<LG>::kPer, kLgSeg = ScanGeom<LG>::kLgSeg, kLgPer = LG;
  __shared__ uint4 l[2][kScanThreads];
  const int t = threadIdx.x;
  const uint64_t base = ((uint64_t)blockIdx.x << kLgSeg) + (uint64_t)t * kPer;
  uint32_t m[9];
  unpack_ext(fr_load(pw).l, m);  // balanced digits
  // local Horner: L = sum_k e_k m^k
  fr h;
  if constexpr (LT) {
    h = fr_load(lt + ((uint64_t)blockIdx.x * kScanThreads + t));
  } else {
    fr e[kPer];
    static_for<0, kPer>([&](auto k_) {
      constexpr int k = decltype(k_)::value;
      e[k] = base + k < n ? fr_load(a + base + k) : fr_zero();
    });
    h = horner29<kPer>(e, m);
  }
  // fold the block carry into the last thread: s(start of thread 255) = L + m^kPer K
  fr K = carry ? fr_load(carry + blockIdx.x) : fr_zero();
  if (t == kScanThreads - 1) h = fr_add(h, mulp(K, fr_load(pw + kLgPer)));
  fr v = block_suffix<LG>(h, l, pw);  // s at the start of this thread's run
  // carry-in at this thread's end = s at the start of thread t+1
  fr cin = t + 1 < kScanThreads ? lds_get(l, t + 1) : K;
  if (total && blockIdx.x == 0 && t == 0) fr_store(total, v);
  // rescan: s_k = e_k + m s_{k+1}, out[base + k] = s_{k+1} (signed limbs,
  // canonical where stored); out may equal a: each half's elements are read
  // before its outputs are written, and the low half's outputs only overwrite
  // elements already read
  v9 sn;
  unpack_ext(cin.l, sn.l);
  constexpr int G = kPer > 1 ? 2 : 1, Q = kPer / G;
  static_for<0, G>([&](auto g_) {
    constexpr int top = kPer - 1 - decltype(g_)::value * Q;  // highest index of this half
    fr e[Q];
    static_for<0, Q>([&](auto j_) {
      constexpr int k = top - decltype(j_)::value;
      e[top - k] = k > 0 && base + k < n ? fr_load(a + base + k) : fr_zero();
    });
    static_for<0, Q>([&](auto j_) {
      constexpr int k = top - decltype(j_)::value;
      if (base + k < n) fr_store(out + base + k, k == kPer - 1 ? cin : v9_canonical(sn));
      if constexpr (k > 0) {
        v9 x;
        unpack_ext(e[top - k].l, x.l);
        sn = v9_add(v9_mul<false, false, false>(sn, m), x);
      }
    });
  });
}

__global__ void zero_kernel(fr* p, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) fr_store(p + i, fr_zero());
}

hipError_t fill_zero(fr* p, uint64_t n, hipStream_t s) {
  hipLaunchKernelGGL(zero_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

// pw[j] = m^(2^j) R' mod r, j < count; m = neg ? -(*src) : *src.  32 m (five
// doublings) is m R' in the stored form; squares stay in that form.  Level L
// of a scan (values combined 4096^L at a time) uses pw + 12 L: its multiplier
// is m^(4096^L), whose powers of two continue the same squaring chain.
// tab (evaluation, 128 threads): also tab[t] = m^t R' for t < 256, which the
// reduce multiplies each thread's value by; the second wave extends it by
// doubling (step j fills [2^j, 2^(j+1)) from pw[j]) while thread 0 raises the
// next square.
__global__ void sfx_powers_kernel(const fr* src, int neg, fr* pw, int count, fr* tab) {
  const int t = threadIdx.x;
  fr m;
  if (t == 0) {
    m = fr_load(src);
    if (neg) m = fr_neg(m);
    for (int d = 0; d < 5; d++) m = fr_add(m, m);
  }
  if (tab && t == 64) {
    fr one;
#pragma unroll
    for (int i = 0; i < 8; i++) one.l[i] = rp_one_limb(i);
    fr_store(tab, one);
  }
  for (int j = 0; j < count; j++) {
    if (t == 0) fr_store(pw + j, m);
    if (tab) __syncthreads();  // pw[j] and the table's entries below 2^j visible
    if (t == 0 && j + 1 < count) m = mulp(m, m);
    if (tab && j < 8 && t >= 64) {
      for (int i = t - 64; i < (1 << j); i += 64)
        fr_store(tab + (1 << j) + i, mulp(fr_load(tab + i), fr_load(pw + j)));
    }
  }
}

constexpr int kPw = kLgSeg + 1;

// levels of a suffix scan over n values, and the power-table entries they use
static int sfx_levels(uint64_t n) {
  int L = 1;
  while (n > (uint64_t)kSeg) {
    n = (n + kSeg - 1) / kSeg;
    L++;
  }
  return L;
}
static int sfx_pw_count(uint64_t n) { return kPw + kLgSeg * (sfx_levels(n) - 1); }
// the entries beyond level 0's, for the first reduce launch (none for one level)
static PwExt pw_ext(fr* pw, int np) { return np > kPw ? PwExt{pw, kPw, np} : PwExt{nullptr, 0, 0}; }

static uint64_t sfx_scratch_elems(uint64_t n) {
  uint64_t tot = (uint64_t)sfx_pw_count(n);
  while (n > (uint64_t)kSeg) {
    n = (n + kSeg - 1) / kSeg;
    tot += 2 * n + n * kScanThreads;  // block values, carries, local Horner values
  }
  return tot;
}

// suffix scan with powers table pw (this level's m): out[i] = s_{i+1} (if out),
// *total = s_0 (if total).  out == nullptr: reduction only.  A top level of at
// most 256 values runs one element per thread.
static hipError_t suffix_scan(const fr* a, fr* out, uint64_t n, const fr* pw, fr* total, fr* scratch, hipStream_t s,
                              PwExt ext = PwExt{nullptr, 0, 0}, const fr* mt = nullptr) {
  if (n <= (uint64_t)kSeg) {
    const bool tiny = n <= (uint64_t)kScanThreads;
    if (out) {
      if (tiny)
        hipLaunchKernelGGL((sfx_down_kernel<0, false>), dim3(1), dim3(kScanThreads), 0, s, a, out, n, pw, (const fr*)nullptr,
                           total, (const fr*)nullptr);
      else
        hipLaunchKernelGGL((sfx_down_kernel<kLgPer, false>), dim3(1), dim3(kScanThreads), 0, s, a, out, n, pw,
                           (const fr*)nullptr, total, (const fr*)nullptr);
    } else {
      // one block: the reduce kernel's tree value is s_0
      const PwExt none{nullptr, 0, 0};
      if (tiny)
        hipLaunchKernelGGL(sfx_reduce_kernel<0>, dim3(1), dim3(kScanThreads), 0, s, a, n, pw, total, none, mt);
      else
        hipLaunchKernelGGL(sfx_reduce_kernel<kLgPer>, dim3(1), dim3(kScanThreads), 0, s, a, n, pw, total, none, mt);
    }
    return hipGetLastError();
  }
  const uint64_t nb = (n + kSeg - 1) / kSeg;
  fr* part = scratch;
  fr* carry = scratch + nb;
  fr* lt = scratch + 2 * nb;
  // with a down-sweep, the reduce runs over the down-sweep's runs and keeps
  // their local Horner values; a reduction alone reads strided (coalesced)
  const unsigned grid = sgrid(n) + (ext.count ? 1 : 0);
  if (out)
    hipLaunchKernelGGL(sfx_run_reduce_kernel, dim3(grid), dim3(kScanThreads), 0, s, a, n, pw, part, lt, ext);
  else
    hipLaunchKernelGGL(sfx_reduce_kernel<kLgPer>, dim3(grid), dim3(kScanThreads), 0, s, a, n, pw, part, ext, mt);
  // carry[b] = s at the end of block b = sum_{b' > b} part[b'] (m^4096)^(b'-b-1)
  hipError_t e = suffix_scan(part, out ? carry : nullptr, nb, pw + kLgSeg, total, lt + nb * kScanThreads, s);
  if (e != hipSuccess || !out) return e;
  hipLaunchKernelGGL((sfx_down_kernel<kLgPer, true>), dim3(sgrid(n)), dim3(kScanThreads), 0, s, a, out, n, pw,
                     (const fr*)carry, (fr*)nullptr, (const fr*)lt);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// evaluation and division, latency form (kLatMin <= n <= kLatMax)
// ---------------------------------------------------------------------------
// A Fr product costs a wave about the same issue time whether or not other
// waves share its SIMD, so at these sizes the scan above is bound by the
// products its busiest SIMD issues in sequence: the powers chain, then the
// reduce (which waits for the powers), then the top level.  Here the powers
// are raised beside the data instead of before it:
//
//   launch 1  data blocks 1..G: thread t of block b owns the run of K
//             consecutive coefficients at (256 b + t) K and stores its
//             Horner value L = sum_k a_k m^k, which needs m alone;
//             table block 0 (on a CU of its own: G + 1 <= 256 blocks) raises
//             m: mu = m^K, its squares mu^(2^s), A_j = mu^j (j < 256) and
//             B_b = mu^(256 b) (b < G), by doubling steps that follow the
//             squaring chain;
//   evaluation (m = x)
//   launch 2  block b: P_b = B_b sum_t L_(256 b + t) A_t;
//   launch 3  y = sum_b P_b;
//   division by X + c (m = -c; s_i = a_i + m s_(i+1), q_i = s_(i+1));
//             launch 1 also keeps each run's upper-half Horner value H
//   launch 2  block b scans its 256 run values with mu: the run suffixes S_t
//             and the block value P_b = S_0;
//   launch 3  two threads per run: each block forms its carry
//             C_b = sum_j P_(b+1+j) B_j; the carry at the end of run t is
//             S_(t+1) + mu^(255 - t) C_b, the lower half starts from
//             H_t + m^(K/2) times it, and both halves rescan in place.
//
// The launch boundaries carry the tables and values between blocks; a
// release/acquire hand-off inside one launch costs more than a launch here
// (it writes back and invalidates the L2 of every XCD involved).
constexpr uint64_t kLatMin = 1ull << 12;
constexpr uint64_t kLatMax = 255ull * 256 * 28;  // runs of at most 28 coefficients (in registers)
constexpr uint32_t kLatMaxG = 255;
// workspace (Fr elements): A, B, P (256 each), mu^(2^s) (16), m^(K/2) (1),
// L (256 G; the division replaces it by the run suffixes S), and for the
// division the upper-half Horner values H (256 G)
constexpr uint64_t kWsA = 0, kWsB = 256, kWsP = 512, kWsPw = 768, kWsHalf = 784, kWsL = 800;

__device__ __forceinline__ fr rp_one_fr() {
  fr o;
#pragma unroll
  for (int i = 0; i < 8; i++) o.l[i] = rp_one_limb(i);
  return o;
}
__device__ __forceinline__ fr to_rp(fr v) {  // x R (stored) -> x R' (five doublings)
  for (int d = 0; d < 5; d++) v = fr_add(v, v);
  return v;
}
// the multiplier m R' of a scan: x, or -c for a division by X + c
__device__ __forceinline__ fr load_m(const fr* x, int neg) {
  fr m = fr_load(x);
  if (neg) m = fr_neg(m);
  return to_rp(m);
}

// Horner over a run held in registers, each element's digit split ordered
// after the previous step (an empty asm that reads the running value), so
// the loads issued before it are waited for one by one as the chain reaches
// them instead of all at once
// *upper (if not null): the value after the upper half, sum_(k >= K/2) a_k m^(k - K/2)
template <int K>
__device__ __forceinline__ fr horner29_inorder(fr (&v)[K], const uint32_t (&w9)[9], fr* upper) {
  v9 h;
  unpack_ext(v[K - 1].l, h.l);
  static_for<0, K - 1>([&](auto k_) {
    constexpr int k = K - 2 - decltype(k_)::value;
    fr e = v[k];
    asm volatile("" : "+v"(e.l[0]), "+v"(e.l[1]), "+v"(e.l[2]), "+v"(e.l[3]), "+v"(e.l[4]), "+v"(e.l[5]),
                 "+v"(e.l[6]), "+v"(e.l[7]) : "v"(h.l[0]));
    v9 xv;
    unpack_ext(e.l, xv.l);
    h = v9_add(v9_mul<false, false, false>(h, w9), xv);
    if constexpr (k == K / 2) {
      if (upper) *upper = v9_canonical(h);
    }
  });
  return v9_canonical(h);
}

// a thread's run of K coefficients in registers, loaded top first; zeros
// past n (only the last data block checks, a uniform branch)
template <int K>
__device__ __forceinline__ void load_run(const fr* a, uint64_t n, uint64_t base, bool full, fr (&v)[K]) {
  if (full) {
    static_for<0, K>([&](auto j_) {
      constexpr int k = K - 1 - decltype(j_)::value;
      v[k] = fr_load(a + base + k);
    });
  } else {
    static_for<0, K>([&](auto j_) {
      constexpr int k = K - 1 - decltype(j_)::value;
      v[k] = base + k < n ? fr_load(a + base + k) : fr_zero();
    });
  }
}

// launch 1; steps: doubling steps of the table block (8 for A, then B's);
// halves (division): also the upper-half Horner value of every run and m^(K/2)
template <int K>
__global__ void __launch_bounds__(kScanThreads) lat_runs_kernel(const fr* a, uint64_t n, const fr* x, int neg,
                                                                fr* ws, uint32_t G, int steps, int halves) {
  static_assert(K % 2 == 0, "runs split in two halves");
  const int t = threadIdx.x;
  if (blockIdx.x == 0) {
    // ---- table block: wave 0 squares mu, waves 1..3 extend A, then B ----
    __shared__ fr pw_s[16];
    __shared__ fr tab[2][256];
    const int w = t >> 6;
    if (w == 0) {
      const fr xp = load_m(x, neg);
      fr mu = xp;  // m^K R', left to right; K is even, so m^(K/2) precedes the last squaring
      for (int bit = 30 - __builtin_clz(K); bit >= 0; bit--) {
        if (bit == 0 && t == 0) fr_store(ws + kWsHalf, mu);
        mu = mulp(mu, mu);
        if ((K >> bit) & 1) mu = mulp(mu, xp);
      }
      if (t == 0) pw_s[0] = mu;
    }
    tab[0][t] = rp_one_fr();
    tab[1][t] = rp_one_fr();
    __syncthreads();
    // step s: entries [2^s', 2^(s'+1)) of A (s < 8) or of B (s >= 8),
    // s' = s mod 8, at most 128 per step, over the 192 lanes of waves 1..3
    for (int s = 0; s < steps; s++) {
      const fr p = pw_s[s];  // mu^(2^s) R'
      if (w == 0) {
        if (s + 1 < steps) {
          const fr q = mulp(p, p);
          if (t == 0) pw_s[s + 1] = q;
        }
      } else {
        const int tb = s < 8 ? 0 : 1, ss = s & 7;
        const int j = (1 << ss) + t - 64;
        if (t - 64 < (1 << ss) && (tb == 0 || j < (int)G)) tab[tb][j] = mulp(tab[tb][j - (1 << ss)], p);
      }
      __syncthreads();
    }
    fr_store(ws + kWsA + t, tab[0][t]);
    if (t < (int)G) fr_store(ws + kWsB + t, tab[1][t]);
    if (t < steps) fr_store(ws + kWsPw + t, pw_s[t]);
    return;
  }
  // ---- data block b: the Horner value of each thread's run ----
  const uint32_t b = blockIdx.x - 1;
  uint32_t w9[9];
  unpack_ext(load_m(x, neg).l, w9);  // balanced digits of m R'
  const uint64_t base = ((uint64_t)b * kScanThreads + t) * K;
  const bool full = (uint64_t)(b + 1) * kScanThreads * K <= n;
  fr v[K];
  load_run<K>(a, n, base, full, v);
  // the Horner chain starts after every load is issued (w9 passes through)
  asm volatile("" : "+v"(w9[0]), "+v"(w9[1]), "+v"(w9[2]), "+v"(w9[3]), "+v"(w9[4]), "+v"(w9[5]), "+v"(w9[6]),
               "+v"(w9[7]), "+v"(w9[8])::"memory");
  fr upper;
  const fr L = horner29_inorder<K>(v, w9, halves ? &upper : nullptr);
  const uint64_t r = (uint64_t)b * kScanThreads + t;
  fr_store(ws + kWsL + r, L);
  if (halves) fr_store(ws + kWsL + (uint64_t)kScanThreads * G + r, upper);
}

// evaluation, launch 2: P_b = B_b sum_t L_(256 b + t) A_t
__global__ void __launch_bounds__(kScanThreads) lat_combine_kernel(fr* ws) {
  __shared__ uint4 l[2][kScanThreads];
  const int t = threadIdx.x;
  const uint32_t b = blockIdx.x;
  fr c = mulp(fr_load(ws + kWsL + (uint64_t)b * kScanThreads + t), fr_load(ws + kWsA + t));
  lds_put(l, t, c);
  __syncthreads();
  for (int s = kScanThreads / 2; s > 0; s >>= 1) {
    if (t < s) {
      c = fr_add(c, lds_get(l, t + s));
      lds_put(l, t, c);
    }
    __syncthreads();
  }
  if (t == 0) fr_store(ws + kWsP + b, mulp(c, fr_load(ws + kWsB + b)));
}

// evaluation, launch 3: y = sum of the G block values
__global__ void __launch_bounds__(kScanThreads) eval_lat_sum_kernel(const fr* ws, uint32_t G, fr* y) {
  __shared__ uint4 l[2][kScanThreads];
  const int t = threadIdx.x;
  fr v = t < (int)G ? fr_load(ws + kWsP + t) : fr_zero();
  lds_put(l, t, v);
  __syncthreads();
  for (int s = kScanThreads / 2; s > 0; s >>= 1) {
    if (t < s) {
      v = fr_add(v, lds_get(l, t + s));
      lds_put(l, t, v);
    }
    __syncthreads();
  }
  if (t == 0) fr_store(y, v);
}

// division, launch 2: the run suffixes of block b, S_t = sum_(t' >= t) L_t' mu^(t' - t)
// (S_t replaces L_t), and the block value P_b = S_0
__global__ void __launch_bounds__(kScanThreads) div_lat_suffix_kernel(fr* ws) {
  __shared__ uint4 l[2][kScanThreads];
  const int t = threadIdx.x;
  const uint64_t r = (uint64_t)blockIdx.x * kScanThreads + t;
  fr v = fr_load(ws + kWsL + r);
  fr pw[8];  // mu^(2^s), loaded once before the levels
#pragma unroll
  for (int s = 0; s < 8; s++) pw[s] = fr_load(ws + kWsPw + s);
  lds_put(l, t, v);
  __syncthreads();
#pragma unroll
  for (int s = 0; s < 8; s++) {
    const int o = 1 << s;
    const fr other = t + o < kScanThreads ? lds_get(l, t + o) : fr_zero();
    __syncthreads();
    if (t + o < kScanThreads) v = fr_add(v, mulp(other, pw[s]));
    lds_put(l, t, v);
    __syncthreads();
  }
  fr_store(ws + kWsL + r, v);
  if (t == 0) fr_store(ws + kWsP + blockIdx.x, v);
}

// division, launch 3: two threads per run, each rescanning half of it in
// place.  Launch block d covers runs [(d mod 2) 128, +128) of launch-1 block
// b = d / 2.  C_b = s at the end of block b = sum_(j >= 0) P_(b+1+j) B_j (one
// product per thread and a sum); the carry at the end of run t is
// S_(t+1) + mu^(255 - t) C_b; the lower half starts from
// s at the middle = H_t + m^(K/2) (that carry).  q_i = s_(i+1),
// s_i = a_i + m s_(i+1).  Block 0 also writes the remainder s_0 = P_0 + M C_0.
// Every thread reads its half-run before writing it, so q may be a.
template <int K>
__global__ void __launch_bounds__(kScanThreads) div_lat_down_kernel(const fr* a, fr* q, uint64_t n, const fr* c,
                                                                    fr* ws, uint32_t G, fr* rem) {
  constexpr int H = K / 2;
  __shared__ uint4 l[2][kScanThreads];
  const int tid = threadIdx.x;
  const uint32_t b = blockIdx.x >> 1;
  const int t = (int)((blockIdx.x & 1) * 128 + (tid >> 1));
  const int upper = tid & 1;
  const uint64_t r = (uint64_t)b * kScanThreads + t;
  const uint64_t base = r * K + (upper ? H : 0);
  fr v[H];
  load_run<H>(a, n, base, (uint64_t)(b + 1) * kScanThreads * K <= n, v);
  fr cb = b + 1 + tid < G ? mulp(fr_load(ws + kWsP + b + 1 + tid), fr_load(ws + kWsB + tid)) : fr_zero();
  lds_put(l, tid, cb);
  __syncthreads();
  for (int s = kScanThreads / 2; s > 0; s >>= 1) {
    if (tid < s) {
      cb = fr_add(cb, lds_get(l, tid + s));
      lds_put(l, tid, cb);
    }
    __syncthreads();
  }
  cb = lds_get(l, 0);
  if (blockIdx.x == 0 && tid == 0 && rem) {
    const fr p0 = fr_load(ws + kWsP);
    fr_store(rem, G > 1 ? fr_add(p0, mulp(cb, fr_load(ws + kWsB + 1))) : p0);
  }
  fr cin = mulp(cb, fr_load(ws + kWsA + (kScanThreads - 1 - t)));
  if (t + 1 < kScanThreads) cin = fr_add(cin, fr_load(ws + kWsL + r + 1));
  if (!upper) cin = fr_add(fr_load(ws + kWsL + (uint64_t)kScanThreads * G + r), mulp(cin, fr_load(ws + kWsHalf)));
  uint32_t w9[9];
  unpack_ext(load_m(c, 1).l, w9);
  v9 sn;
  unpack_ext(cin.l, sn.l);
  static_for<0, H>([&](auto j_) {
    constexpr int k = H - 1 - decltype(j_)::value;
    if (base + k < n) fr_store(q + base + k, k == H - 1 ? cin : v9_canonical(sn));
    if constexpr (k > 0) {
      v9 xv;
      unpack_ext(v[k].l, xv.l);
      sn = v9_add(v9_mul<false, false, false>(sn, w9), xv);
    }
  });
}

// run length K: the smallest multiple of 4 that keeps G <= 255 data blocks
static uint32_t lat_run(uint64_t n) {
  const uint64_t per4 = (uint64_t)kLatMaxG * kScanThreads * 4;
  return (uint32_t)(4 * ((n + per4 - 1) / per4));
}
static uint32_t lat_blocks(uint64_t n, uint32_t K) {
  return (uint32_t)((n + (uint64_t)kScanThreads * K - 1) / ((uint64_t)kScanThreads * K));
}
static uint64_t lat_ws_elems(uint64_t n) { return kWsL + 2ull * kScanThreads * lat_blocks(n, lat_run(n)); }

template <template <int> class F, typename... Args>
static void lat_dispatch(uint32_t K, Args... args) {
  switch (K) {
    case 4: F<4>::run(args...); break;
    case 8: F<8>::run(args...); break;
    case 12: F<12>::run(args...); break;
    case 16: F<16>::run(args...); break;
    case 20: F<20>::run(args...); break;
    case 24: F<24>::run(args...); break;
    default: F<28>::run(args...); break;
  }
}
template <int K>
struct LatRuns {
  static void run(dim3 g, hipStream_t s, const fr* a, uint64_t n, const fr* x, int neg, fr* ws, uint32_t G,
                  int steps, int halves) {
    hipLaunchKernelGGL(lat_runs_kernel<K>, g, dim3(kScanThreads), 0, s, a, n, x, neg, ws, G, steps, halves);
  }
};
template <int K>
struct LatDown {
  static void run(dim3 g, hipStream_t s, const fr* a, fr* q, uint64_t n, const fr* c, fr* ws, uint32_t G, fr* rem) {
    hipLaunchKernelGGL(div_lat_down_kernel<K>, g, dim3(kScanThreads), 0, s, a, q, n, c, ws, G, rem);
  }
};

// B and A need 8 + ceil(log2 G) doubling steps
static int lat_steps(uint32_t G) { return 8 + (G > 1 ? 32 - __builtin_clz(G - 1) : 0); }

static hipError_t evaluate_lat(const fr* a, uint64_t n, const fr* x, fr* y, fr* ws, hipStream_t s) {
  const uint32_t K = lat_run(n), G = lat_blocks(n, K);
  lat_dispatch<LatRuns>(K, dim3(G + 1), s, a, n, x, 0, ws, G, lat_steps(G), 0);
  hipLaunchKernelGGL(lat_combine_kernel, dim3(G), dim3(kScanThreads), 0, s, ws);
  hipLaunchKernelGGL(eval_lat_sum_kernel, dim3(1), dim3(kScanThreads), 0, s, (const fr*)ws, G, y);
  return hipGetLastError();
}

// in place: inout <- quotient by (X + c), rem <- remainder (if not null)
static hipError_t div_lat(fr* inout, uint64_t n, const fr* c, fr* rem, fr* ws, hipStream_t s) {
  const uint32_t K = lat_run(n), G = lat_blocks(n, K);
  lat_dispatch<LatRuns>(K, dim3(G + 1), s, (const fr*)inout, n, c, 1, ws, G, lat_steps(G), 1);
  hipLaunchKernelGGL(div_lat_suffix_kernel, dim3(G), dim3(kScanThreads), 0, s, ws);
  lat_dispatch<LatDown>(K, dim3(2 * G), s, (const fr*)inout, inout, n, c, ws, G, rem);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// entry points used by the C-ABI
// ---------------------------------------------------------------------------

uint64_t scan_scratch_elems(uint64_t n) {
  // the evaluation's m^t table sits between the powers and the scan
  uint64_t a = pfx_scratch_elems(n), b = sfx_scratch_elems(n) + kScanThreads;
  if (n >= kLatMin && n <= kLatMax && b < lat_ws_elems(n)) b = lat_ws_elems(n);
  return a > b ? a : b;
}

uint64_t grand_product_scratch_elems(uint64_t n) { return pfx_scratch_elems(n); }

hipError_t grand_product(fr* inout, fr* incl_out, uint64_t n, fr* scratch, hipStream_t s) {
  return prefix_product(inout, inout, incl_out, n, nullptr, scratch, s);
}

hipError_t div_linear(fr* inout, uint64_t n, const fr* c, fr* rem, fr* scratch, hipStream_t s) {
  if (n >= kLatMin && n <= kLatMax) return div_lat(inout, n, c, rem, scratch, s);
  fr* pw = scratch;
  const int np = sfx_pw_count(n);
  hipLaunchKernelGGL(sfx_powers_kernel, dim3(1), dim3(64), 0, s, c, 1, pw, std::min(np, kPw), (fr*)nullptr);
  return suffix_scan(inout, inout, n, pw, rem, scratch + np, s, pw_ext(pw, np));
}

hipError_t evaluate(const fr* coeffs, uint64_t n, const fr* x, fr* y, fr* scratch, hipStream_t s) {
  if (n >= kLatMin && n <= kLatMax) return evaluate_lat(coeffs, n, x, y, scratch, s);
  fr* pw = scratch;
  const int np = sfx_pw_count(n);
  // the reduce's factors m^t (t < 256) as a table beside the powers
  fr* mt = scratch + np;
  hipLaunchKernelGGL(sfx_powers_kernel, dim3(1), dim3(128), 0, s, x, 0, pw, std::min(np, kPw), mt);
  return suffix_scan(coeffs, nullptr, n, pw, y, mt + kScanThreads, s, pw_ext(pw, np), mt);
}

// one scan for an evaluation + division pair: q(X) = (p(X) - p(z)) / (X - z)
// (q[n-1] = 0) and y = p(z); the suffix s_i = a_i + z s_{i+1} gives both
hipError_t eval_div(const fr* in, fr* q, uint64_t n, const fr* z, fr* y, fr* scratch, hipStream_t s) {
  fr* pw = scratch;
  const int np = sfx_pw_count(n);
  hipLaunchKernelGGL(sfx_powers_kernel, dim3(1), dim3(64), 0, s, z, 0, pw, std::min(np, kPw), (fr*)nullptr);
  return suffix_scan(in, q, n, pw, y, scratch + np, s, pw_ext(pw, np));
}

}  // namespace pl
