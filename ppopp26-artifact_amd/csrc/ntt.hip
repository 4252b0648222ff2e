// NTT host side: plans, domain tables, pass sequencing; small-size kernel
// and table builders.  The pass kernel itself is in ntt_pass.hpp.
#include <mutex>

#include "ntt_pass.hpp"

namespace pl {

// x times a premultiplication factor (pbd word of c R'), canonical: x c
__device__ __forceinline__ fr premul_one(const fr& x, const fr& f) {
  v9 a;
  unpack_ext(x.l, a.l);
  uint32_t w[9];
  pbd_unpack(f.l, w);
  return v9_canonical(v9_mul<false, false, false>(a, w));
}

// Sizes 2^1 .. 2^3: one thread per output, direct sum.  Blocks hold whole
// polynomials (blockDim is a multiple of 8), all reads precede the barrier,
// so in-place calls are safe.
__global__ void __launch_bounds__(256) ntt_tiny_kernel(PassArgs a, uint32_t npoly) {
  const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t lg = a.lgN;
  const uint32_t n = 1u << lg;
  const bool valid = tid < ((uint64_t)npoly << lg);
  const uint64_t poly = tid >> lg;
  const uint32_t k = (uint32_t)(tid & (n - 1));
  fr acc = fr_zero();
  if (valid) {
    const fr* src = a.in + poly * a.in_pstride;
    const fr* pm = a.premul ? a.premul + ((uint64_t)(poly & a.pm_mask) << lg) : nullptr;
    for (uint32_t j = 0; j < n; j++) {
      fr v = fr_load(src + j);
      if (pm) v = premul_one(v, fr_load(pm + j));
      const fr w = fr_load(a.inner + (((j * k) & (n - 1)) << (8 - lg)));
      acc = fr_add(acc, fr_mul(v, w));
    }
    if (a.do_scale) acc = fr_mul_u(acc, a.scale);
  }
  __syncthreads();
  if (valid) {
    const uint32_t lgI = a.lgI;
    fr* dst = a.out + ((poly >> lgI) << (lg + lgI)) + (poly & ((1ull << lgI) - 1));
    fr_store(dst + ((uint64_t)k << lgI), acc);
  }
}

// Single-pass sizes with a per-element factor table (small coset NTTs):
// out[poly][i] = in[poly * in_pstride + i] * premul[(poly & pm_mask) << lg | i]
__global__ void premul_kernel(const fr* in, fr* out, uint64_t in_pstride, const fr* premul, uint32_t pm_mask,
                              uint32_t lg, uint64_t total) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  uint64_t poly = i >> lg, pos = i & ((1ull << lg) - 1);
  fr v = premul_one(fr_load(in + poly * in_pstride + pos), fr_load(premul + ((uint64_t)(poly & pm_mask) << lg) + pos));
  fr_store(out + i, v);
}

// ---- table construction (domain creation; not on the hot path) ----------

// out[i] = base^(i << sh) * scale, i < n  (square and multiply)
__global__ void pow_direct_kernel(fr* out, uint64_t n, fr base, uint32_t sh, fr scale) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  fr_store(out + i, fr_mul(fr_pow_u64(base, i << sh), scale));
}

// out[i] = w^(e(i)) * scale with w^e = W0[e & 16383] * W1[e >> 14];
// e(i) = (i & (S-1)) * (i >> lgS) when lgS > 0 (pass twiddles), else i.
__global__ void pow_window_kernel(fr* out, uint64_t n, const fr* w0, const fr* w1, uint32_t lgS,
                                  fr scale, uint32_t do_scale) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t e = lgS ? (i & ((1ull << lgS) - 1)) * (i >> lgS) : i;
  fr v = fr_mul(fr_load(w0 + (e & 16383)), fr_load(w1 + (e >> 14)));
  if (do_scale) v = fr_mul(v, scale);
  fr_store(out + i, v);
}

// pass-kernel tables: w R' mod r -> pbd words (balanced digits), in place
__global__ void pbd_kernel(fr* t, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const fr v = fr_load(t + i);
  fr o;
  to_pbd(v.l, o.l);
  fr_store(t + i, o);
}

static inline fr to_dev(const host::fr64& h) {
  fr f;
  memcpy(f.l, h.l, 32);
  return f;
}

static inline unsigned grid_for(uint64_t n, unsigned block) { return (unsigned)((n + block - 1) / block); }

NttPlan plan_from_splits(int lg, const int* m, int P) {
  NttPlan p;
  p.lg = lg;
  p.P = P;
  int lgM = lg;
  for (int i = 0; i < P; i++) {
    p.m[i] = m[i];
    p.lgM[i] = lgM;
    p.lgS[i] = lgM - m[i];
    lgM -= m[i];
  }
  return p;
}

// Pass splits chosen by the autotuner (P == 0: the default split below).
static std::mutex g_plan_mu;
static int g_plan_P[29];
static int g_plan_m[29][PL_MAX_PASSES];

bool plan_valid(int lg, const int* m, int P) {
  if (lg < 4 || lg > 28 || P < 1 || P > PL_MAX_PASSES) return false;
  int sum = 0;
  for (int i = 0; i < P; i++) {
    if (m[i] < kLgPerT || m[i] > kLgPerT + (kLgPerT < 4 ? kLgPerT : 4)) return false;
    sum += m[i];
  }
  return sum == lg;
}

bool plan_set(int lg, const int* m, int P) {
  if (lg < 0 || lg > 28) return false;
  if (P != 0 && !plan_valid(lg, m, P)) return false;
  std::lock_guard<std::mutex> g(g_plan_mu);
  g_plan_P[lg] = P;
  for (int i = 0; i < P; i++) g_plan_m[lg][i] = m[i];
  return true;
}

// default: ceil(lg/8) passes of 4..8 points, sizes differing by <= 1, larger first
NttPlan default_plan(int lg) {
  NttPlan p;
  p.lg = lg;
  if (lg < 4) return p;
  const int MX = kLgPerT + (kLgPerT < 4 ? kLgPerT : 4);
  const int P = (lg + MX - 1) / MX;
  int m[PL_MAX_PASSES];
  const int base = lg / P, rem = lg % P;
  for (int i = 0; i < P; i++) m[i] = base + (i < rem ? 1 : 0);
  return plan_from_splits(lg, m, P);
}

NttPlan make_plan(int lg) {
  if (lg >= 4 && lg <= 28) {
    std::lock_guard<std::mutex> g(g_plan_mu);
    if (g_plan_P[lg]) return plan_from_splits(lg, g_plan_m[lg], g_plan_P[lg]);
  }
  return default_plan(lg);
}

// out[i] = base^e(i) * scale for the n entries of a table (see pow_window_kernel)
static hipError_t fill_pow(fr* out, uint64_t n, uint32_t lgS, const host::fr64& base, const host::fr64* scale,
                           fr* win, hipStream_t s) {
  fr one = to_dev(host::one());
  fr b = to_dev(base);
  hipLaunchKernelGGL(pow_direct_kernel, dim3(grid_for(16384, 256)), dim3(256), 0, s, win, (uint64_t)16384, b, 0u, one);
  hipLaunchKernelGGL(pow_direct_kernel, dim3(grid_for(16384, 256)), dim3(256), 0, s, win + 16384, (uint64_t)16384, b,
                     14u, one);
  fr sc = scale ? to_dev(*scale) : one;
  hipLaunchKernelGGL(pow_window_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, out, n, (const fr*)win,
                     (const fr*)(win + 16384), lgS, sc, scale ? 1u : 0u);
  return hipGetLastError();
}

hipError_t fill_pow_table(fr* out, uint64_t n, const host::fr64& base, const host::fr64& scale, hipStream_t s) {
  fr* win = nullptr;
  hipError_t e = hipMalloc((void**)&win, 2 * 16384 * sizeof(fr));
  if (e != hipSuccess) return e;
  e = fill_pow(out, n, 0, base, &scale, win, s);
  hipError_t e2 = hipStreamSynchronize(s);
  (void)hipFree(win);
  return e != hipSuccess ? e : e2;
}

hipError_t domain_build(Domain& d, int lg, hipStream_t s) {
  d.lg = lg;
  d.plan = make_plan(lg);
  const NttPlan& P = d.plan;
  size_t count = 4 * 256 + 2 * 16384;  // inner tables (both forms), w_N power windows
  for (int p = 0; p + 1 < P.P; p++) count += 2 * (1ull << P.lgM[p]);
  d.bytes = count * sizeof(fr);
  hipError_t e = hipMalloc((void**)&d.block, d.bytes);
  if (e != hipSuccess) return e;
  fr* cur = d.block;
  d.pw0 = cur; cur += 16384;
  d.pw1 = cur; cur += 16384;
  d.inner_fwd = cur; cur += 256;
  d.inner_inv = cur; cur += 256;
  d.inner29_fwd = cur; cur += 256;
  d.inner29_inv = cur; cur += 256;
  for (int p = 0; p + 1 < P.P; p++) {
    d.tw_fwd[p] = cur; cur += 1ull << P.lgM[p];
    d.tw_inv[p] = cur; cur += 1ull << P.lgM[p];
  }
  fr* win = nullptr;
  e = hipMalloc((void**)&win, 2 * 16384 * sizeof(fr));
  if (e != hipSuccess) return e;
  host::fr64 w256 = host::root_of_unity(8);
  host::fr64 w256i = host::inv(w256);
  d.n_inv = host::domain_size_inverse(lg);
  const host::fr64 c32 = host::from_u64(32);  // stored form -> w R' mod r
  const host::fr64 n_inv32 = host::mul(d.n_inv, c32);
  {
    // powers of w_N in two 2^14 windows (permutation ratio, coset tables)
    fr b = to_dev(host::root_of_unity(lg)), one = to_dev(host::one());
    hipLaunchKernelGGL(pow_direct_kernel, dim3(grid_for(16384, 256)), dim3(256), 0, s, d.pw0, (uint64_t)16384, b, 0u,
                       one);
    hipLaunchKernelGGL(pow_direct_kernel, dim3(grid_for(16384, 256)), dim3(256), 0, s, d.pw1, (uint64_t)16384, b, 14u,
                       one);
  }
  if ((e = fill_pow(d.inner_fwd, 256, 0, w256, nullptr, win, s)) != hipSuccess) goto done;
  if ((e = fill_pow(d.inner_inv, 256, 0, w256i, nullptr, win, s)) != hipSuccess) goto done;
  if ((e = fill_pow(d.inner29_fwd, 256, 0, w256, &c32, win, s)) != hipSuccess) goto done;
  if ((e = fill_pow(d.inner29_inv, 256, 0, w256i, &c32, win, s)) != hipSuccess) goto done;
  for (int p = 0; p + 1 < P.P; p++) {
    host::fr64 w = host::root_of_unity(P.lgM[p]);
    host::fr64 wi = host::inv(w);
    uint64_t n = 1ull << P.lgM[p];
    if ((e = fill_pow(d.tw_fwd[p], n, P.lgS[p], w, &c32, win, s)) != hipSuccess) goto done;
    // the inverse transform's 2^-lg rides on the first pass's twiddles
    if ((e = fill_pow(d.tw_inv[p], n, P.lgS[p], wi, p == 0 ? &n_inv32 : &c32, win, s)) != hipSuccess) goto done;
  }
  {
    // the pass kernels' tables (inner29_*, tw_*) are contiguous to the end
    const uint64_t n29 = (uint64_t)(cur - d.inner29_fwd);
    hipLaunchKernelGGL(pbd_kernel, dim3(grid_for(n29, 256)), dim3(256), 0, s, d.inner29_fwd, n29);
    if ((e = hipGetLastError()) != hipSuccess) goto done;
  }
  e = hipStreamSynchronize(s);
done:
  (void)hipFree(win);
  return e;
}

// cp[t << lg_n | i] = (shift * w_M^t)^i, M = 2^(lg_n + lg_blowup), as pbd words
// of the value times R' (the factor form of the pass kernels' product)
hipError_t coset_table_build(fr** cp, size_t* bytes, int lg_n, int lg_blowup, const host::fr64& shift,
                             hipStream_t s) {
  return coset_table_build_n(cp, bytes, lg_n, 1u << lg_blowup, shift, host::root_of_unity(lg_n + lg_blowup), s);
}

hipError_t coset_table_build_n(fr** cp, size_t* bytes, int lg_n, uint32_t B, const host::fr64& shift,
                               const host::fr64& step, hipStream_t s) {
  const uint64_t n = 1ull << lg_n;
  *bytes = (n * B) * sizeof(fr);
  hipError_t e = hipMalloc((void**)cp, *bytes);
  if (e != hipSuccess) return e;
  fr* win = nullptr;
  if ((e = hipMalloc((void**)&win, 2 * 16384 * sizeof(fr))) != hipSuccess) return e;
  const host::fr64 c32 = host::from_u64(32);  // stored form -> c R' mod r
  host::fr64 st = shift;
  for (uint64_t t = 0; t < B; t++) {
    if ((e = fill_pow(*cp + t * n, n, 0, st, &c32, win, s)) != hipSuccess) break;
    st = host::mul(st, step);
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL(pbd_kernel, dim3(grid_for(n * B, 256)), dim3(256), 0, s, *cp, n * B);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(win);
  return e;
}

// lg_blowup < 0: blowup 3, cosets shift * w_3N^t (N = 2^lg_n)
hipError_t lde_domain_build(LdeDomain& d, int lg_n, int lg_blowup, const host::fr64& shift, hipStream_t s) {
  d.lg_n = lg_n;
  d.lg_blowup = lg_blowup < 0 ? 0 : lg_blowup;
  d.blowup = lg_blowup < 0 ? 3 : 1 << lg_blowup;
  d.shift = shift;
  hipError_t e = domain_build(d.ntt, lg_n, s);
  if (e != hipSuccess) return e;
  if (lg_blowup < 0) return coset_table_build_n(&d.cp, &d.bytes, lg_n, 3, shift, host::root_of_unity_3n(lg_n), s);
  return coset_table_build(&d.cp, &d.bytes, lg_n, lg_blowup, shift, s);
}

__global__ void interleave3_kernel(const fr* __restrict__ in, fr* __restrict__ out, uint64_t n) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= 3 * n) return;
  const uint64_t j = k / 3, t = k - 3 * j;
  fr_store(out + k, fr_load(in + t * n + j));
}

hipError_t interleave3(const fr* in, fr* out, uint64_t n, hipStream_t s) {
  hipLaunchKernelGGL(interleave3_kernel, dim3(grid_for(3 * n, 256)), dim3(256), 0, s, in, out, n);
  return hipGetLastError();
}

// The reference's windowed power tables (parameters.cuh:7-33; the first
// 4 * 2^14 elements of its Params): w_{2^28}^i, w_{2^28}^(i 2^14), 7^i, 7^(i 2^14)
hipError_t params_fill(fr* out, hipStream_t s) {
  const fr one = to_dev(host::one());
  const fr w = to_dev(host::root_of_unity(28));
  const fr g = to_dev(host::from_u64(7));
  const uint64_t W = 16384;
  hipLaunchKernelGGL(pow_direct_kernel, dim3(grid_for(W, 256)), dim3(256), 0, s, out, W, w, 0u, one);
  hipLaunchKernelGGL(pow_direct_kernel, dim3(grid_for(W, 256)), dim3(256), 0, s, out + W, W, w, 14u, one);
  hipLaunchKernelGGL(pow_direct_kernel, dim3(grid_for(W, 256)), dim3(256), 0, s, out + 2 * W, W, g, 0u, one);
  hipLaunchKernelGGL(pow_direct_kernel, dim3(grid_for(W, 256)), dim3(256), 0, s, out + 3 * W, W, g, 14u, one);
  return hipGetLastError();
}

void lde_domain_free(LdeDomain& d) {
  domain_free(d.ntt);
  if (d.cp) (void)hipFree(d.cp);
  d.cp = nullptr;
}

void domain_free(Domain& d) {
  if (d.block) (void)hipFree(d.block);
  d.block = nullptr;
}

static void launch_pass_la(int la, bool inv, bool last, const PassArgs& a, hipStream_t s) {
  switch (la) {
    case 0: launch_pass_la0(inv, last, a, s); break;
    case 1: launch_pass_la1(inv, last, a, s); break;
    case 2: launch_pass_la2(inv, last, a, s); break;
    case 3: launch_pass_la3(inv, last, a, s); break;
    default: launch_pass_la4(inv, last, a, s); break;
  }
}

// Kernel arguments of pass p of the plan, reading `in` (poly stride
// in_pstride, optional premul table) and writing `out`.
static void fill_pass_args(const Domain& d, const NttCall& c, int p, const fr* in, uint64_t in_pstride,
                           const fr* premul, fr* out, PassArgs& a) {
  const NttPlan& P = d.plan;
  const int lg = d.lg;
  memset(&a, 0, sizeof(a));
  a.inner = c.inverse ? d.inner_inv : d.inner_fwd;
  a.inner29 = c.inverse ? d.inner29_inv : d.inner29_fwd;
  a.lgN = lg;
  a.lgI = c.lgI;
  a.pm_mask = c.pm_mask;
  a.scale = to_dev(d.n_inv);
  {
    const fr sc = to_dev(host::mul(d.n_inv, host::from_u64(32)));
    to_pbd(sc.l, a.scale29.l);
  }
  a.in_ext = (P.P == 0 || p == 0) ? 1 : 0;
  a.in = in;
  a.out = out;
  a.in_pstride = in_pstride;
  a.premul = premul;
  if (P.P == 0) {
    a.do_scale = c.inverse ? 1 : 0;
    return;
  }
  const bool last = p == P.P - 1;
  a.tw = last ? nullptr : (c.inverse ? d.tw_inv[p] : d.tw_fwd[p]);
  a.lgMp = P.lgM[p];
  a.lgS = P.lgS[p];
  a.nlines = (uint64_t)c.npoly << (lg - P.m[p]);
  a.do_scale = (c.inverse && P.P == 1) ? 1 : 0;
  if (last) {
    a.ndig = p;
    for (int q = 0; q < p; q++) a.lgrad[q] = P.m[q];
  }
}

// Pass p of the plan.  Intermediate passes keep the in-place layout, so in
// may equal out for them; the last pass writes natural order.
hipError_t ntt_launch_one(const Domain& d, const NttCall& c, int p, const fr* in, uint64_t in_pstride,
                          const fr* premul, fr* out, hipStream_t s) {
  const NttPlan& P = d.plan;
  const int lg = d.lg;
  if (premul && (P.P == 0 || p == P.P - 1)) {
    // the last-pass and tiny kernels carry no factor table (register budget):
    // scale the input into the scratch buffer first.  The scratch copy also
    // makes a shared-input call (in_pstride 0) safe when out aliases in.
    if (!c.tmp) return hipErrorInvalidValue;
    uint64_t total = (uint64_t)c.npoly << lg;
    hipLaunchKernelGGL(premul_kernel, dim3(grid_for(total, 256)), dim3(256), 0, s, in, c.tmp, in_pstride, premul,
                       c.pm_mask, (uint32_t)lg, total);
    in = c.tmp;
    in_pstride = 1ull << lg;
    premul = nullptr;
  }
  PassArgs a;
  fill_pass_args(d, c, p, in, in_pstride, premul, out, a);
  if (P.P == 0) {
    uint64_t total = (uint64_t)c.npoly << lg;
    hipLaunchKernelGGL(ntt_tiny_kernel, dim3(grid_for(total, 256)), dim3(256), 0, s, a, c.npoly);
    return hipGetLastError();
  }
  launch_pass_la(P.m[p] - kLgPerT, c.inverse, p == P.P - 1, a, s);
  return hipGetLastError();
}

// Pass p of a single-polynomial transform on `nlines` of its lines: the
// in/out pointers advanced by in_off/out_off elements (the pass twiddles
// follow the output column).  Pass 0's lines are columns j at stride
// N >> m[0], so a group of consecutive columns is in_off = out_off = j0; the
// last pass's line lp writes lp + (N >> m[P-1]) k and reads sub-row
// digit-reversed(lp), a bit permutation, so a group of lines g*W .. g*W+W-1
// (W a power of two) is out_off = g*W, in_off = pi(g*W) << m[P-1].
hipError_t ntt_launch_part(const Domain& d, const NttCall& c, int p, const fr* in, fr* out, uint64_t in_off,
                           uint64_t out_off, uint64_t nlines, hipStream_t s) {
  const NttPlan& P = d.plan;
  if (P.P < 2 || c.premul || c.npoly != 1 || p < 0 || p >= P.P) return hipErrorInvalidValue;
  PassArgs a;
  fill_pass_args(d, c, p, in + in_off, 1ull << d.lg, nullptr, out + out_off, a);
  if (a.tw) a.tw += out_off;
  a.nlines = nlines;
  launch_pass_la(P.m[p] - kLgPerT, c.inverse, p == P.P - 1, a, s);
  return hipGetLastError();
}

// the last pass's sub-row of line lp (the digit reversal of the earlier
// passes' radices, as the kernel undoes it)
uint64_t last_pass_row(const NttPlan& P, uint64_t lp) {
  uint64_t r = 0, q = lp;
  for (int p = 0; p + 1 < P.P; p++) {
    const int lr = P.m[p];
    r = (r << lr) | (q & ((1ull << lr) - 1));
    q >>= lr;
  }
  return r;
}

bool pass_args_regular(const Domain& d, const NttCall& c, int p, PassArgs& a) {
  const NttPlan& P = d.plan;
  if (P.P == 0 || (c.premul && P.P == 1) || p < 0 || p >= P.P) return false;
  const bool last = p == P.P - 1;
  fill_pass_args(d, c, p, p == 0 ? c.in : c.tmp, p == 0 ? c.in_pstride : (1ull << d.lg), p == 0 ? c.premul : nullptr,
                 last ? c.out : c.tmp, a);
  return true;
}

hipError_t ntt_launch(const Domain& d, const NttCall& c, hipStream_t s) {
  const int np = d.plan.P == 0 ? 1 : d.plan.P;
  const uint64_t N = 1ull << d.lg;
  for (int p = 0; p < np; p++) {
    const bool last = p == np - 1;
    hipError_t e = ntt_launch_one(d, c, p, p == 0 ? c.in : c.tmp, p == 0 ? c.in_pstride : N,
                                  p == 0 ? c.premul : nullptr, last ? c.out : c.tmp, s);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace pl
