// Baseline transforms for the CLIs' --bench-baseline / --bench-naive modes
// (the reference's standard4_ntt, ntt.cu:1039-1138, and naive_lde,
// lde.cu:1367-1490): a bit-reversal kernel and one radix-2 butterfly launch
// per stage, twiddles from the domain's power windows.  Same results as the
// fast path; kept only as the workload baseline.
#include "pl_internal.hpp"

namespace pl {

__global__ void bitrev_copy_kernel(const fr* in, fr* out, uint32_t lg) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (1ull << lg)) return;
  uint64_t r = __builtin_bitreverse64(i) >> (64 - lg);
  fr_store(out + r, fr_load(in + i));
}

// stage s of a DIT NTT on bit-reversed data: pairs (i0, i0 + 2^s)
__global__ void radix2_stage_kernel(fr* d, uint32_t lg, uint32_t s, const fr* w0, const fr* w1) {
  uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (1ull << (lg - 1))) return;
  const uint64_t h = 1ull << s;
  const uint64_t j = t & (h - 1);
  const uint64_t i0 = ((t >> s) << (s + 1)) + j;
  const uint64_t e = j << (lg - s - 1);  // w_{2h}^j = w_N^(j N / 2h)
  fr w = fr_mul(fr_load(w0 + (e & 16383)), fr_load(w1 + (e >> 14)));
  fr a = fr_load(d + i0), b = fr_mul(fr_load(d + i0 + h), w);
  fr_store(d + i0, fr_add(a, b));
  fr_store(d + i0 + h, fr_sub(a, b));
}

__global__ void replicate_kernel(fr* d, uint64_t n, uint64_t m) {
  uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x + n;
  if (k >= m) return;
  fr_store(d + k, fr_load(d + (k & (n - 1))));
}

hipError_t ntt_naive(const Domain& dom, const fr* in, fr* out, hipStream_t s) {
  const uint32_t lg = dom.lg;
  const uint64_t n = 1ull << lg;
  hipLaunchKernelGGL(bitrev_copy_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, in, out, lg);
  for (uint32_t st = 0; st < lg; st++)
    hipLaunchKernelGGL(radix2_stage_kernel, dim3((unsigned)((n / 2 + 255) / 256)), dim3(256), 0, s, out, lg, st,
                       (const fr*)dom.pw0, (const fr*)dom.pw1);
  return hipGetLastError();
}

hipError_t lde_naive(const Domain& dom, const fr* in, fr* out, int lg_blowup, hipStream_t s) {
  hipError_t e = ntt_naive(dom, in, out, s);
  if (e != hipSuccess) return e;
  const uint64_t n = 1ull << dom.lg, m = n << lg_blowup;
  hipLaunchKernelGGL(replicate_kernel, dim3((unsigned)((m - n + 255) / 256)), dim3(256), 0, s, out, n, m);
  return hipGetLastError();
}

}  // namespace pl
