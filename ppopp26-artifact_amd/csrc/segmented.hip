// The segmented NTT's per-chunk step: Ntt::stage_no_rotate.
//
// [Ntt::stage_no_rotate, PLONK/utils/function.cu:312-314 ->
//  ntt_zkp_stage_no_rotate, zksnark_ntt.cu:194-232, 440-454 ->
//  ntt_stage_no_rotate, ntt_kernel/ntt.cuh:642-665 ->
//  CTkernel_no_rotate, ct_mixed_radix_wide.cuh:884-974 ->
//  _CT_NTT_without_rotate, ct_mixed_radix_wide.cuh:610-751]
//
// A chunk of C elements gathered by the caller (the driver's
// ct_stage_indices, ntt.cu:168-178) holds C/2 element pairs (2t, 2t+1).
// One call runs DIT stages [stage, stage + it) of a size-2^lg_domain forward
// NTT on it and leaves the result in the gathered order (the caller's scatter
// applies rotate_stage_output_index, ntt.cu:190-195).  Thread t of the chunk
// (global number g = chunk_id * C/2 + t, 32-bit as in the reference):
//   * stage > 0: r0 *= f0, r1 *= f1 with, for j = 2 (g mod 2^(it-1)) and
//     p = (g mod 2^(stage+it-1)) >> (it-1),
//       f0 = w_{2^28}^(rev_{28-stage}(j) p), f1 = f0 w_{2^28}^(p 2^(27-stage))
//     except at (radix, stage) = (6,6), (6,12), (7,7), (8,8), where the
//     reference reads its w_{2^L}^(row col) tables (L = 12, 18, 14, 16; rows
//     of 2^radix): f_b = w_{2^L}^(row col) of flat index
//     p 2^radix + rev_{nb}(j + b), nb = max(L - lg_domain, 0) + it;
//   * a butterfly (r0 + r1, r0 - r1), then for s = 1 .. it-1 an exchange with
//     thread t xor 2^(s-1) of the same block (rank = t_block mod 2^s: ranks
//     below 2^(s-1) keep r0 and take the partner's r0 as r1, the others keep
//     r1 and take the partner's r1 as r0) and a butterfly with twiddle
//     w_{2^(s+1)}^rank on r1;
//   * blocks hold min(C/2, 2^(radix-1)) threads, radix = max(6, it).
// Every factor is a power of w_{2^28}, read from the Params windows
// (w^e = T0[e mod 2^14] T1[e >> 14]), so the arithmetic is exact and the
// result is bit-identical to the reference's.  Exchanges go through LDS.
//
// This step serves callers of the reference's chunked, host-streamed NTT;
// the library's own transforms (pl_ntt, pl_lde, batches) keep whole
// polynomials resident and do not use it.
#include "pl_internal.hpp"

namespace pl {

namespace {

constexpr int kMaxBlock = 512;  // 2^(10-1): radix 10

__device__ __forceinline__ uint32_t rev32(uint32_t i, uint32_t nbits) {
  return nbits ? (__builtin_bitreverse32(i) >> (32 - nbits)) : 0u;
}

// w_{2^28}^e, e < 2^28, from the Params windows T0 = params[0..2^14),
// T1 = params[2^14..2^15) (w^(i 2^14))
__device__ __forceinline__ fr w28(const fr* params, uint32_t e) {
  return fr_mul(fr_load(params + 16384 + (e >> 14)), fr_load(params + (e & 16383u)));
}

struct StageArgs {
  fr* chunk;
  const fr* params;
  uint32_t offset;  // chunk_id * C/2, 32-bit as the reference's index_t
  int it, stage, radix, lg_domain;
  int mode, L;      // factor rule, see the file header
};

__global__ void __launch_bounds__(kMaxBlock) stage_no_rotate_kernel(StageArgs a) {
  __shared__ fr sx[kMaxBlock];
  const uint32_t tl = threadIdx.x;
  const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + tl;
  const uint32_t g = (uint32_t)tid + a.offset;
  fr r0 = fr_load(a.chunk + 2 * tid);
  fr r1 = fr_load(a.chunk + 2 * tid + 1);
  if (a.mode != 0) {
    const int it = a.it, stage = a.stage;
    const uint32_t j = (g & ((1u << (it - 1)) - 1)) * 2;
    const uint32_t p = (g & ((1u << (stage + it - 1)) - 1)) >> (it - 1);
    fr f0, f1;
    if (a.mode == 1) {
      const uint32_t nbits = 28 - stage;
      f0 = w28(a.params, rev32(j, nbits) * p);
      f1 = fr_mul(w28(a.params, p << (nbits - 1)), f0);
    } else {
      const int sh = a.L - a.lg_domain > 0 ? a.L - a.lg_domain : 0;
      const uint32_t nb = (uint32_t)(sh + it);
      const uint64_t i0 = ((uint64_t)p << a.radix) + rev32(j, nb);
      const uint64_t i1 = ((uint64_t)p << a.radix) + rev32(j + 1, nb);
      const uint64_t m = (1ull << a.radix) - 1;
      // w_{2^L}^(row col) = w_{2^28}^(row col 2^(28-L) mod 2^28)
      const uint64_t emask = (1ull << 28) - 1;
      f0 = w28(a.params, (uint32_t)((((i0 >> a.radix) * (i0 & m)) << (28 - a.L)) & emask));
      f1 = w28(a.params, (uint32_t)((((i1 >> a.radix) * (i1 & m)) << (28 - a.L)) & emask));
    }
    r0 = fr_mul(r0, f0);
    r1 = fr_mul(r1, f1);
  }
  {
    const fr t = r1;
    r1 = fr_sub(r0, t);
    r0 = fr_add(r0, t);
  }
  for (int s = 1; s < a.it; s++) {
    const uint32_t laneMask = 1u << (s - 1), thrdMask = (1u << s) - 1;
    const uint32_t rank = tl & thrdMask;
    const bool pos = rank < laneMask;
    __syncthreads();  // the previous round's reads are done
    sx[tl] = pos ? r1 : r0;
    __syncthreads();
    const fr x = sx[tl ^ laneMask];
    if (pos)
      r1 = x;
    else
      r0 = x;
    // w_{2^(s+1)}^rank
    const fr t = fr_mul(w28(a.params, rank << (27 - s)), r1);
    r1 = fr_sub(r0, t);
    r0 = fr_add(r0, t);
  }
  fr_store(a.chunk + 2 * tid, r0);
  fr_store(a.chunk + 2 * tid + 1, r1);
}

}  // namespace

}  // namespace pl

// [Ntt::stage_no_rotate, function.cu:312-314]
extern "C" int pl_ntt_stage_no_rotate(void* d_chunk, uint64_t chunk_elems, int lg_domain, int iterations, int stage,
                                      uint64_t chunk_id, const void* d_params, pl_stream_t stream) {
  const int it = iterations;
  const int radix = it < 6 ? 6 : it;
  if (!d_chunk || !d_params || ((uintptr_t)d_chunk & 15) || ((uintptr_t)d_params & 15)) return PL_ERR_ARG;
  if (it < 1 || !(radix == 6 || radix == 7 || radix == 8 || radix == 10)) return PL_ERR_UNSUPPORTED;
  if (lg_domain < 1 || lg_domain > 28 || stage < 0 || stage + it > lg_domain) return PL_ERR_ARG;
  if (chunk_elems < 2 || (chunk_elems & (chunk_elems - 1)) || chunk_elems < (1ull << it)) return PL_ERR_ARG;
  const uint64_t nthreads = chunk_elems / 2;
  if (nthreads > (1ull << 31)) return PL_ERR_ARG;
  const uint64_t block = nthreads < (1ull << (radix - 1)) ? nthreads : (1ull << (radix - 1));
  pl::StageArgs a;
  a.chunk = (pl::fr*)d_chunk;
  a.params = (const pl::fr*)d_params;
  a.offset = (uint32_t)(chunk_id * nthreads);
  a.it = it;
  a.stage = stage;
  a.radix = radix;
  a.lg_domain = lg_domain;
  a.mode = stage == 0 ? 0 : 1;
  a.L = 0;
  if (stage != 0) {
    if (radix == 6 && stage == 6) { a.mode = 2; a.L = 12; }
    if (radix == 6 && stage == 12) { a.mode = 2; a.L = 18; }
    if (radix == 7 && stage == 7) { a.mode = 2; a.L = 14; }
    if (radix == 8 && stage == 8) { a.mode = 2; a.L = 16; }
  }
  hipLaunchKernelGGL(pl::stage_no_rotate_kernel, dim3((unsigned)(nthreads / block)), dim3((unsigned)block), 0,
                     (hipStream_t)stream, a);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? PL_OK : pl::hip_status(e);
}
