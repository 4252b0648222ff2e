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

// ---- the driver's index algebra (ntt.cu:168-195) -----------------------------
// The chunked flow's layout contract, so the formulas are the reference's:
// thread tid of a step (stage, it) owns the global element pair (i0, i1),
// i1 = i0 + 2^stage; a step's output element g goes to g with bits
// [stage, stage + it) rotated right by one.
__host__ __device__ inline void ct_pair(uint64_t tid, int stage, int it, uint64_t& i0, uint64_t& i1) {
  const uint64_t out = 1ull << (stage + it - 1);
  const uint64_t pos = (tid & (out - 1)) >> (it - 1);
  const uint64_t a = (tid & ~(out - 1)) + ((tid << stage) & (out - 1));
  i0 = a * 2 + pos;
  i1 = i0 + (1ull << stage);
}

__host__ __device__ inline uint64_t rotate_out(uint64_t g, int stage, int it) {
  const uint64_t mask = (1ull << (stage + it)) - (1ull << stage);
  const uint64_t r = g & mask;
  return (g & ~mask) | (((r >> 1) | (r << (it - 1))) & mask);
}

// rotate_out's inverse (bits [stage, stage + it) rotated left by one)
__host__ __device__ inline uint64_t rotate_back(uint64_t g, int stage, int it) {
  const uint64_t mask = (1ull << (stage + it)) - (1ull << stage);
  const uint64_t r = g & mask;
  return (g & ~mask) | (((r << 1) | (r >> (it - 1))) & mask);
}

namespace {

// chunk[2t + b] = src[i_b(tid0 + t)]  (permute_stage_input_chunk, ntt.cu:335-354)
__global__ void stage_gather_kernel(const fr* __restrict__ src, fr* __restrict__ chunk, uint64_t half, uint64_t tid0,
                                    int stage, int it) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= half) return;
  uint64_t i0, i1;
  ct_pair(tid0 + t, stage, it, i0, i1);
  fr_store(chunk + 2 * t, fr_load(src + i0));
  fr_store(chunk + 2 * t + 1, fr_load(src + i1));
}

// dst[rotate_out(i_b(tid0 + t)) - sub] = chunk[2t + b]
// (scatter_stage_output_chunk, ntt.cu:593-614; scatter_stage_kernel with
// sub = chunk * chunk_elems, ntt.cu:647-700)
__global__ void stage_scatter_kernel(const fr* __restrict__ chunk, fr* __restrict__ dst, uint64_t half,
                                     uint64_t tid0, int stage, int it, uint64_t sub) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= half) return;
  uint64_t i0, i1;
  ct_pair(tid0 + t, stage, it, i0, i1);
  fr_store(dst + (rotate_out(i0, stage, it) - sub), fr_load(chunk + 2 * t));
  fr_store(dst + (rotate_out(i1, stage, it) - sub), fr_load(chunk + 2 * t + 1));
}

// dst[2t + b] = src[rotate_back(i_b) - sub], i_b the next step's pair: a
// chunk left in the previous step's rotated order regathered for the next
// step without a trip through the whole vector (transition_stage_kernel,
// ntt.cu:702-756)
__global__ void stage_transition_kernel(const fr* __restrict__ src, fr* __restrict__ dst, uint64_t half,
                                        uint64_t tid0, int prev_stage, int prev_it, int next_stage, int next_it,
                                        uint64_t sub) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= half) return;
  uint64_t i0, i1;
  ct_pair(tid0 + t, next_stage, next_it, i0, i1);
  fr_store(dst + 2 * t, fr_load(src + (rotate_back(i0, prev_stage, prev_it) - sub)));
  fr_store(dst + 2 * t + 1, fr_load(src + (rotate_back(i1, prev_stage, prev_it) - sub)));
}

// out[rev_lg(i)] = in[i]: coalesced reads, 32-byte scattered writes
// (the host bit reversal of our_ntt, bit_rev_step1_parallel +
// calculate_map1, function.cuh:221-244, function.cu:503-509)
__global__ void bit_reverse_kernel(const fr* __restrict__ in, fr* __restrict__ out, uint64_t n, int lg) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t r = lg ? (__builtin_bitreverse64(i) >> (64 - lg)) : 0;
  fr_store(out + r, fr_load(in + i));
}

inline dim3 grid_of(uint64_t n) { return dim3((unsigned)((n + 255) / 256)); }

bool valid_iterations(int it) { return it >= 1 && it <= 10 && it != 9; }

}  // namespace

// The reference's chunked NTT flow run on the device: bit reversal, then
// pass 1 per chunk of C = 2^(lg_i + lg_k) consecutive elements (two
// stage_no_rotate steps of lg_i/2 and lg_i - lg_i/2 stages with the
// transition between them and a chunk-local scatter, run_first_ntt_pass,
// ntt.cu:801-859), then pass 2 per chunk (gather for stage lg_i, one step of
// lg_n - lg_i stages, scatter into the output,
// run_second_ntt_pass_async_scatter, ntt.cu:861-977).  Every permutation is
// a device kernel; the vector stays in HBM.
hipError_t ntt_segmented(fr* inout, int lg_n, int lg_i, int lg_k, const fr* params, fr* work, hipStream_t s);

}  // namespace pl

// [Ntt::stage_no_rotate, function.cu:312-314]
namespace pl {

static hipError_t stage_launch(fr* chunk, uint64_t chunk_elems, int lg_domain, int it, int stage, uint64_t chunk_id,
                               const fr* params, hipStream_t s) {
  const int radix = it < 6 ? 6 : it;
  const uint64_t nthreads = chunk_elems / 2;
  const uint64_t block = nthreads < (1ull << (radix - 1)) ? nthreads : (1ull << (radix - 1));
  StageArgs a;
  a.chunk = chunk;
  a.params = params;
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
  hipLaunchKernelGGL(stage_no_rotate_kernel, dim3((unsigned)(nthreads / block)), dim3((unsigned)block), 0, s, a);
  return hipGetLastError();
}

hipError_t ntt_segmented(fr* inout, int lg_n, int lg_i, int lg_k, const fr* params, fr* work, hipStream_t s) {
  const uint64_t N = 1ull << lg_n, C = 1ull << (lg_i + lg_k), half = C / 2, chunks = N / C;
  const int first = lg_i / 2, second = lg_i - first, it2 = lg_n - lg_i;
  fr* W = work;        // the bit-reversed vector, then pass 1's output
  fr* S = work + N;    // one chunk of scratch
  hipLaunchKernelGGL(bit_reverse_kernel, grid_of(N), dim3(256), 0, s, (const fr*)inout, W, N, lg_n);
  hipError_t e = hipGetLastError();
  for (uint64_t c = 0; c < chunks && e == hipSuccess; c++) {
    fr* chunk = W + c * C;
    e = stage_launch(chunk, C, lg_n, first, 0, c, params, s);
    if (e != hipSuccess) break;
    hipLaunchKernelGGL(stage_transition_kernel, grid_of(half), dim3(256), 0, s, (const fr*)chunk, S, half, c * half, 0,
                       first, first, second, c * C);
    if ((e = stage_launch(S, C, lg_n, second, first, c, params, s)) != hipSuccess) break;
    hipLaunchKernelGGL(stage_scatter_kernel, grid_of(half), dim3(256), 0, s, (const fr*)S, chunk, half, c * half, first,
                       second, c * C);
    e = hipGetLastError();
  }
  for (uint64_t c = 0; c < chunks && e == hipSuccess; c++) {
    hipLaunchKernelGGL(stage_gather_kernel, grid_of(half), dim3(256), 0, s, (const fr*)W, S, half, c * half, lg_i, it2);
    if ((e = stage_launch(S, C, lg_n, it2, lg_i, c, params, s)) != hipSuccess) break;
    hipLaunchKernelGGL(stage_scatter_kernel, grid_of(half), dim3(256), 0, s, (const fr*)S, inout, half, c * half, lg_i,
                       it2, (uint64_t)0);
    e = hipGetLastError();
  }
  return e;
}

}  // namespace pl

namespace {
bool al16(const void* p) { return ((uintptr_t)p & 15) == 0; }
int st_of(hipError_t e) { return e == hipSuccess ? PL_OK : pl::hip_status(e); }
}  // namespace

extern "C" {

// [Ntt::stage_no_rotate, function.cu:312-314]
int pl_ntt_stage_no_rotate(void* d_chunk, uint64_t chunk_elems, int lg_domain, int iterations, int stage,
                           uint64_t chunk_id, const void* d_params, pl_stream_t stream) {
  const int it = iterations;
  if (!d_chunk || !d_params || !al16(d_chunk) || !al16(d_params)) return PL_ERR_ARG;
  if (!pl::valid_iterations(it)) return PL_ERR_UNSUPPORTED;
  if (lg_domain < 1 || lg_domain > 28 || stage < 0 || stage + it > lg_domain) return PL_ERR_ARG;
  if (chunk_elems < 2 || (chunk_elems & (chunk_elems - 1)) || chunk_elems < (1ull << it)) return PL_ERR_ARG;
  if (chunk_elems / 2 > (1ull << 31)) return PL_ERR_ARG;
  return st_of(pl::stage_launch((pl::fr*)d_chunk, chunk_elems, lg_domain, it, stage, chunk_id,
                                (const pl::fr*)d_params, (hipStream_t)stream));
}

void pl_ct_stage_indices(uint64_t tid, int stage, int iterations, uint64_t* idx0, uint64_t* idx1) {
  uint64_t a = 0, b = 0;
  if (iterations >= 1 && stage >= 0 && stage + iterations <= 63) pl::ct_pair(tid, stage, iterations, a, b);
  if (idx0) *idx0 = a;
  if (idx1) *idx1 = b;
}

uint64_t pl_rotate_stage_output_index(uint64_t global_idx, int stage, int iterations) {
  if (iterations < 1 || stage < 0 || stage + iterations > 63) return global_idx;
  return pl::rotate_out(global_idx, stage, iterations);
}

// shared checks of the chunk permutations: a power-of-two chunk of at least
// one pair, 16-byte aligned buffers, a step inside 2^63
static int chunk_args(const void* a, const void* b, uint64_t chunk_elems, int stage, int it) {
  if (!a || !b || a == b || !al16(a) || !al16(b)) return PL_ERR_ARG;
  if (chunk_elems < 2 || (chunk_elems & (chunk_elems - 1)) || it < 1 || stage < 0 || stage + it > 40)
    return PL_ERR_ARG;
  return PL_OK;
}

int pl_stage_gather(const void* d_src, void* d_chunk, uint64_t chunk_elems, uint64_t chunk_id, int stage,
                    int iterations, pl_stream_t stream) {
  int st = chunk_args(d_src, d_chunk, chunk_elems, stage, iterations);
  if (st != PL_OK) return st;
  const uint64_t half = chunk_elems / 2;
  hipLaunchKernelGGL(pl::stage_gather_kernel, pl::grid_of(half), dim3(256), 0, (hipStream_t)stream,
                     (const pl::fr*)d_src, (pl::fr*)d_chunk, half, chunk_id * half, stage, iterations);
  return st_of(hipGetLastError());
}

int pl_stage_scatter(const void* d_chunk, void* d_dst, uint64_t chunk_elems, uint64_t chunk_id, int stage,
                     int iterations, int chunk_local, pl_stream_t stream) {
  int st = chunk_args(d_chunk, d_dst, chunk_elems, stage, iterations);
  if (st != PL_OK) return st;
  const uint64_t half = chunk_elems / 2;
  hipLaunchKernelGGL(pl::stage_scatter_kernel, pl::grid_of(half), dim3(256), 0, (hipStream_t)stream,
                     (const pl::fr*)d_chunk, (pl::fr*)d_dst, half, chunk_id * half, stage, iterations,
                     chunk_local ? chunk_id * chunk_elems : (uint64_t)0);
  return st_of(hipGetLastError());
}

int pl_stage_transition(const void* d_src, void* d_dst, uint64_t chunk_elems, uint64_t chunk_id, int prev_stage,
                        int prev_iterations, int next_stage, int next_iterations, pl_stream_t stream) {
  int st = chunk_args(d_src, d_dst, chunk_elems, next_stage, next_iterations);
  if (st != PL_OK) return st;
  if (prev_iterations < 1 || prev_stage < 0 || prev_stage + prev_iterations > 40) return PL_ERR_ARG;
  const uint64_t half = chunk_elems / 2;
  hipLaunchKernelGGL(pl::stage_transition_kernel, pl::grid_of(half), dim3(256), 0, (hipStream_t)stream,
                     (const pl::fr*)d_src, (pl::fr*)d_dst, half, chunk_id * half, prev_stage, prev_iterations,
                     next_stage, next_iterations, chunk_id * chunk_elems);
  return st_of(hipGetLastError());
}

int pl_bit_reverse(const void* d_in, void* d_out, int lg, pl_stream_t stream) {
  if (!d_in || !d_out || d_in == d_out || !al16(d_in) || !al16(d_out) || lg < 0 || lg > 32) return PL_ERR_ARG;
  const uint64_t n = 1ull << lg;
  hipLaunchKernelGGL(pl::bit_reverse_kernel, pl::grid_of(n), dim3(256), 0, (hipStream_t)stream,
                     (const pl::fr*)d_in, (pl::fr*)d_out, n, lg);
  return st_of(hipGetLastError());
}

// the reference's split constraints (ntt.cu:987-1001): both pass-1 steps and
// the pass-2 step are supported stage_no_rotate widths, and the chunk fits
// (and the pass-2 step's 2^(lg_n - lg_i) points lie in one chunk)
static bool segmented_valid(int lg_n, int lg_i, int lg_k) {
  if (lg_n < 2 || lg_n > 28 || lg_i < 2 || lg_i >= lg_n || lg_k < 0 || lg_i + lg_k > lg_n) return false;
  if (lg_n - lg_i > lg_i + lg_k) return false;
  return pl::valid_iterations(lg_i / 2) && pl::valid_iterations(lg_i - lg_i / 2) &&
         pl::valid_iterations(lg_n - lg_i);
}

size_t pl_ntt_segmented_work_bytes(int lg_n, int lg_i, int lg_k) {
  if (!segmented_valid(lg_n, lg_i, lg_k)) return 0;
  return ((1ull << lg_n) + (1ull << (lg_i + lg_k))) * sizeof(pl::fr);
}

int pl_ntt_segmented(void* d_inout, int lg_n, int lg_i, int lg_k, const void* d_params, void* d_work,
                     pl_stream_t stream) {
  if (!d_inout || !d_params || !al16(d_inout) || !al16(d_params) || !al16(d_work)) return PL_ERR_ARG;
  if (!segmented_valid(lg_n, lg_i, lg_k)) return PL_ERR_UNSUPPORTED;
  hipStream_t s = (hipStream_t)stream;
  void* w = d_work;
  if (!w) {
    hipError_t e = hipMallocAsync(&w, pl_ntt_segmented_work_bytes(lg_n, lg_i, lg_k), s);
    if (e != hipSuccess) return pl::hip_status(e);
  }
  hipError_t e = pl::ntt_segmented((pl::fr*)d_inout, lg_n, lg_i, lg_k, (const pl::fr*)d_params, (pl::fr*)w, s);
  if (!d_work) (void)hipFreeAsync(w, s);
  return st_of(e);
}

}  // extern "C"
