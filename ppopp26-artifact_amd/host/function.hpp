// Operator API on SyncedMemory, mirroring the reference's
// PLONK/utils/function.cuh (names, argument meaning, buffer ownership).
// Every GPU operator runs through libpl's C-ABI (include/pl.h).
//
// Semantics kept from the reference:
//  * Ntt::forward returns a NEW device buffer (natural order in and out);
//    forward_ writes into `output`.  Intt likewise with w^-1 and 2^-lg.
//  * accumulate_mul_poly works in place and returns the same handle
//    (exclusive prefix product); accumulate_mul_poly_ also leaves the
//    inclusive products in `output`.
//  * evaluate returns a new one-element device buffer (0 for empty input).
//  * poly_div_poly divides by (X + c) in place and returns `divid`; c is one
//    element and is not modified.
//  * Element-wise ops run on the CPU when the first operand's head is on the
//    CPU and on the GPU otherwise (function.cu:5-111); x/0 = 0.
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "interface.hpp"

struct pl_domain;

SyncedMemory to_mont(SyncedMemory input);
SyncedMemory to_base(SyncedMemory input, hipStream_t stream = (hipStream_t)0);
SyncedMemory neg_mod(SyncedMemory input);
SyncedMemory inv_mod(SyncedMemory input);
SyncedMemory exp_mod(SyncedMemory input, int exp);

SyncedMemory add_mod(SyncedMemory input1, SyncedMemory input2, hipStream_t stream = (hipStream_t)0);
void add_mod_(SyncedMemory input1, SyncedMemory input2, hipStream_t stream = (hipStream_t)0);
SyncedMemory sub_mod(SyncedMemory input1, SyncedMemory input2, hipStream_t stream = (hipStream_t)0);
void sub_mod_(SyncedMemory input1, SyncedMemory input2, hipStream_t stream = (hipStream_t)0);
SyncedMemory mul_mod(SyncedMemory input1, SyncedMemory input2, hipStream_t stream = (hipStream_t)0);
void mul_mod_(SyncedMemory input1, SyncedMemory input2, hipStream_t stream = (hipStream_t)0);
SyncedMemory div_mod(SyncedMemory input1, SyncedMemory input2, hipStream_t stream = (hipStream_t)0);
void div_mod_(SyncedMemory input1, SyncedMemory input2, hipStream_t stream = (hipStream_t)0);

SyncedMemory add_mod_scalar(SyncedMemory input1, SyncedMemory input2, hipStream_t stream = (hipStream_t)0);
void add_mod_scalar_(SyncedMemory input1, SyncedMemory input2, hipStream_t stream = (hipStream_t)0);
SyncedMemory sub_mod_scalar(SyncedMemory input1, SyncedMemory input2, hipStream_t stream = (hipStream_t)0);
void sub_mod_scalar_(SyncedMemory input1, SyncedMemory input2, hipStream_t stream = (hipStream_t)0);
SyncedMemory mul_mod_scalar(SyncedMemory input1, SyncedMemory input2, hipStream_t stream = (hipStream_t)0);
void mul_mod_scalar_(SyncedMemory input1, SyncedMemory input2, hipStream_t stream = (hipStream_t)0);
SyncedMemory div_mod_scalar(SyncedMemory input1, SyncedMemory input2, hipStream_t stream = (hipStream_t)0);
void div_mod_scalar_(SyncedMemory input1, SyncedMemory input2, hipStream_t stream = (hipStream_t)0);

SyncedMemory evaluate(SyncedMemory poly, SyncedMemory x);
SyncedMemory poly_div_poly(SyncedMemory divid, SyncedMemory c);
SyncedMemory accumulate_mul_poly(SyncedMemory product, hipStream_t stream = (hipStream_t)0);
void accumulate_mul_poly_(SyncedMemory product, SyncedMemory output, hipStream_t stream = (hipStream_t)0);

SyncedMemory gen_sequence(uint64_t N, SyncedMemory x, hipStream_t stream = (hipStream_t)0);
void gen_sequence_(uint64_t N, SyncedMemory x, SyncedMemory y, hipStream_t stream = (hipStream_t)0);
SyncedMemory repeat_to_poly(SyncedMemory x, uint64_t N, hipStream_t stream = (hipStream_t)0);
void repeat_to_poly_(SyncedMemory x, SyncedMemory y, uint64_t N, hipStream_t stream = (hipStream_t)0);

SyncedMemory pad_poly(SyncedMemory x, uint64_t N, hipStream_t stream = (hipStream_t)0);
SyncedMemory repeat_zero(uint64_t N);
SyncedMemory cat(SyncedMemory a, SyncedMemory b);
SyncedMemory slice(SyncedMemory a, uint64_t len, bool forward);

namespace pl {
// per-size domain cache shared by copies of one Ntt / Intt
struct DomainCache {
  std::mutex mu;
  std::map<int, pl_domain*> doms;
  ~DomainCache();
  pl_domain* get(int lg);
};
}  // namespace pl

// [Ntt, function.cuh:81-100] domain_size: the largest log2 size to be used
// (the reference passes 32 and its tables then serve every size <= 2^28).
// Params holds the reference's windowed power tables in its layout
// (WINDOW_NUM x WINDOW_SIZE powers of w_{2^28}, then of g = 7), so code that
// reads ntt.Params.mutable_gpu_data() at those offsets keeps working.
class Ntt {
 public:
  SyncedMemory Params;
  explicit Ntt(int domain_size, hipStream_t stream = (hipStream_t)0);
  SyncedMemory forward(SyncedMemory input, hipStream_t stream = (hipStream_t)0);
  void forward_(SyncedMemory input, SyncedMemory output, hipStream_t stream = (hipStream_t)0);
  // [Ntt::stage_no_rotate, function.cu:312-314] forward DIT stages
  // [stage, stage + iterations) of a size-2^lg_domain_size NTT on a chunk
  // gathered with the driver's ct_stage_indices, in place, in gathered order
  // (pl_ntt_stage_no_rotate).  Forward only, as in the reference; an Intt
  // object throws.
  void stage_no_rotate(SyncedMemory input, int lg_domain_size, int iterations, int stage, int chunk_id,
                       hipStream_t stream = (hipStream_t)0);
  pl_domain* domain(int lg) { return cache_->get(lg); }

 protected:
  bool inverse_ = false;
  std::shared_ptr<pl::DomainCache> cache_;
};

// [Intt, function.cuh:102-114]
class Intt : public Ntt {
 public:
  explicit Intt(int domain_size, hipStream_t stream = (hipStream_t)0);
  void _forward_(SyncedMemory input, hipStream_t stream = (hipStream_t)0);  // in place
};

// [Ntt_coset / Intt_coset, function.cuh:116-142] The coset transform over
// g<w_N>, g = 7, N = 2^coset_size (pl_coset_*):
//   Ntt_coset   X[k] = sum_i x_i (g w_N^k)^i
//   Intt_coset  x_i = g^-i iNTT(X)_i
// init scales chunk chunk_id of the natural-order vector by the coset powers
// (g^(+-(chunk_id * C + i)), C = the chunk's length) in place; forward1 /
// forward2 / forward3 run the transform in three calls over the whole vector
// (the first pass with the coset powers fused, in place; the middle passes,
// in place; the last pass into `output`, natural order) and compose to the
// whole transform, which forward() runs in one call; init_and_forward writes
// the coset LDE of the 2^lg_domain_size coefficients in `input`, blowup
// 2^lambda, shift g, into `output` (natural order).  The stage / chunk_id
// arguments of forward1/2/3 are the reference's chunked schedule, which the
// whole-vector steps do not need (accepted and ignored).
namespace pl {
struct CosetDomains;
}
class Ntt_coset {
 public:
  SyncedMemory Params;
  int lg_domain_size;
  Ntt_coset(int domain_size, int coset_size);
  void init(SyncedMemory input, int lg_domain_size, int chunk_id, bool is_intt = false,
            hipStream_t stream = (hipStream_t)0);
  void init_and_forward(SyncedMemory input, SyncedMemory output, int lg_domain_size, int chunk_id, int lambda,
                        int stage, bool is_intt, hipStream_t stream = (hipStream_t)0);
  void forward1(SyncedMemory input, int lg_domain_size, int stage, int chunk_id, hipStream_t stream = (hipStream_t)0);
  void forward2(SyncedMemory input, int lg_domain_size, int stage, int chunk_id, hipStream_t stream = (hipStream_t)0);
  void forward3(SyncedMemory input, SyncedMemory output, int lg_domain_size, int stage, int chunk_id, int lambda,
                hipStream_t stream = (hipStream_t)0);
  SyncedMemory forward(SyncedMemory input, hipStream_t stream = (hipStream_t)0);

 protected:
  bool inverse_ = false;
  std::shared_ptr<pl::CosetDomains> doms_;
};

class Intt_coset : public Ntt_coset {
 public:
  Intt_coset(int domain_size, int coset_size, hipStream_t stream = (hipStream_t)0);
  void init(SyncedMemory input, int lg_domain_size, int chunk_id, bool is_intt = true,
            hipStream_t stream = (hipStream_t)0);
};

// host helpers (function.cuh:162-244)
uint32_t bit_rev(uint32_t i, uint32_t nbits);
void calculate_map1(uint32_t* map, int lg_N);

template <typename fr>
void bit_rev_step1_parallel(fr* d_out, fr* d_in, uint32_t* map, int lg_N, int num_threads, uint64_t idx) {
  const uint64_t N = uint64_t{1} << lg_N;
  if (num_threads < 1) num_threads = 1;
  std::vector<std::thread> ts;
  const uint64_t chunk = N / num_threads;
  for (int t = 0; t < num_threads; t++) {
    const uint64_t b = t * chunk, e = t == num_threads - 1 ? N : b + chunk;
    ts.emplace_back([=]() {
      for (uint64_t i = b; i < e; i++) d_out[map[idx + i]] = d_in[i];
    });
  }
  for (auto& th : ts) th.join();
}
