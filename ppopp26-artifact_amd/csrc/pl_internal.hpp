// Internal declarations shared by the HIP translation units of libpl.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fr.hpp"
#include "fr_host.hpp"
#include "../../include/pl.h"

namespace pl {

#define PL_MAX_PASSES 8

// Four-step DIF decomposition of a size-2^lg NTT into passes of 2^m[p]
// points per column (4 <= m <= 8).  Pass p works on sub-problems of size
// 2^lgM[p] (contiguous), doing column DFTs of size 2^m[p] at stride
// 2^lgS[p] = 2^(lgM[p] - m[p]); the last pass has lgS = 0 and scatters its
// output to the digit-reversed position.
struct NttPlan {
  int lg = 0;
  int P = 0;
  int m[PL_MAX_PASSES] = {};
  int lgM[PL_MAX_PASSES] = {};
  int lgS[PL_MAX_PASSES] = {};
};

NttPlan make_plan(int lg);

// Device-resident tables for NTT / iNTT of one size (the reference's Ntt /
// Intt Params, PLONK/utils/function.cu:278, 328; parameters.cuh:142-222).
struct Domain {
  int lg = 0;
  NttPlan plan;
  fr* tw_fwd[PL_MAX_PASSES] = {};  // pass twiddles, 2^lgM[p] entries, p < P-1
  fr* tw_inv[PL_MAX_PASSES] = {};  // inverse; pass 0 also carries 2^-lg when P >= 2
  fr* inner_fwd = nullptr;         // omega_256^e, e < 256
  fr* inner_inv = nullptr;
  fr* block = nullptr;             // single allocation holding all tables
  size_t bytes = 0;
  host::fr64 n_inv;
  int device = 0;
};

// Launch geometry for one call: a batch of `npoly` polynomials of size 2^lg.
struct NttCall {
  const fr* in = nullptr;      // first pass input
  fr* out = nullptr;           // final output
  fr* tmp = nullptr;           // npoly * 2^lg scratch (intermediate passes)
  uint64_t in_pstride = 0;     // element stride between input polys (0: shared)
  uint32_t npoly = 1;
  const fr* premul = nullptr;  // per-input-element factor table [(poly & pm_mask) << lg | pos]
  uint32_t pm_mask = 0;
  uint32_t lgI = 0;            // output interleave: out[(pos << lgI) + (poly & (2^lgI-1))]
  bool inverse = false;
};

hipError_t ntt_launch(const Domain& d, const NttCall& c, hipStream_t s);
hipError_t ntt_launch_one(const Domain& d, const NttCall& c, int p, const fr* in, uint64_t in_pstride,
                          const fr* premul, fr* out, hipStream_t s);
hipError_t domain_build(Domain& d, int lg, hipStream_t s);
void domain_free(Domain& d);

// element-wise helper kernels (ops.hip)
hipError_t fill_pow_table(fr* out, uint64_t n, const host::fr64& base, const host::fr64& scale, hipStream_t s);

}  // namespace pl
