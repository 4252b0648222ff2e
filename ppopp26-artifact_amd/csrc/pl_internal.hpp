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
// points per column (3 <= m <= 6).  Pass p works on sub-problems of size
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

NttPlan make_plan(int lg);         // the tuned split when one is set, else default_plan
NttPlan default_plan(int lg);
NttPlan plan_from_splits(int lg, const int* m, int P);
bool plan_valid(int lg, const int* m, int P);
// set the split used by domains built from now on (P == 0: back to the default)
bool plan_set(int lg, const int* m, int P);

// Device-resident tables for NTT / iNTT of one size (the reference's Ntt /
// Intt Params, PLONK/utils/function.cu:278, 328; parameters.cuh:142-222).
struct Domain {
  int lg = 0;
  NttPlan plan;
  // The pass kernels' factors are held as w R' mod r (R' = 2^261, the radix-2^29
  // product's Montgomery constant; fr29.hpp), i.e. the stored form times 32.
  fr* tw_fwd[PL_MAX_PASSES] = {};  // pass twiddles, 2^lgM[p] entries, p < P-1
  fr* tw_inv[PL_MAX_PASSES] = {};  // inverse; pass 0 also carries 2^-lg when P >= 2
  fr* pw0 = nullptr;               // w_N^i = pw0[i & 16383] * pw1[i >> 14]
  fr* pw1 = nullptr;
  fr* inner_fwd = nullptr;         // omega_256^e, e < 256 (stored form; small-size kernel)
  fr* inner_inv = nullptr;
  fr* inner29_fwd = nullptr;       // the same times 32 (pass kernels)
  fr* inner29_inv = nullptr;
  fr* block = nullptr;             // single allocation holding all tables
  size_t bytes = 0;
  host::fr64 n_inv;
  int device = 0;
};

// Coset tables for the LDE: cp[t << lg_n | i] = (shift * w_M^t)^i.
struct LdeDomain {
  int lg_n = 0, lg_blowup = 0;
  int blowup = 0;           // cosets: 2^lg_blowup, or 3 (shifts shift * w_3N^t, output interleave 3)
  Domain ntt;               // size-2^lg_n transform
  fr* cp = nullptr;         // 2^(lg_n + lg_blowup) entries
  size_t bytes = 0;
  host::fr64 shift;
};

hipError_t lde_domain_build(LdeDomain& d, int lg_n, int lg_blowup, const host::fr64& shift, hipStream_t s);
// the coset factor table alone (*bytes = its size)
hipError_t coset_table_build(fr** cp, size_t* bytes, int lg_n, int lg_blowup, const host::fr64& shift,
                             hipStream_t s);
// B cosets: cp[t << lg_n | i] = (shift * step^t)^i as pbd words of the value times R'
hipError_t coset_table_build_n(fr** cp, size_t* bytes, int lg_n, uint32_t B, const host::fr64& shift,
                               const host::fr64& step, hipStream_t s);
// out[3 j + t] = in[t N + j], t < 3, j < N (the blowup-3 LDE's interleave)
hipError_t interleave3(const fr* in, fr* out, uint64_t n, hipStream_t s);
void lde_domain_free(LdeDomain& d);

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
// pass p of a one-polynomial call over nlines of its lines, pointers advanced
// by in_off / out_off elements (column groups of pass 0, line groups of the
// last pass; see ntt.hip)
hipError_t ntt_launch_part(const Domain& d, const NttCall& c, int p, const fr* in, fr* out, uint64_t in_off,
                           uint64_t out_off, uint64_t nlines, hipStream_t s);
uint64_t last_pass_row(const NttPlan& P, uint64_t lp);
// host-to-host NTT copy groups (2^lg_k) per size: tuned, else the default
constexpr int kHostLgKDefault = 2;
constexpr int kHostLgKMax = 20;
int host_groups_get(int lg);
bool host_groups_set(int lg, int lg_k);  // lg_k < 0: back to the default
hipError_t domain_build(Domain& d, int lg, hipStream_t s);
// maps a HIP error to a pl_status and records its text for pl_last_hip_error
int hip_status(hipError_t e);
hipError_t params_fill(fr* out, hipStream_t s);
hipError_t ntt_naive(const Domain& dom, const fr* in, fr* out, hipStream_t s);
hipError_t lde_naive(const Domain& dom, const fr* in, fr* out, int lg_blowup, hipStream_t s);
void domain_free(Domain& d);

// scans (scan.hip)
uint64_t scan_scratch_elems(uint64_t n);
// zero-fill by kernel (no memset nodes: callers may capture in hipGraphs)
hipError_t fill_zero(fr* p, uint64_t n, hipStream_t s);
hipError_t grand_product(fr* inout, fr* incl_out, uint64_t n, fr* scratch, hipStream_t s);
uint64_t grand_product_scratch_elems(uint64_t n);  // the prefix product's partials (a few per 4096 elements)
hipError_t div_linear(fr* inout, uint64_t n, const fr* c, fr* rem, fr* scratch, hipStream_t s);
hipError_t evaluate(const fr* coeffs, uint64_t n, const fr* x, fr* y, fr* scratch, hipStream_t s);
hipError_t eval_div(const fr* in, fr* q, uint64_t n, const fr* z, fr* y, fr* scratch, hipStream_t s);

// element-wise ops, batch inversion, permutation ratio (ops.hip)
uint64_t ops_scratch_elems(uint64_t n);
hipError_t vec_op(int op, const fr* a, const fr* b, fr* out, uint64_t n, fr* scratch, hipStream_t s);
hipError_t exp_mod(const fr* a, fr* out, uint64_t n, uint32_t e, hipStream_t s);
hipError_t gen_sequence(const fr* x, fr* out, uint64_t n, hipStream_t s);  // n <= 2^28
hipError_t repeat_to_poly(const fr* x, fr* out, uint64_t n, hipStream_t s);
hipError_t perm_ratio(const fr* const w[3], const fr* const sg[3], const fr* beta, const fr* gamma,
                      const fr* w0, const fr* w1, const host::fr64 k[3], fr* ratio, uint64_t n, fr* scratch,
                      hipStream_t s);
// z = the exclusive grand product of the permutation ratio (no ratio vector)
hipError_t perm_grand_product(const fr* const w[3], const fr* const sg[3], const fr* beta, const fr* gamma,
                              const fr* w0, const fr* w1, const host::fr64 k[3], fr* z, uint64_t n, fr* scratch,
                              hipStream_t s);

// synthetic inputs and digests (util.hip)
uint64_t uniform_work_bytes(uint64_t n);
hipError_t fill_uniform(fr* out, uint64_t n, uint64_t seed, uint32_t* work, hipStream_t s);
hipError_t digest(const fr* d, uint64_t n, uint64_t* out, hipStream_t s);
hipError_t mad_probe(double* tmad_per_s, hipStream_t s);

// element-wise helper kernels
hipError_t fill_pow_table(fr* out, uint64_t n, const host::fr64& base, const host::fr64& scale, hipStream_t s);

}  // namespace pl

// the C-ABI's opaque domain handle (include/pl.h)
struct pl_domain {
  pl::Domain d;
};
