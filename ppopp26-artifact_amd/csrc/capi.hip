// C-ABI entry points of libpl (declared in include/pl.h).
#include <string.h>

#include <new>

#include "ntt_pass.hpp"

struct pl_lde_domain {
  pl::LdeDomain d;
};

namespace {

thread_local char g_last_error[256] = "";

int hip_fail(hipError_t e) {
  snprintf(g_last_error, sizeof(g_last_error), "%s", hipGetErrorString(e));
  return e == hipErrorOutOfMemory ? PL_ERR_NOMEM : PL_ERR_HIP;
}

bool aligned16(const void* p) { return ((uintptr_t)p & 15) == 0; }

// [a, a + na) and [b, b + nb) share a byte
bool ranges_overlap(const void* a, uint64_t na, const void* b, uint64_t nb) {
  const uintptr_t x = (uintptr_t)a, y = (uintptr_t)b;
  return x < y + nb && y < x + na;
}

}  // namespace

namespace pl {
int hip_status(hipError_t e) { return hip_fail(e); }
}  // namespace pl

extern "C" {

const char* pl_status_string(int status) {
  switch (status) {
    case PL_OK: return "ok";
    case PL_ERR_ARG: return "invalid argument";
    case PL_ERR_HIP: return "HIP runtime error";
    case PL_ERR_UNSUPPORTED: return "unsupported size";
    case PL_ERR_NOMEM: return "device out of memory";
    case PL_ERR_NO_DEVICE: return "no HIP device";
    default: return "unknown status";
  }
}

int pl_version(void) { return 100; }

const char* pl_last_hip_error(void) { return g_last_error; }

int pl_domain_create(int lg, pl_domain** out) {
  if (!out || lg < 1 || lg > 28) return lg > 28 ? PL_ERR_UNSUPPORTED : PL_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return PL_ERR_NO_DEVICE;
  pl_domain* d = new (std::nothrow) pl_domain();
  if (!d) return PL_ERR_NOMEM;
  (void)hipGetDevice(&d->d.device);
  hipError_t e = pl::domain_build(d->d, lg, nullptr);
  if (e != hipSuccess) {
    pl::domain_free(d->d);
    delete d;
    return hip_fail(e);
  }
  *out = d;
  return PL_OK;
}

int pl_domain_destroy(pl_domain* d) {
  if (!d) return PL_ERR_ARG;
  pl::domain_free(d->d);
  delete d;
  return PL_OK;
}

int pl_params_fill(void* d_out, pl_stream_t stream) {
  if (!d_out || !aligned16(d_out)) return PL_ERR_ARG;
  hipError_t e = pl::params_fill((pl::fr*)d_out, (hipStream_t)stream);
  return e == hipSuccess ? PL_OK : hip_fail(e);
}

int pl_domain_lg(const pl_domain* d) { return d ? d->d.lg : PL_ERR_ARG; }

size_t pl_domain_bytes(const pl_domain* d) { return d ? d->d.bytes : 0; }

// Every split of lg >= 6 may have two or more passes (columns of 2^3..2^6
// points), and such a transform needs an intermediate buffer; the size does
// not depend on the split in use, so a workspace sized here fits a domain
// built under any split.
size_t pl_ntt_work_bytes(int lg, uint32_t npoly) {
  if (lg < 1 || lg > 28) return 0;
  return lg >= 2 * pl::kLgPerT ? ((size_t)npoly << lg) * sizeof(pl::fr) : 0;
}

int pl_ntt_batch(const void* d_in, void* d_out, int lg, uint32_t npoly, int inverse, const pl_domain* d,
                 void* d_work, pl_stream_t stream) {
  if (!d || !d_in || !d_out || npoly == 0) return PL_ERR_ARG;
  if (d->d.lg != lg) return PL_ERR_ARG;
  if (!aligned16(d_in) || !aligned16(d_out) || !aligned16(d_work)) return PL_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  // the domain's own split decides whether the passes need the buffer
  size_t wb = d->d.plan.P >= 2 ? ((size_t)npoly << lg) * sizeof(pl::fr) : 0;
  pl::fr* tmp = (pl::fr*)d_work;
  bool own = false;
  if (wb && !tmp) {
    hipError_t e = hipMallocAsync((void**)&tmp, wb, s);
    if (e != hipSuccess) return hip_fail(e);
    own = true;
  }
  pl::NttCall c;
  c.in = (const pl::fr*)d_in;
  c.out = (pl::fr*)d_out;
  c.tmp = tmp;
  c.in_pstride = 1ull << lg;
  c.npoly = npoly;
  c.inverse = inverse != 0;
  hipError_t e = pl::ntt_launch(d->d, c, s);
  if (own) (void)hipFreeAsync(tmp, s);
  return e == hipSuccess ? PL_OK : hip_fail(e);
}

int pl_ntt_num_passes(int lg) {
  if (lg < 1 || lg > 28) return PL_ERR_UNSUPPORTED;
  pl::NttPlan p = pl::make_plan(lg);
  return p.P == 0 ? 1 : p.P;
}

int pl_domain_num_passes(const pl_domain* d) {
  if (!d) return PL_ERR_ARG;
  return d->d.plan.P == 0 ? 1 : d->d.plan.P;
}

int pl_ntt_pass(const void* d_in, void* d_out, int lg, uint32_t npoly, int pass, int inverse, const pl_domain* d,
                pl_stream_t stream) {
  if (!d || !d_in || !d_out || npoly == 0 || d->d.lg != lg) return PL_ERR_ARG;
  if (!aligned16(d_in) || !aligned16(d_out)) return PL_ERR_ARG;
  const int np = d->d.plan.P == 0 ? 1 : d->d.plan.P;  // the domain's own split
  if (pass < 0 || pass >= np) return PL_ERR_ARG;
  // the last pass scatters across the whole polynomial: not in place
  if (pass == np - 1 && np > 1 && d_in == d_out) return PL_ERR_ARG;
  pl::NttCall c;
  c.npoly = npoly;
  c.inverse = inverse != 0;
  hipError_t e = pl::ntt_launch_one(d->d, c, pass, (const pl::fr*)d_in, 1ull << lg, nullptr, (pl::fr*)d_out,
                                    (hipStream_t)stream);
  return e == hipSuccess ? PL_OK : hip_fail(e);
}

int pl_ntt(const void* d_in, void* d_out, int lg, int inverse, const pl_domain* d, void* d_work,
           pl_stream_t stream) {
  return pl_ntt_batch(d_in, d_out, lg, 1, inverse, d, d_work, stream);
}

int pl_lde_domain_create(int lg_n, int lg_blowup, const void* h_shift, pl_lde_domain** out) {
  if (!out || lg_n < 1 || lg_blowup < 1) return PL_ERR_ARG;
  if (lg_blowup > 4 || lg_n + lg_blowup > 28) return PL_ERR_UNSUPPORTED;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return PL_ERR_NO_DEVICE;
  pl::host::fr64 shift;
  if (h_shift) {
    memcpy(shift.l, h_shift, 32);
    if (pl::host::geq_p(shift.l)) return PL_ERR_ARG;
  } else {
    // the reference's coset factor 7^(2^lambda) (lde.cu:642-645, 659)
    shift = pl::host::pow_u64(pl::host::from_u64(7), 1ull << lg_blowup);
  }
  pl_lde_domain* d = new (std::nothrow) pl_lde_domain();
  if (!d) return PL_ERR_NOMEM;
  hipError_t e = pl::lde_domain_build(d->d, lg_n, lg_blowup, shift, nullptr);
  if (e != hipSuccess) {
    pl::lde_domain_free(d->d);
    delete d;
    return hip_fail(e);
  }
  *out = d;
  return PL_OK;
}

int pl_lde_domain_destroy(pl_lde_domain* d) {
  if (!d) return PL_ERR_ARG;
  pl::lde_domain_free(d->d);
  delete d;
  return PL_OK;
}

size_t pl_lde_domain_bytes(const pl_lde_domain* d) { return d ? d->d.bytes + d->d.ntt.bytes : 0; }

size_t pl_lde_work_bytes(int lg_n, int lg_blowup) {
  if (lg_n < 1 || lg_blowup < 1 || lg_n + lg_blowup > 28) return 0;
  return (size_t)sizeof(pl::fr) << (lg_n + lg_blowup);
}

// blowup 3 (three coset NTTs of size N with shifts g w_3N^t; no
// counterpart in the reference, whose our_lde takes power-of-two blowups)
int pl_lde3_domain_create(int lg_n, const void* h_shift, pl_lde_domain** out) {
  if (!out || lg_n < 1) return PL_ERR_ARG;
  if (lg_n > 26) return PL_ERR_UNSUPPORTED;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return PL_ERR_NO_DEVICE;
  pl::host::fr64 shift = pl::host::from_u64(7);
  if (h_shift) {
    memcpy(shift.l, h_shift, 32);
    if (pl::host::geq_p(shift.l)) return PL_ERR_ARG;
  }
  pl_lde_domain* d = new (std::nothrow) pl_lde_domain();
  if (!d) return PL_ERR_NOMEM;
  hipError_t e = pl::lde_domain_build(d->d, lg_n, -1, shift, nullptr);
  if (e != hipSuccess) {
    pl::lde_domain_free(d->d);
    delete d;
    return hip_fail(e);
  }
  *out = d;
  return PL_OK;
}

size_t pl_lde_domain_work_bytes(const pl_lde_domain* d) {
  if (!d) return 0;
  const size_t m = (size_t)d->d.blowup << d->d.lg_n;
  // blowup 3: the pass scratch and the cosets' outputs before the interleave
  return sizeof(pl::fr) * (d->d.blowup == 3 ? 2 * m : m);
}

int pl_lde_domain_blowup(const pl_lde_domain* d) { return d ? d->d.blowup : PL_ERR_ARG; }

int pl_lde(const void* d_coeffs, void* d_out, const pl_lde_domain* d, void* d_work, pl_stream_t stream) {
  if (!d || !d_coeffs || !d_out) return PL_ERR_ARG;
  if (!aligned16(d_coeffs) || !aligned16(d_out) || !aligned16(d_work)) return PL_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  const int lg_n = d->d.lg_n, lb = d->d.lg_blowup;
  const bool three = d->d.blowup == 3;
  size_t wb = pl_lde_domain_work_bytes(d);
  pl::fr* tmp = (pl::fr*)d_work;
  bool own = false;
  if (!tmp) {
    hipError_t e = hipMallocAsync((void**)&tmp, wb, s);
    if (e != hipSuccess) return hip_fail(e);
    own = true;
  }
  const uint64_t n = 1ull << lg_n;
  pl::NttCall c;
  c.in = (const pl::fr*)d_coeffs;
  c.tmp = tmp;
  c.in_pstride = 0;  // every coset reads the same coefficients
  c.premul = d->d.cp;
  c.inverse = false;
  if (three) {
    // three coset NTTs side by side (coset t at t N), then out[3 j + t]
    c.out = tmp + 3 * n;
    c.npoly = 3;
    c.pm_mask = 3;
    c.lgI = 0;
  } else {
    c.out = (pl::fr*)d_out;
    c.npoly = 1u << lb;
    c.pm_mask = (1u << lb) - 1;
    c.lgI = lb;
  }
  hipError_t e = pl::ntt_launch(d->d.ntt, c, s);
  if (e == hipSuccess && three) e = pl::interleave3(c.out, (pl::fr*)d_out, n, s);
  if (own) (void)hipFreeAsync(tmp, s);
  return e == hipSuccess ? PL_OK : hip_fail(e);
}

size_t pl_scan_work_bytes(uint64_t n) { return (size_t)pl::scan_scratch_elems(n) * sizeof(pl::fr); }

namespace {
// stream-ordered scratch when the caller passes none
struct Scratch {
  pl::fr* p = nullptr;
  bool own = false;
  hipStream_t s = nullptr;
  hipError_t get(void* user, size_t bytes, hipStream_t st) {
    s = st;
    if (user) {
      p = (pl::fr*)user;
      return hipSuccess;
    }
    own = true;
    return hipMallocAsync((void**)&p, bytes, st);
  }
  ~Scratch() {
    if (own && p) (void)hipFreeAsync(p, s);
  }
};
}  // namespace

int pl_grand_product(void* d_inout, uint64_t n, void* d_incl_out, void* d_work, pl_stream_t stream) {
  if (!d_inout || n == 0) return PL_ERR_ARG;
  if (!aligned16(d_inout) || !aligned16(d_incl_out) || !aligned16(d_work)) return PL_ERR_ARG;
  Scratch sc;
  hipError_t e = sc.get(d_work, pl_scan_work_bytes(n), (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e);
  e = pl::grand_product((pl::fr*)d_inout, (pl::fr*)d_incl_out, n, sc.p, (hipStream_t)stream);
  return e == hipSuccess ? PL_OK : hip_fail(e);
}

int pl_eval(const void* d_poly, uint64_t n, const void* d_x, void* d_y, void* d_work, pl_stream_t stream) {
  if (!d_x || !d_y || (n && !d_poly)) return PL_ERR_ARG;
  if (!aligned16(d_poly) || !aligned16(d_x) || !aligned16(d_y) || !aligned16(d_work)) return PL_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  if (n == 0) {
    // evaluate() of an empty polynomial is zero (function.cu:171-176)
    hipError_t e = pl::fill_zero((pl::fr*)d_y, 1, s);
    return e == hipSuccess ? PL_OK : hip_fail(e);
  }
  Scratch sc;
  hipError_t e = sc.get(d_work, pl_scan_work_bytes(n), s);
  if (e != hipSuccess) return hip_fail(e);
  e = pl::evaluate((const pl::fr*)d_poly, n, (const pl::fr*)d_x, (pl::fr*)d_y, sc.p, s);
  return e == hipSuccess ? PL_OK : hip_fail(e);
}

int pl_div_linear(void* d_inout, uint64_t n, const void* d_c, void* d_rem, void* d_work, pl_stream_t stream) {
  if (!d_inout || !d_c || n == 0) return PL_ERR_ARG;
  if (!aligned16(d_inout) || !aligned16(d_c) || !aligned16(d_rem) || !aligned16(d_work)) return PL_ERR_ARG;
  Scratch sc;
  hipError_t e = sc.get(d_work, pl_scan_work_bytes(n), (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e);
  e = pl::div_linear((pl::fr*)d_inout, n, (const pl::fr*)d_c, (pl::fr*)d_rem, sc.p, (hipStream_t)stream);
  return e == hipSuccess ? PL_OK : hip_fail(e);
}

// The element-wise kernels run one thread per element in blocks of 256;
// n <= 2^36 (2 TiB of elements, beyond any one GPU's memory) keeps the
// block count inside 32 bits.
constexpr uint64_t kMaxElemwise = 1ull << 36;

size_t pl_vec_work_bytes(uint64_t n) {
  uint64_t a = pl::ops_scratch_elems(n), b = pl::scan_scratch_elems(n);
  return (size_t)(a > b ? a : b) * sizeof(pl::fr);
}

int pl_vec_op(int op, const void* d_a, const void* d_b, void* d_out, uint64_t n, void* d_work, pl_stream_t stream) {
  if (op < PL_VEC_ADD || op > PL_VEC_FROM_MONT) return PL_ERR_ARG;
  const bool needs_b = !(op == PL_VEC_NEG || op == PL_VEC_INV || op == PL_VEC_TO_MONT || op == PL_VEC_FROM_MONT);
  if (n > kMaxElemwise || (n && (!d_a || !d_out || (needs_b && !d_b)))) return PL_ERR_ARG;
  if (!aligned16(d_a) || !aligned16(d_b) || !aligned16(d_out) || !aligned16(d_work)) return PL_ERR_ARG;
  if (n == 0) return PL_OK;
  Scratch sc;
  const bool inv = op == PL_VEC_DIV || op == PL_VEC_INV || op == PL_VEC_DIV_SCALAR;
  if (inv) {
    hipError_t e = sc.get(d_work, pl_vec_work_bytes(n), (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e);
  }
  hipError_t e = pl::vec_op(op, (const pl::fr*)d_a, (const pl::fr*)d_b, (pl::fr*)d_out, n, sc.p, (hipStream_t)stream);
  return e == hipSuccess ? PL_OK : hip_fail(e);
}

int pl_exp_mod(const void* d_in, void* d_out, uint64_t n, int exp, pl_stream_t stream) {
  if (exp < 0 || n > kMaxElemwise || (n && (!d_in || !d_out)) || !aligned16(d_in) || !aligned16(d_out))
    return PL_ERR_ARG;
  hipError_t e = pl::exp_mod((const pl::fr*)d_in, (pl::fr*)d_out, n, (uint32_t)exp, (hipStream_t)stream);
  return e == hipSuccess ? PL_OK : hip_fail(e);
}

int pl_gen_sequence(uint64_t n, const void* d_x, void* d_out, pl_stream_t stream) {
  if (n > (1ull << 28) || (n && (!d_x || !d_out)) || !aligned16(d_x) || !aligned16(d_out)) return PL_ERR_ARG;
  if (n && ranges_overlap(d_x, 32, d_out, n * 32)) return PL_ERR_ARG;
  hipError_t e = pl::gen_sequence((const pl::fr*)d_x, (pl::fr*)d_out, n, (hipStream_t)stream);
  return e == hipSuccess ? PL_OK : hip_fail(e);
}

int pl_repeat_to_poly(const void* d_x, void* d_out, uint64_t n, pl_stream_t stream) {
  if (n > kMaxElemwise || (n && (!d_x || !d_out)) || !aligned16(d_x) || !aligned16(d_out)) return PL_ERR_ARG;
  if (n && ranges_overlap(d_x, 32, d_out, n * 32)) return PL_ERR_ARG;
  hipError_t e = pl::repeat_to_poly((const pl::fr*)d_x, (pl::fr*)d_out, n, (hipStream_t)stream);
  return e == hipSuccess ? PL_OK : hip_fail(e);
}

// shared argument checks of the permutation entry points
static int perm_args(const void* const* d_w, const void* const* d_s, const void* d_beta, const void* d_gamma,
                     const void* h_k, int lg, const pl_domain* d, void* d_out, void* d_work, pl::host::fr64 k[3]) {
  if (!d_w || !d_s || !d_beta || !d_gamma || !h_k || !d || !d_out || d->d.lg != lg) return PL_ERR_ARG;
  for (int j = 0; j < 3; j++)
    if (!d_w[j] || !d_s[j] || !aligned16(d_w[j]) || !aligned16(d_s[j])) return PL_ERR_ARG;
  if (!aligned16(d_beta) || !aligned16(d_gamma) || !aligned16(d_out) || !aligned16(d_work)) return PL_ERR_ARG;
  memcpy(k, h_k, 3 * sizeof(pl::host::fr64));
  for (int j = 0; j < 3; j++)
    if (pl::host::geq_p(k[j].l)) return PL_ERR_ARG;
  return PL_OK;
}

int pl_perm_ratio(const void* const* d_w, const void* const* d_s, const void* d_beta, const void* d_gamma,
                  const void* h_k, int lg, const pl_domain* d, void* d_ratio, void* d_work, pl_stream_t stream) {
  pl::host::fr64 k[3];
  const int st = perm_args(d_w, d_s, d_beta, d_gamma, h_k, lg, d, d_ratio, d_work, k);
  if (st != PL_OK) return st;
  const uint64_t n = 1ull << lg;
  Scratch sc;
  hipError_t e = sc.get(d_work, pl_vec_work_bytes(n), (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e);
  const pl::fr* w[3] = {(const pl::fr*)d_w[0], (const pl::fr*)d_w[1], (const pl::fr*)d_w[2]};
  const pl::fr* sg[3] = {(const pl::fr*)d_s[0], (const pl::fr*)d_s[1], (const pl::fr*)d_s[2]};
  e = pl::perm_ratio(w, sg, (const pl::fr*)d_beta, (const pl::fr*)d_gamma, d->d.pw0, d->d.pw1, k, (pl::fr*)d_ratio,
                     n, sc.p, (hipStream_t)stream);
  return e == hipSuccess ? PL_OK : hip_fail(e);
}

int pl_perm_grand_product(const void* const* d_w, const void* const* d_s, const void* d_beta,
                          const void* d_gamma, const void* h_k, int lg, const pl_domain* d, void* d_z,
                          void* d_work, pl_stream_t stream) {
  pl::host::fr64 k[3];
  const int st = perm_args(d_w, d_s, d_beta, d_gamma, h_k, lg, d, d_z, d_work, k);
  if (st != PL_OK) return st;
  const uint64_t n = 1ull << lg;
  Scratch sc;
  hipError_t e = sc.get(d_work, pl_vec_work_bytes(n), (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e);
  const pl::fr* w[3] = {(const pl::fr*)d_w[0], (const pl::fr*)d_w[1], (const pl::fr*)d_w[2]};
  const pl::fr* sg[3] = {(const pl::fr*)d_s[0], (const pl::fr*)d_s[1], (const pl::fr*)d_s[2]};
  e = pl::perm_grand_product(w, sg, (const pl::fr*)d_beta, (const pl::fr*)d_gamma, d->d.pw0, d->d.pw1, k,
                             (pl::fr*)d_z, n, sc.p, (hipStream_t)stream);
  return e == hipSuccess ? PL_OK : hip_fail(e);
}

size_t pl_uniform_work_bytes(uint64_t n) { return (size_t)pl::uniform_work_bytes(n); }

int pl_fill_uniform(void* d_out, uint64_t n, uint64_t seed, void* d_work, pl_stream_t stream) {
  if ((!d_out && n) || n >= (1ull << 31) || !aligned16(d_out) || ((uintptr_t)d_work & 3)) return PL_ERR_ARG;
  if (n == 0) return PL_OK;
  hipStream_t s = (hipStream_t)stream;
  void* w = d_work;
  if (!w) {
    hipError_t e = hipMallocAsync(&w, pl_uniform_work_bytes(n), s);
    if (e != hipSuccess) return hip_fail(e);
  }
  hipError_t e = pl::fill_uniform((pl::fr*)d_out, n, seed, (uint32_t*)w, s);
  if (!d_work) (void)hipFreeAsync(w, s);
  return e == hipSuccess ? PL_OK : hip_fail(e);
}

int pl_digest(const void* d_data, uint64_t n, void* d_out, pl_stream_t stream) {
  if ((!d_data && n) || !d_out || ((uintptr_t)d_data & 7) || ((uintptr_t)d_out & 7)) return PL_ERR_ARG;
  hipError_t e = pl::digest((const pl::fr*)d_data, n, (uint64_t*)d_out, (hipStream_t)stream);
  return e == hipSuccess ? PL_OK : hip_fail(e);
}

int pl_probe_mad_rate(double* tmad_per_s, pl_stream_t stream) {
  if (!tmad_per_s) return PL_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return PL_ERR_NO_DEVICE;
  hipError_t e = pl::mad_probe(tmad_per_s, (hipStream_t)stream);
  return e == hipSuccess ? PL_OK : hip_fail(e);
}

int pl_ntt_naive(const void* d_in, void* d_out, int lg, const pl_domain* d, pl_stream_t stream) {
  if (!d || !d_in || !d_out || d_in == d_out || d->d.lg != lg) return PL_ERR_ARG;
  if (!aligned16(d_in) || !aligned16(d_out)) return PL_ERR_ARG;
  hipError_t e = pl::ntt_naive(d->d, (const pl::fr*)d_in, (pl::fr*)d_out, (hipStream_t)stream);
  return e == hipSuccess ? PL_OK : hip_fail(e);
}

int pl_lde_naive(const void* d_coeffs, void* d_out, int lg_n, int lg_blowup, const pl_domain* d, pl_stream_t stream) {
  if (!d || !d_coeffs || !d_out || d_coeffs == d_out || d->d.lg != lg_n || lg_blowup < 1 || lg_n + lg_blowup > 28)
    return PL_ERR_ARG;
  if (!aligned16(d_coeffs) || !aligned16(d_out)) return PL_ERR_ARG;
  hipError_t e = pl::lde_naive(d->d, (const pl::fr*)d_coeffs, (pl::fr*)d_out, lg_blowup, (hipStream_t)stream);
  return e == hipSuccess ? PL_OK : hip_fail(e);
}

}  // extern "C"
