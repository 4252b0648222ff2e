// Coset NTT / iNTT over a shifted domain g<w_N> (C-ABI pl_coset_*), the
// surface of the reference's Ntt_coset / Intt_coset (function.cuh:116-142,
// function.cu:354-395; ntt_zkp_coset_* in zksnark_ntt.cu:521-908):
//
//   forward  X[k] = sum_i x_i (g w_N^k)^i          (coset powers, then NTT)
//   inverse  x_i  = g^-i (1/N) sum_k X_k w_N^-ik    (iNTT, then coset powers)
//
// with g = 7 (the group generator whose windowed powers the reference's
// coset steps read, LDE_distribute_powers, kernels.cuh:122-144) unless the
// caller gives another shift.  The forward transform runs the library's pass
// kernels with the coset powers fused into the first pass (the premultiplied
// pass of the LDE: a table of pbd words of g^i R'); the inverse multiplies the
// iNTT's output by a table of g^-i.  The reference splits the coset transform
// into chunked steps (init, forward1/2/3, each a kernel set over a chunk);
// here the same three-call shape maps onto whole-vector passes
// (pl_coset_ntt_step): step 1 the first pass (coset powers fused), step 2 the
// middle passes, step 3 the last pass into natural order.
#include <string.h>

#include <new>

#include "pl_internal.hpp"

struct pl_coset_domain {
  pl::Domain ntt;
  pl::fr* cp = nullptr;   // pbd words of g^i R' (the pass kernels' factor form), i < N
  pl::fr* pw = nullptr;   // g^i, stored form
  pl::fr* ipw = nullptr;  // g^-i, stored form
  size_t bytes = 0;
};

namespace {

constexpr int kMul = PL_VEC_MUL;

void coset_free(pl_coset_domain& d) {
  pl::domain_free(d.ntt);
  for (pl::fr* p : {d.cp, d.pw, d.ipw})
    if (p) (void)hipFree(p);
  d.cp = d.pw = d.ipw = nullptr;
}

hipError_t coset_build(pl_coset_domain& d, int lg, const pl::host::fr64& shift) {
  hipError_t e = pl::domain_build(d.ntt, lg, nullptr);
  if (e != hipSuccess) return e;
  size_t cpb = 0;
  if ((e = pl::coset_table_build(&d.cp, &cpb, lg, 0, shift, nullptr)) != hipSuccess) return e;
  const uint64_t n = 1ull << lg;
  if ((e = hipMalloc((void**)&d.pw, n * sizeof(pl::fr))) != hipSuccess) return e;
  if ((e = hipMalloc((void**)&d.ipw, n * sizeof(pl::fr))) != hipSuccess) return e;
  const pl::host::fr64 one = pl::host::one();
  if ((e = pl::fill_pow_table(d.pw, n, shift, one, nullptr)) != hipSuccess) return e;
  if ((e = pl::fill_pow_table(d.ipw, n, pl::host::inv(shift), one, nullptr)) != hipSuccess) return e;
  d.bytes = cpb + 2 * n * sizeof(pl::fr) + d.ntt.bytes;
  return hipSuccess;
}

bool al16(const void* p) { return ((uintptr_t)p & 15) == 0; }
int st_of(hipError_t e) { return e == hipSuccess ? PL_OK : pl::hip_status(e); }

}  // namespace

extern "C" {

// [Ntt_coset::Ntt_coset / Intt_coset::Intt_coset, function.cu:354-355, 378-379]
int pl_coset_domain_create(int lg, const void* h_shift, pl_coset_domain** out) {
  if (!out || lg < 1) return PL_ERR_ARG;
  if (lg > 28) return PL_ERR_UNSUPPORTED;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return PL_ERR_NO_DEVICE;
  pl::host::fr64 shift;
  if (h_shift) {
    memcpy(shift.l, h_shift, 32);
    if (pl::host::geq_p(shift.l) || pl::host::is_zero(shift)) return PL_ERR_ARG;
  } else {
    shift = pl::host::from_u64(7);
  }
  pl_coset_domain* d = new (std::nothrow) pl_coset_domain();
  if (!d) return PL_ERR_NOMEM;
  hipError_t e = coset_build(*d, lg, shift);
  if (e != hipSuccess) {
    coset_free(*d);
    delete d;
    return pl::hip_status(e);
  }
  *out = d;
  return PL_OK;
}

int pl_coset_domain_destroy(pl_coset_domain* d) {
  if (!d) return PL_ERR_ARG;
  coset_free(*d);
  delete d;
  return PL_OK;
}

size_t pl_coset_domain_bytes(const pl_coset_domain* d) { return d ? d->bytes : 0; }

size_t pl_coset_work_bytes(int lg) {
  if (lg < 1 || lg > 28) return 0;
  return sizeof(pl::fr) << lg;
}

// [Ntt_coset / Intt_coset: init + forward1 + forward2 + forward3 as one call]
int pl_coset_ntt(const void* d_in, void* d_out, int inverse, const pl_coset_domain* d, void* d_work,
                 pl_stream_t stream) {
  if (!d || !d_in || !d_out || !al16(d_in) || !al16(d_out) || !al16(d_work)) return PL_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  const int lg = d->ntt.lg;
  const uint64_t n = 1ull << lg;
  pl::fr* tmp = (pl::fr*)d_work;
  const bool own = !tmp;
  if (own) {
    hipError_t e = hipMallocAsync((void**)&tmp, pl_coset_work_bytes(lg), s);
    if (e != hipSuccess) return pl::hip_status(e);
  }
  pl::NttCall c;
  c.in = (const pl::fr*)d_in;
  c.out = (pl::fr*)d_out;
  c.tmp = tmp;
  c.in_pstride = n;
  c.npoly = 1;
  c.inverse = inverse != 0;
  if (!inverse) {
    c.premul = d->cp;  // g^i on the first pass's loads
    c.pm_mask = 0;
  }
  hipError_t e = pl::ntt_launch(d->ntt, c, s);
  if (e == hipSuccess && inverse) e = pl::vec_op(kMul, (const pl::fr*)d_out, d->ipw, (pl::fr*)d_out, n, nullptr, s);
  if (own) (void)hipFreeAsync(tmp, s);
  return st_of(e);
}

// [Ntt_coset::init / Intt_coset::init, function.cu:357-359, 381-383;
//  NTT_LDE_init's coset powers, ntt.cuh:213-283]: chunk chunk_id of the
// natural-order vector, chunk[i] *= g^(chunk_id * chunk_elems + i) (inverse:
// g^-(...)), in place
int pl_coset_scale(void* d_chunk, uint64_t chunk_elems, uint64_t chunk_id, int inverse, const pl_coset_domain* d,
                   pl_stream_t stream) {
  if (!d || !d_chunk || !al16(d_chunk) || chunk_elems == 0) return PL_ERR_ARG;
  const uint64_t n = 1ull << d->ntt.lg;
  if (chunk_elems > n || chunk_id >= n / chunk_elems + (n % chunk_elems ? 1 : 0)) return PL_ERR_ARG;
  const uint64_t off = chunk_id * chunk_elems;
  if (off + chunk_elems > n) return PL_ERR_ARG;
  const pl::fr* t = (inverse ? d->ipw : d->pw) + off;
  return st_of(pl::vec_op(kMul, (const pl::fr*)d_chunk, t, (pl::fr*)d_chunk, chunk_elems, nullptr,
                          (hipStream_t)stream));
}

// [Ntt_coset::forward1/2/3, Intt_coset::forward1/2/3, function.cu:365-375,
//  385-395]: the coset transform in three calls over the whole vector:
//   step 1  first pass (forward: coset powers fused), d_in -> d_out (may alias)
//   step 2  middle passes, in place on d_in (d_out ignored)
//   step 3  last pass into natural order, d_in -> d_out (distinct; inverse:
//           times g^-i)
// Steps 1..3 in order equal pl_coset_ntt.  A size whose transform is one
// pass does the coset powers in step 1 (forward) and the pass in step 3.
int pl_coset_ntt_step(const void* d_in, void* d_out, int step, int inverse, const pl_coset_domain* d,
                      pl_stream_t stream) {
  if (!d || !d_in || !al16(d_in) || (d_out && !al16(d_out)) || step < 1 || step > 3) return PL_ERR_ARG;
  if (step != 2 && !d_out) return PL_ERR_ARG;
  if (step == 3 && d_in == d_out) return PL_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  const pl::NttPlan& P = d->ntt.plan;
  const uint64_t n = 1ull << d->ntt.lg;
  const int np = P.P == 0 ? 1 : P.P;
  pl::NttCall c;
  c.npoly = 1;
  c.inverse = inverse != 0;
  hipError_t e = hipSuccess;
  if (np == 1) {
    // one pass: the powers as their own step, the pass in step 3
    if (step == 1) {
      if (!inverse) e = pl::vec_op(kMul, (const pl::fr*)d_in, d->pw, (pl::fr*)d_out, n, nullptr, s);
      else if (d_in != d_out) e = hipMemcpyAsync(d_out, d_in, n * sizeof(pl::fr), hipMemcpyDeviceToDevice, s);
    } else if (step == 3) {
      e = pl::ntt_launch_one(d->ntt, c, 0, (const pl::fr*)d_in, n, nullptr, (pl::fr*)d_out, s);
    }
  } else if (step == 1) {
    e = pl::ntt_launch_one(d->ntt, c, 0, (const pl::fr*)d_in, n, inverse ? nullptr : d->cp, (pl::fr*)d_out, s);
  } else if (step == 2) {
    for (int p = 1; p + 1 < np && e == hipSuccess; p++)
      e = pl::ntt_launch_one(d->ntt, c, p, (const pl::fr*)d_in, n, nullptr, (pl::fr*)d_in, s);
  } else {
    e = pl::ntt_launch_one(d->ntt, c, np - 1, (const pl::fr*)d_in, n, nullptr, (pl::fr*)d_out, s);
  }
  if (e == hipSuccess && step == 3 && inverse)
    e = pl::vec_op(kMul, (const pl::fr*)d_out, d->ipw, (pl::fr*)d_out, n, nullptr, s);
  return st_of(e);
}

}  // extern "C"
