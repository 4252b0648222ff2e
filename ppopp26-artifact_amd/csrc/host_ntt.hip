// Host-to-host NTT (pl_ntt_host): the scope of the reference's our_ntt
// timer (ntt.cu:979-1037), which covers host bit reversal, two streamed
// passes over chunks and host gathers/scatters, with every element crossing
// PCIe twice each way.
//
// Here the whole polynomial fits in HBM, so each element crosses once each
// way and the transform runs device-resident, with the copies overlapped:
//  * copy-in stream: the first pass's columns j (stride N >> m0) arrive in
//    2^lg_k groups of consecutive columns, each one strided 2-D copy; the
//    compute stream runs the first pass on a group as soon as it has landed;
//  * the middle passes run on the whole vector;
//  * the last pass runs on 2^lg_k groups of consecutive lines; a group's
//    natural-order outputs are M' rows of consecutive elements, one 2-D copy
//    on the copy-out stream, which starts while the next group computes;
//  * two device slots: polynomial i + 1 is copied in while polynomial i is
//    transformed and copied out (PCIe runs both directions at once).
// The reference's two-slot, two-stream pipeline (run_first_ntt_pass,
// ntt.cu:801-859; run_second_ntt_pass_async_scatter, ntt.cu:861-977) has
// the same overlap structure over chunks of a host-resident vector.
#include <algorithm>
#include <mutex>
#include <vector>

#include "ntt_pass.hpp"

namespace pl {

// copy-group count (lg_k) per size for callers that pass lg_k < 0: set by
// the autotuner or a cache file, else kHostLgKDefault (process-wide)
static std::mutex g_hg_mu;
static int g_host_lg_k[29] = {-1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1,
                              -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1};

int host_groups_get(int lg) {
  if (lg < 1 || lg > 28) return kHostLgKDefault;
  std::lock_guard<std::mutex> g(g_hg_mu);
  return g_host_lg_k[lg] >= 0 ? g_host_lg_k[lg] : kHostLgKDefault;
}

bool host_groups_set(int lg, int lg_k) {
  if (lg < 1 || lg > 28 || lg_k > kHostLgKMax) return false;
  std::lock_guard<std::mutex> g(g_hg_mu);
  g_host_lg_k[lg] = lg_k < 0 ? -1 : lg_k;
  return true;
}

namespace {

struct Streams {
  hipStream_t cin = nullptr, cout = nullptr;
  std::vector<hipEvent_t> ev;
  ~Streams() {
    // the call synchronizes before returning; nothing is pending here
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    if (cin) (void)hipStreamDestroy(cin);
    if (cout) (void)hipStreamDestroy(cout);
  }
  hipError_t event(hipEvent_t* out) {
    hipError_t e = hipEventCreateWithFlags(out, hipEventDisableTiming);
    if (e == hipSuccess) ev.push_back(*out);
    return e;
  }
};

uint64_t pow2_groups(int lg_k, uint64_t lines, uint64_t lines_per_tile) {
  uint64_t g = 1ull << std::min(std::max(lg_k, 0), 20);
  const uint64_t tiles = std::max<uint64_t>(1, lines / std::max<uint64_t>(1, lines_per_tile));
  return std::max<uint64_t>(1, std::min(g, tiles));
}

}  // namespace

static hipError_t ntt_host(const Domain& d, const fr* h_in, fr* h_out, uint32_t npoly, bool inverse, int lg_k,
                           fr* work, hipStream_t s) {
  const int lg = d.lg;
  const uint64_t N = 1ull << lg;
  const NttPlan& P = d.plan;
  const int nslots = npoly > 1 ? 2 : 1;
  Streams st;
  hipError_t e = hipStreamCreateWithFlags(&st.cin, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&st.cout, hipStreamNonBlocking);
  hipEvent_t start, slot_free[2] = {}, done = nullptr;
  if (e == hipSuccess) e = st.event(&start);
  for (int k = 0; k < nslots && e == hipSuccess; k++) e = st.event(&slot_free[k]);
  if (e == hipSuccess) e = st.event(&done);
  if (e != hipSuccess) return e;
  // the copies start after the caller's earlier work on `stream`
  if ((e = hipEventRecord(start, s)) != hipSuccess) return e;
  if ((e = hipStreamWaitEvent(st.cin, start, 0)) != hipSuccess) return e;

  const bool grouped = P.P >= 2;
  const int m0 = grouped ? P.m[0] : 0, mL = grouped ? P.m[P.P - 1] : 0;
  const uint64_t S0 = N >> m0, L = N >> mL;
  const uint64_t G0 = grouped ? pow2_groups(lg_k, S0, kTile >> m0) : 1;
  const uint64_t G1 = grouped ? pow2_groups(lg_k, L, kTile >> mL) : 1;
  const uint64_t W0 = S0 / G0, W1 = L / G1;
  std::vector<hipEvent_t> ev_in(G0), ev_out(G1);
  for (auto& v : ev_in)
    if (e == hipSuccess) e = st.event(&v);
  for (auto& v : ev_out)
    if (e == hipSuccess) e = st.event(&v);
  NttCall c;
  c.npoly = 1;
  c.inverse = inverse;
  for (uint32_t i = 0; i < npoly && e == hipSuccess; i++) {
    const int k = (int)(i % nslots);
    fr* data = work + (uint64_t)k * 2 * N;
    fr* tmp = data + N;
    const fr* src = h_in + (uint64_t)i * N;
    fr* dst = h_out + (uint64_t)i * N;
    if (i >= (uint32_t)nslots && (e = hipStreamWaitEvent(st.cin, slot_free[k], 0)) != hipSuccess) break;
    if (!grouped) {
      c.in = data;
      c.out = data;
      c.tmp = tmp;
      c.in_pstride = N;
      e = hipMemcpyAsync(data, src, N * sizeof(fr), hipMemcpyHostToDevice, st.cin);
      if (e == hipSuccess) e = hipEventRecord(ev_in[0], st.cin);
      if (e == hipSuccess) e = hipStreamWaitEvent(s, ev_in[0], 0);
      if (e == hipSuccess) e = ntt_launch(d, c, s);
      if (e == hipSuccess) e = hipEventRecord(ev_out[0], s);
      if (e == hipSuccess) e = hipStreamWaitEvent(st.cout, ev_out[0], 0);
      if (e == hipSuccess) e = hipMemcpyAsync(dst, data, N * sizeof(fr), hipMemcpyDeviceToHost, st.cout);
    } else {
      // first pass by column groups as they arrive
      for (uint64_t g = 0; g < G0 && e == hipSuccess; g++) {
        const uint64_t j0 = g * W0;
        e = hipMemcpy2DAsync(data + j0, S0 * sizeof(fr), src + j0, S0 * sizeof(fr), W0 * sizeof(fr),
                             (size_t)1 << m0, hipMemcpyHostToDevice, st.cin);
        if (e == hipSuccess) e = hipEventRecord(ev_in[g], st.cin);
        if (e == hipSuccess) e = hipStreamWaitEvent(s, ev_in[g], 0);
        if (e == hipSuccess) e = ntt_launch_part(d, c, 0, data, tmp, j0, j0, W0, s);
      }
      // middle passes in place on the whole vector
      for (int p = 1; p + 1 < P.P && e == hipSuccess; p++) e = ntt_launch_one(d, c, p, tmp, N, nullptr, tmp, s);
      // last pass by line groups; each group's outputs leave at once
      for (uint64_t g = 0; g < G1 && e == hipSuccess; g++) {
        const uint64_t l0 = g * W1;
        e = ntt_launch_part(d, c, P.P - 1, tmp, data, last_pass_row(P, l0) << mL, l0, W1, s);
        if (e == hipSuccess) e = hipEventRecord(ev_out[g], s);
        if (e == hipSuccess) e = hipStreamWaitEvent(st.cout, ev_out[g], 0);
        if (e == hipSuccess)
          e = hipMemcpy2DAsync(dst + l0, L * sizeof(fr), data + l0, L * sizeof(fr), W1 * sizeof(fr),
                               (size_t)1 << mL, hipMemcpyDeviceToHost, st.cout);
      }
    }
    if (e == hipSuccess) e = hipEventRecord(slot_free[k], st.cout);
  }
  // the caller's stream continues after the last copy-out
  if (e == hipSuccess) e = hipEventRecord(done, st.cout);
  if (e == hipSuccess) e = hipStreamWaitEvent(s, done, 0);
  hipError_t e2 = hipStreamSynchronize(st.cout);
  hipError_t e3 = hipStreamSynchronize(st.cin);
  hipError_t e4 = hipStreamSynchronize(s);
  if (e == hipSuccess) e = e2;
  if (e == hipSuccess) e = e3;
  if (e == hipSuccess) e = e4;
  return e;
}

}  // namespace pl

extern "C" {

size_t pl_ntt_host_work_bytes(int lg, uint32_t npoly) {
  if (lg < 1 || lg > 28 || npoly == 0) return 0;
  return (size_t)(npoly > 1 ? 2 : 1) * 2 * (sizeof(pl::fr) << lg);
}

// [our_ntt, ntt.cu:979-1037]
int pl_ntt_host(const void* h_in, void* h_out, int lg, uint32_t npoly, int inverse, const pl_domain* d, int lg_k,
                void* d_work, pl_stream_t stream) {
  if (!h_in || !h_out || !d || npoly == 0 || d->d.lg != lg) return PL_ERR_ARG;
  if (((uintptr_t)d_work & 15) != 0) return PL_ERR_ARG;
  if (lg_k < 0) lg_k = pl::host_groups_get(lg);  // tuned or default
  hipStream_t s = (hipStream_t)stream;
  void* w = d_work;
  // without a caller workspace the slots come from the stream-ordered pool
  // (no device-wide synchronization, reused across calls)
  if (!w) {
    hipError_t e = hipMallocAsync(&w, pl_ntt_host_work_bytes(lg, npoly), s);
    if (e != hipSuccess) return pl::hip_status(e);
  }
  hipError_t e = pl::ntt_host(d->d, (const pl::fr*)h_in, (pl::fr*)h_out, npoly, inverse != 0, lg_k, (pl::fr*)w, s);
  if (!d_work) {
    hipError_t e2 = hipFreeAsync(w, s);
    if (e == hipSuccess) e = e2;
  }
  return e == hipSuccess ? PL_OK : pl::hip_status(e);
}

int pl_ntt_host_groups(int lg) {
  if (lg < 1 || lg > 28) return PL_ERR_ARG;
  return pl::host_groups_get(lg);
}

int pl_ntt_host_groups_set(int lg, int lg_k) { return pl::host_groups_set(lg, lg_k) ? PL_OK : PL_ERR_ARG; }

}  // extern "C"
