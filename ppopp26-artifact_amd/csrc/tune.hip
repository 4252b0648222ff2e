// Pass-split autotuner for the NTT (C-ABI pl_plan_*, pl_autotune).
//
// A size-2^lg transform runs as P passes of 2^m[p]-point columns, sum m = lg,
// kLgPerT <= m <= kLgPerT + min(kLgPerT, 4) (3..6 with 8 elements per lane).
// Every split computes the same exact transform; they differ in twiddle-table
// traffic, tile shape per pass and launch count.  The tuner
// times the candidate splits of each size on the device (a batch of
// max(1, 2^22 / N) polynomials per call, so small sizes are measured at the
// throughput the batched form sees rather than at launch latency), keeps
// the fastest, and caches the choices in a text file:
//     # comment
//     <lg> <P> <m0> ... <m(P-1)> <ms>
// Domains built after a split is set use it; domains already built keep theirs.
// With PL_TUNE_HOST it also picks, per size, the copy-group count 2^lg_k of
// the host-to-host NTT (the reference's segment count k of our_ntt) by
// timing pl_ntt_host over lg_k = 0..6, cached as "h <lg> <lg_k> <ms>" lines.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <functional>
#include <vector>

#include "ntt_pass.hpp"

namespace pl {

// compositions of lg into P parts in [MN, MX]; non-increasing only when
// nonincreasing is set
static void compositions(int lg, int P, bool nonincreasing, std::vector<std::vector<int>>& out) {
  std::vector<int> cur;
  std::function<void(int, int)> rec = [&](int left, int k) {
    if (k == P) {
      if (left == 0) out.push_back(cur);
      return;
    }
    const int MX = kLgPerT + (kLgPerT < 4 ? kLgPerT : 4), MN = kLgPerT;
    const int hi = (nonincreasing && !cur.empty()) ? std::min(MX, cur.back()) : MX;
    for (int m = MN; m <= hi; m++) {
      if (left - m < MN * (P - k - 1) || left - m > MX * (P - k - 1)) continue;
      cur.push_back(m);
      rec(left - m, k + 1);
      cur.pop_back();
    }
  };
  rec(lg, 0);
}

std::vector<std::vector<int>> plan_candidates(int lg) {
  std::vector<std::vector<int>> c;
  if (lg < 4 || lg > 28) return c;
  const int MX = kLgPerT + (kLgPerT < 4 ? kLgPerT : 4);
  const int P0 = (lg + MX - 1) / MX;
  compositions(lg, P0, false, c);
  if (P0 + 1 <= PL_MAX_PASSES) compositions(lg, P0 + 1, true, c);
  return c;
}

// the cache file's contents: a pass split per size, and the host-to-host
// copy-group count per size ("h <lg> <lg_k> <ms>" lines; -1: none)
struct TuneCache {
  std::vector<std::vector<int>> plans = std::vector<std::vector<int>>(29);
  std::vector<double> ms = std::vector<double>(29, 0.0);
  std::vector<int> host_lg_k = std::vector<int>(29, -1);
  std::vector<double> host_ms = std::vector<double>(29, 0.0);
};

static bool read_cache(const char* path, TuneCache& c) {
  c = TuneCache();
  auto& plans = c.plans;
  auto& ms = c.ms;
  FILE* f = fopen(path, "r");
  if (!f) return false;
  char line[256];
  while (fgets(line, sizeof(line), f)) {
    if (line[0] == '#') continue;
    if (line[0] == 'h') {
      int lg, k;
      double t = 0;
      if (sscanf(line + 1, "%d %d %lf", &lg, &k, &t) >= 2 && lg >= 1 && lg <= 28 && k >= 0 && k <= kHostLgKMax) {
        c.host_lg_k[lg] = k;
        c.host_ms[lg] = t;
      }
      continue;
    }
    int v[12], n = 0;
    double t = 0;
    char* p = line;
    int lg, P, used;
    if (sscanf(p, "%d %d%n", &lg, &P, &used) != 2) continue;
    p += used;
    if (lg < 4 || lg > 28 || P < 1 || P > PL_MAX_PASSES) continue;
    for (n = 0; n < P; n++) {
      if (sscanf(p, "%d%n", &v[n], &used) != 1) break;
      p += used;
    }
    if (n != P || !plan_valid(lg, v, P)) continue;
    (void)sscanf(p, "%lf", &t);
    plans[lg].assign(v, v + P);
    ms[lg] = t;
  }
  fclose(f);
  return true;
}

static bool write_cache(const char* path, const TuneCache& c) {
  FILE* f = fopen(path, "w");
  if (!f) return false;
  fprintf(f, "# libpl NTT pass splits: lg P m0..m(P-1) ms-per-call (batch of max(1, 2^22/N) polynomials)\n");
  for (int lg = 4; lg <= 28; lg++) {
    if (c.plans[lg].empty()) continue;
    fprintf(f, "%d %d", lg, (int)c.plans[lg].size());
    for (int m : c.plans[lg]) fprintf(f, " %d", m);
    fprintf(f, " %.4f\n", c.ms[lg]);
  }
  bool head = false;
  for (int lg = 1; lg <= 28; lg++) {
    if (c.host_lg_k[lg] < 0) continue;
    if (!head) {
      fprintf(f, "# host-to-host NTT copy groups: h lg lg_k ms-per-call (one polynomial, pinned host buffers)\n");
      head = true;
    }
    fprintf(f, "h %d %d %.4f\n", lg, c.host_lg_k[lg], c.host_ms[lg]);
  }
  fclose(f);
  return true;
}

// time the host-to-host NTT of one polynomial with 2^lg_k copy groups:
// median ms per call over reps calls after one untimed call (each call
// returns once the results are in host memory)
static int time_host(const pl_domain* hd, int lg, int lg_k, void* h_in, void* h_out, void* d_work, int reps,
                     double* ms) {
  int st = pl_ntt_host(h_in, h_out, lg, 1, 0, hd, lg_k, d_work, nullptr);
  std::vector<double> t;
  for (int r = 0; r < reps && st == PL_OK; r++) {
    const auto a = std::chrono::steady_clock::now();
    st = pl_ntt_host(h_in, h_out, lg, 1, 0, hd, lg_k, d_work, nullptr);
    t.push_back(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - a).count());
  }
  if (st != PL_OK) return st;
  std::nth_element(t.begin(), t.begin() + t.size() / 2, t.end());
  *ms = t[t.size() / 2];
  return PL_OK;
}

// time one split: average ms per call of a batch of npoly transforms
static hipError_t time_split(int lg, const std::vector<int>& m, uint32_t npoly, fr* in, fr* out, fr* tmp, int reps,
                             hipEvent_t e0, hipEvent_t e1, double* ms) {
  if (!plan_set(lg, m.data(), (int)m.size())) return hipErrorInvalidValue;
  Domain d;
  hipError_t e = domain_build(d, lg, nullptr);
  NttCall c;
  c.in = in;
  c.out = out;
  c.tmp = tmp;
  c.in_pstride = 1ull << lg;
  c.npoly = npoly;
  if (e == hipSuccess) e = ntt_launch(d, c, nullptr);  // warm-up
  if (e == hipSuccess) e = hipEventRecord(e0, nullptr);
  for (int r = 0; r < reps && e == hipSuccess; r++) e = ntt_launch(d, c, nullptr);
  if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float t = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  *ms = t / reps;
  domain_free(d);
  return e;
}

}  // namespace pl

extern "C" {

int pl_plan_get(int lg, int* m, int* P) {
  if (!m || !P || lg < 1 || lg > 28) return PL_ERR_ARG;
  pl::NttPlan p = pl::make_plan(lg);
  *P = p.P;
  for (int i = 0; i < p.P; i++) m[i] = p.m[i];
  return PL_OK;
}

int pl_plan_set(int lg, const int* m, int P) {
  if (P != 0 && !m) return PL_ERR_ARG;
  return pl::plan_set(lg, m, P) ? PL_OK : PL_ERR_ARG;
}

int pl_plan_candidates(int lg, int* out, int max_cand) {
  auto c = pl::plan_candidates(lg);
  const int n = (int)c.size();
  if (out)
    for (int i = 0; i < n && i < max_cand; i++) {
      for (int k = 0; k < PL_MAX_PASSES; k++) out[i * PL_MAX_PASSES + k] = k < (int)c[i].size() ? c[i][k] : 0;
    }
  return n;
}

int pl_plan_load(const char* path) {
  if (!path) return PL_ERR_ARG;
  pl::TuneCache c;
  if (!pl::read_cache(path, c)) return PL_ERR_ARG;
  int n = 0;
  for (int lg = 4; lg <= 28; lg++)
    if (!c.plans[lg].empty() && pl::plan_set(lg, c.plans[lg].data(), (int)c.plans[lg].size())) n++;
  for (int lg = 1; lg <= 28; lg++)
    if (c.host_lg_k[lg] >= 0 && pl::host_groups_set(lg, c.host_lg_k[lg])) n++;
  return n;
}

int pl_autotune(int lg_lo, int lg_hi, const char* cache_path, int reps, int flags) {
  if (lg_lo < 4 || lg_hi > 28 || lg_lo > lg_hi || reps < 1 || (flags & ~(PL_TUNE_REMEASURE | PL_TUNE_HOST)))
    return PL_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return PL_ERR_NO_DEVICE;
  pl::TuneCache cache;
  if (cache_path && !(flags & PL_TUNE_REMEASURE)) pl::read_cache(cache_path, cache);
  auto& plans = cache.plans;
  auto& ms = cache.ms;
  // buffers for the largest batch: npoly * N = max(N, 2^22) elements
  uint64_t cap = 0;
  for (int lg = lg_lo; lg <= lg_hi; lg++) cap = std::max<uint64_t>(cap, std::max<uint64_t>(1ull << lg, 1ull << 22));
  pl::fr *in = nullptr, *out = nullptr, *tmp = nullptr;
  uint32_t* w = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMalloc(&in, cap * sizeof(pl::fr));
  if (e == hipSuccess) e = hipMalloc(&out, cap * sizeof(pl::fr));
  if (e == hipSuccess) e = hipMalloc(&tmp, cap * sizeof(pl::fr));
  if (e == hipSuccess) e = hipMalloc(&w, pl::uniform_work_bytes(cap));
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = pl::fill_uniform(in, cap, 0x504C4E4B54554E45ull, w, nullptr);
  for (int lg = lg_lo; lg <= lg_hi && e == hipSuccess; lg++) {
    if (!plans[lg].empty()) continue;  // cached
    auto cands = pl::plan_candidates(lg);
    const uint32_t npoly = (uint32_t)std::max<uint64_t>(1, (1ull << 22) >> lg);
    double best = 1e30;
    for (auto& m : cands) {
      double t = 0;
      e = pl::time_split(lg, m, npoly, in, out, tmp, reps, e0, e1, &t);
      if (e != hipSuccess) break;
      if (t < best) {
        best = t;
        plans[lg] = m;
        ms[lg] = t;
      }
    }
  }
  // apply every known choice (cached and measured); sizes without one keep the default
  for (int lg = 4; lg <= 28; lg++) {
    if (!plans[lg].empty()) pl::plan_set(lg, plans[lg].data(), (int)plans[lg].size());
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  for (void* p : {(void*)in, (void*)out, (void*)tmp, (void*)w})
    if (p) (void)hipFree(p);
  int st = e == hipSuccess ? PL_OK : pl::hip_status(e);
  // host-to-host copy groups, with the splits just chosen
  const int hlo = std::max(lg_lo, 12), hhi = std::min(lg_hi, 26);
  if (st == PL_OK && (flags & PL_TUNE_HOST) && hlo <= hhi) {
    const size_t hb = sizeof(pl::fr) << hhi;
    void *h_in = nullptr, *h_out = nullptr, *d_work = nullptr;
    e = hipHostMalloc(&h_in, hb, hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc(&h_out, hb, hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc(&d_work, pl_ntt_host_work_bytes(hhi, 1));
    if (e == hipSuccess) memset(h_in, 0, hb);  // canonical values; the time does not depend on them
    st = e == hipSuccess ? PL_OK : pl::hip_status(e);
    for (int lg = hlo; lg <= hhi && st == PL_OK; lg++) {
      if (cache.host_lg_k[lg] >= 0) continue;  // cached
      pl_domain hd;
      e = pl::domain_build(hd.d, lg, nullptr);
      if (e != hipSuccess) {
        pl::domain_free(hd.d);
        st = pl::hip_status(e);
        break;
      }
      double best = 1e30;
      for (int k = 0; k <= 6 && st == PL_OK; k++) {
        double t = 0;
        st = pl::time_host(&hd, lg, k, h_in, h_out, d_work, reps, &t);
        if (st == PL_OK && t < best) {
          best = t;
          cache.host_lg_k[lg] = k;
          cache.host_ms[lg] = t;
        }
      }
      pl::domain_free(hd.d);
    }
    if (h_in) (void)hipHostFree(h_in);
    if (h_out) (void)hipHostFree(h_out);
    if (d_work) (void)hipFree(d_work);
  }
  for (int lg = 1; lg <= 28; lg++)
    if (cache.host_lg_k[lg] >= 0) pl::host_groups_set(lg, cache.host_lg_k[lg]);
  if (st == PL_OK && cache_path) pl::write_cache(cache_path, cache);
  return st;
}

}  // extern "C"
