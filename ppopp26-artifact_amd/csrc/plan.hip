// Phase planner (batched and served forms) and the process-wide domain cache.
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>

#include "plan.hpp"

namespace pl {

// ---- domain cache ----------------------------------------------------------
// keyed by (device, lg) and (device, lg_n, lg_blowup); held until the process
// ends
static std::mutex g_cache_mu;
static std::map<std::pair<int, int>, Domain*> g_domains;
struct CosetTable {
  fr* cp = nullptr;
  size_t bytes = 0;
};
static std::map<std::tuple<int, int, int>, CosetTable> g_coset;

static int current_device() {
  int dev = 0;
  (void)hipGetDevice(&dev);
  return dev;
}

Domain* cached_domain(int lg, hipError_t* err) {
  std::lock_guard<std::mutex> g(g_cache_mu);
  const auto key = std::make_pair(current_device(), lg);
  auto it = g_domains.find(key);
  if (it != g_domains.end()) return it->second;
  Domain* d = new Domain();
  hipError_t e = domain_build(*d, lg, nullptr);
  if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
  if (e != hipSuccess) {
    domain_free(*d);
    delete d;
    if (err) *err = e;
    return nullptr;
  }
  d->device = key.first;
  g_domains[key] = d;
  return d;
}

const fr* cached_coset_table(int lg_n, int lg_blowup, hipError_t* err) {
  std::lock_guard<std::mutex> g(g_cache_mu);
  const auto key = std::make_tuple(current_device(), lg_n, lg_blowup);
  auto it = g_coset.find(key);
  if (it != g_coset.end()) return it->second.cp;
  CosetTable t;
  // the reference's coset factor 7^(2^lambda) (lde.cu:642-645, 659)
  host::fr64 shift = host::pow_u64(host::from_u64(7), 1ull << lg_blowup);
  hipError_t e = coset_table_build(&t.cp, &t.bytes, lg_n, lg_blowup, shift, nullptr);
  if (e != hipSuccess) {
    if (t.cp) (void)hipFree(t.cp);
    if (err) *err = e;
    return nullptr;
  }
  g_coset[key] = t;
  return t.cp;
}

size_t cached_domain_bytes() {
  std::lock_guard<std::mutex> g(g_cache_mu);
  size_t b = 0;
  for (auto& kv : g_domains) b += kv.second->bytes;
  for (auto& kv : g_coset) b += kv.second.bytes;
  return b;
}

// ---- planning --------------------------------------------------------------

hipError_t phase_plan(PhasePlan& ph, const std::vector<Item>& items, Bump& bump, Bump* tables) {
  Bump& tb = tables ? *tables : bump;
  hipError_t e = hipSuccess;
  ph = PhasePlan();
  ph.items = items;
  std::map<std::tuple<int, int, int, int, int>, size_t> gidx;
  int n_eval = 0;
  for (size_t i = 0; i < ph.items.size(); i++) {
    const Item& it = ph.items[i];
    const uint64_t n = 1ull << it.lg;
    if (it.op == IT_EVAL_DIV) {
      fr* scr = (fr*)bump.get(scan_scratch_elems(n) * sizeof(fr));
      ph.scans.push_back({IT_EVAL_DIV, (int)i, nullptr, scr, n});
      n_eval++;
      continue;
    }
    if (it.op == IT_PERM_GP) {
      const Domain* d = cached_domain(it.lg, &e);
      if (!d) return e;
      const uint64_t se = ops_scratch_elems(n);
      fr* scr = (fr*)bump.get(se * sizeof(fr));
      ph.scans.push_back({IT_PERM_GP, (int)i, d, scr, n});
      continue;
    }
    NttCall c;
    const Domain* d = nullptr;
    c.in = it.in;
    c.out = it.out;
    if (it.op == IT_LDE) {
      d = cached_domain(it.lg, &e);
      const fr* cp = d ? cached_coset_table(it.lg, it.lg_blowup, &e) : nullptr;
      if (!cp) return e;
      c.in_pstride = 0;  // every coset reads the same coefficients
      c.npoly = 1u << it.lg_blowup;
      c.premul = cp;
      c.pm_mask = c.npoly - 1;
      c.lgI = it.lg_blowup;
    } else {
      d = cached_domain(it.lg, &e);
      if (!d) return e;
      c.in_pstride = n;
      c.npoly = 1;
      c.inverse = it.op == IT_INTT;
    }
    const bool need_tmp = d->plan.P >= 2 || c.premul;
    if (need_tmp)
      c.tmp = it.use_shared_tmp ? it.shared_tmp : (fr*)bump.get(((size_t)c.npoly << it.lg) * sizeof(fr));
    ph.plain_calls.push_back({d, c});
    PassArgs a;
    if (!pass_args_regular(*d, c, 0, a)) {
      ph.irregular.push_back({d, c});
      continue;
    }
    for (int p = 0; p < d->plan.P; p++) {
      pass_args_regular(*d, c, p, a);
      const int la = d->plan.m[p] - kLgPerT;
      const bool last = p == d->plan.P - 1;
      const auto key = std::make_tuple(p, la, c.inverse ? 1 : 0, last ? 1 : 0, a.do_scale ? 1 : 0);
      auto f = gidx.find(key);
      if (f == gidx.end()) {
        RaggedGroup g;
        g.level = p;
        g.la = la;
        g.inv = c.inverse;
        g.last = last;
        g.scale = a.do_scale != 0;
        f = gidx.emplace(key, ph.groups.size()).first;
        ph.groups.push_back(g);
      }
      RaggedGroup& g = ph.groups[f->second];
      const int lgc = kLgTile - d->plan.m[p];
      g.seg_block0.push_back(g.nblocks);
      g.segs.push_back(a);
      g.nblocks += (uint32_t)((a.nlines + (1ull << lgc) - 1) >> lgc);
    }
  }
  // launch order: every group of depth 0, then depth 1, ...
  std::stable_sort(ph.groups.begin(), ph.groups.end(),
                   [](const RaggedGroup& x, const RaggedGroup& y) { return x.level < y.level; });
  for (auto& g : ph.groups) {
    g.d_segs = (PassArgs*)tb.get(g.segs.size() * sizeof(PassArgs));
    g.d_seg_block0 = (uint32_t*)tb.get(g.seg_block0.size() * sizeof(uint32_t));
  }
  if (n_eval) ph.negz_all = (fr*)bump.get(n_eval * sizeof(fr));
  return bump.ok() && tb.ok() ? hipSuccess : hipErrorOutOfMemory;
}

__global__ void neg_copy_kernel(const fr* z, fr* out) {
  if (threadIdx.x == 0) fr_store(out, fr_neg(fr_load(z)));
}

hipError_t phase_upload(PhasePlan& ph, hipStream_t s) {
  hipError_t e = hipSuccess;
  for (auto& g : ph.groups) {
    if (e == hipSuccess)
      e = hipMemcpyAsync(g.d_segs, g.segs.data(), g.segs.size() * sizeof(PassArgs), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(g.d_seg_block0, g.seg_block0.data(), g.seg_block0.size() * sizeof(uint32_t),
                         hipMemcpyHostToDevice, s);
  }
  return e;
}

static void launch_ragged(const RaggedGroup& g, hipStream_t s) {
  const uint32_t ns = (uint32_t)g.segs.size();
  switch (g.la) {
    case 0: launch_pass_ragged_la0(g.inv, g.last, g.scale, g.d_segs, g.d_seg_block0, ns, g.nblocks, s); break;
    case 1: launch_pass_ragged_la1(g.inv, g.last, g.scale, g.d_segs, g.d_seg_block0, ns, g.nblocks, s); break;
    case 2: launch_pass_ragged_la2(g.inv, g.last, g.scale, g.d_segs, g.d_seg_block0, ns, g.nblocks, s); break;
    case 3: launch_pass_ragged_la3(g.inv, g.last, g.scale, g.d_segs, g.d_seg_block0, ns, g.nblocks, s); break;
    default: launch_pass_ragged_la4(g.inv, g.last, g.scale, g.d_segs, g.d_seg_block0, ns, g.nblocks, s); break;
  }
}

static hipError_t perm_gp(const Item& it, const ScanJob& sc, hipStream_t s) {
  return perm_grand_product(it.w, it.sg, it.beta, it.gamma, sc.dom->pw0, sc.dom->pw1, it.k, it.out, sc.n, sc.scratch,
                            s);
}

// The transform launches go in before the scans.  Captured the other way
// round, a kernel trace of the launched graph showed the scan chain (short
// kernels, one after another) running almost to its end before the first
// ragged pass started (0.46 ms of a 2.7 ms batch); in this order the scans
// run beside the passes.
hipError_t phase_enqueue(const PhasePlan& ph, hipStream_t main, hipStream_t side, hipEvent_t fork, hipEvent_t join) {
  hipError_t e;
  if (!ph.scans.empty()) {
    if ((e = hipEventRecord(fork, main)) != hipSuccess) return e;
    if ((e = hipStreamWaitEvent(side, fork, 0)) != hipSuccess) return e;
  }
  for (auto& ir : ph.irregular)
    if ((e = ntt_launch(*ir.first, ir.second, main)) != hipSuccess) return e;
  for (auto& g : ph.groups) launch_ragged(g, main);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if (!ph.scans.empty()) {
    for (auto& sc : ph.scans) {
      const Item& it = ph.items[sc.item];
      e = sc.op == IT_EVAL_DIV ? eval_div(it.in, it.out, sc.n, it.z, it.y, sc.scratch, side) : perm_gp(it, sc, side);
      if (e != hipSuccess) return e;
    }
    if ((e = hipEventRecord(join, side)) != hipSuccess) return e;
    if ((e = hipStreamWaitEvent(main, join, 0)) != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t phase_enqueue_plain(const PhasePlan& ph, hipStream_t s) {
  hipError_t e;
  for (auto& pc : ph.plain_calls)
    if ((e = ntt_launch(*pc.first, pc.second, s)) != hipSuccess) return e;
  int k = 0;
  for (auto& sc : ph.scans) {
    const Item& it = ph.items[sc.item];
    if (sc.op == IT_PERM_GP) {
      if ((e = perm_gp(it, sc, s)) != hipSuccess) return e;
      continue;
    }
    // the reference's call pair: evaluate, then poly_div_poly by (X + c),
    // c = -z read from the item's point at run time (the caller may change
    // it between runs)
    e = evaluate(it.in, sc.n, it.z, it.y, sc.scratch, s);
    if (e == hipSuccess) e = hipMemcpyAsync(it.out, it.in, sc.n * sizeof(fr), hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(neg_copy_kernel, dim3(1), dim3(64), 0, s, it.z, ph.negz_all + k);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = div_linear(it.out, sc.n, ph.negz_all + k, nullptr, sc.scratch, s);
    k++;
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace pl
