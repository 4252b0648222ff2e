// Batched form: the uneven-size polynomials of one proof in a few fat
// launches, replayed as one hipGraph.
//
// A batch is planned once: each NTT / iNTT / LDE item is split into its
// passes; passes at the same depth with the same column radix, direction and
// last-pass flag become segments of ONE ragged launch (a block-to-segment
// table picks each block's polynomial and tile), so small polynomials share
// the chip instead of each leaving it mostly idle.  Evaluation + division
// items are fused into one affine suffix scan each (quotient of p(X) by
// (X - z) plus the remainder p(z)) and run on a forked stream, overlapping
// the transforms.  The whole sequence is captured into a hipGraph and
// replayed by a single hipGraphLaunch.  Every buffer and table is allocated
// at planning time; a run allocates nothing.
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <new>
#include <vector>

#include "ntt_pass.hpp"

struct pl_batch;

namespace pl {

// ---- library-wide domain caches (batched and served forms) ----------------
static std::mutex g_cache_mu;
// keyed by (device, lg) and (device, lg_n, lg_blowup); held until the
// process ends or cache_clear()
static std::map<std::pair<int, int>, Domain*> g_domains;
static std::map<std::tuple<int, int, int>, LdeDomain*> g_lde;

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

LdeDomain* cached_lde_domain(int lg_n, int lg_blowup, hipError_t* err) {
  std::lock_guard<std::mutex> g(g_cache_mu);
  const auto key = std::make_tuple(current_device(), lg_n, lg_blowup);
  auto it = g_lde.find(key);
  if (it != g_lde.end()) return it->second;
  LdeDomain* d = new LdeDomain();
  host::fr64 shift = host::pow_u64(host::from_u64(7), 1ull << lg_blowup);
  hipError_t e = lde_domain_build(*d, lg_n, lg_blowup, shift, nullptr);
  if (e != hipSuccess) {
    lde_domain_free(*d);
    delete d;
    if (err) *err = e;
    return nullptr;
  }
  g_lde[key] = d;
  return d;
}

size_t cached_bytes() {
  std::lock_guard<std::mutex> g(g_cache_mu);
  size_t b = 0;
  for (auto& kv : g_domains) b += kv.second->bytes;
  for (auto& kv : g_lde) b += kv.second->bytes + kv.second->ntt.bytes;
  return b;
}

void cache_clear() {
  std::lock_guard<std::mutex> g(g_cache_mu);
  for (auto& kv : g_domains) {
    domain_free(*kv.second);
    delete kv.second;
  }
  for (auto& kv : g_lde) {
    lde_domain_free(*kv.second);
    delete kv.second;
  }
  g_domains.clear();
  g_lde.clear();
}

struct RaggedGroup {
  int level = 0, la = 0;
  bool inv = false, last = false;
  std::vector<PassArgs> segs;
  std::vector<uint32_t> tiles;
  PassArgs* d_segs = nullptr;
  uint32_t* d_block_seg = nullptr;
  uint32_t* d_seg_block0 = nullptr;
  uint32_t nblocks = 0;
};

struct BatchPlan {
  std::vector<pl_batch_item> items;
  std::vector<NttCall> irregular;           // tiny sizes: plain launches
  std::vector<const Domain*> irregular_dom;
  std::vector<RaggedGroup> groups;
  struct Scan {
    const fr* in;
    fr* q;
    uint64_t n;
    const fr* z;
    fr* y;
    fr* scratch;
    fr* negz;  // -z for the plain form's poly_div_poly
  };
  std::vector<Scan> scans;
  // plain form: one call sequence per item
  struct Plain {
    int op;
    const Domain* d;
    const LdeDomain* ld;
    NttCall c;
  };
  std::vector<Plain> plain;
  std::vector<void*> allocs;
  size_t alloc_bytes = 0;
  int n_kernels = 0;
  int device = 0;
  hipStream_t s_main = nullptr, s_side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

static void launch_ragged(const RaggedGroup& g, hipStream_t s) {
  switch (g.la) {
    case 0: launch_pass_ragged_la0(g.inv, g.last, g.d_segs, g.d_block_seg, g.d_seg_block0, g.nblocks, s); break;
    case 1: launch_pass_ragged_la1(g.inv, g.last, g.d_segs, g.d_block_seg, g.d_seg_block0, g.nblocks, s); break;
    case 2: launch_pass_ragged_la2(g.inv, g.last, g.d_segs, g.d_block_seg, g.d_seg_block0, g.nblocks, s); break;
    case 3: launch_pass_ragged_la3(g.inv, g.last, g.d_segs, g.d_block_seg, g.d_seg_block0, g.nblocks, s); break;
    default: launch_pass_ragged_la4(g.inv, g.last, g.d_segs, g.d_block_seg, g.d_seg_block0, g.nblocks, s); break;
  }
}

// enqueue the batched work on s_main (+ s_side for the scans)
static hipError_t enqueue_batched(BatchPlan& b) {
  hipError_t e;
  if (!b.scans.empty()) {
    if ((e = hipEventRecord(b.ev_fork, b.s_main)) != hipSuccess) return e;
    if ((e = hipStreamWaitEvent(b.s_side, b.ev_fork, 0)) != hipSuccess) return e;
    for (auto& sc : b.scans)
      if ((e = eval_div(sc.in, sc.q, sc.n, sc.z, sc.y, sc.scratch, b.s_side)) != hipSuccess) return e;
  }
  for (size_t i = 0; i < b.irregular.size(); i++)
    if ((e = ntt_launch(*b.irregular_dom[i], b.irregular[i], b.s_main)) != hipSuccess) return e;
  for (auto& g : b.groups) launch_ragged(g, b.s_main);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if (!b.scans.empty()) {
    if ((e = hipEventRecord(b.ev_join, b.s_side)) != hipSuccess) return e;
    if ((e = hipStreamWaitEvent(b.s_main, b.ev_join, 0)) != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t batch_run_plain(BatchPlan& b, hipStream_t s) {
  for (auto& p : b.plain) {
    hipError_t e = ntt_launch(p.ld ? p.ld->ntt : *p.d, p.c, s);
    if (e != hipSuccess) return e;
  }
  for (auto& sc : b.scans) {
    hipError_t e = evaluate(sc.in, sc.n, sc.z, sc.y, sc.scratch, s);
    if (e == hipSuccess) e = hipMemcpyAsync(sc.q, sc.in, sc.n * sizeof(fr), hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = div_linear(sc.q, sc.n, sc.negz, nullptr, sc.scratch, s);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

__global__ void neg_copy_kernel(const fr* z, fr* out) {
  if (threadIdx.x == 0) fr_store(out, fr_neg(fr_load(z)));
}

static void* balloc(BatchPlan& b, size_t bytes, hipError_t& e) {
  void* p = nullptr;
  if (e == hipSuccess) e = hipMalloc(&p, bytes ? bytes : 32);
  if (e == hipSuccess) {
    b.allocs.push_back(p);
    b.alloc_bytes += bytes;
  }
  return p;
}

hipError_t batch_plan(BatchPlan& b, const pl_batch_item* items, int n, bool graph) {
  hipError_t e = hipSuccess;
  b.items.assign(items, items + n);
  std::map<std::tuple<int, int, int, int>, size_t> gidx;
  for (int i = 0; i < n && e == hipSuccess; i++) {
    const pl_batch_item& it = items[i];
    if (it.op == PL_BATCH_EVAL_DIV) {
      BatchPlan::Scan sc;
      sc.in = (const fr*)it.d_in;
      sc.q = (fr*)it.d_out;
      sc.n = 1ull << it.lg;
      sc.z = (const fr*)it.d_z;
      sc.y = (fr*)it.d_y;
      sc.scratch = (fr*)balloc(b, scan_scratch_elems(sc.n) * sizeof(fr), e);
      sc.negz = (fr*)balloc(b, sizeof(fr), e);
      if (e == hipSuccess) hipLaunchKernelGGL(neg_copy_kernel, dim3(1), dim3(64), 0, nullptr, sc.z, sc.negz);
      b.scans.push_back(sc);
      continue;
    }
    NttCall c;
    const Domain* d = nullptr;
    const LdeDomain* ld = nullptr;
    c.in = (const fr*)it.d_in;
    c.out = (fr*)it.d_out;
    if (it.op == PL_BATCH_LDE) {
      ld = cached_lde_domain(it.lg, it.lg_blowup, &e);
      if (!ld) break;
      d = &ld->ntt;
      c.in_pstride = 0;
      c.npoly = 1u << it.lg_blowup;
      c.premul = ld->cp;
      c.pm_mask = c.npoly - 1;
      c.lgI = it.lg_blowup;
    } else {
      d = cached_domain(it.lg, &e);
      if (!d) break;
      c.in_pstride = 1ull << it.lg;
      c.npoly = 1;
      c.inverse = it.op == PL_BATCH_INTT;
    }
    c.tmp = (fr*)balloc(b, ((size_t)c.npoly << it.lg) * sizeof(fr), e);
    b.plain.push_back({it.op, it.op == PL_BATCH_LDE ? nullptr : d, ld, c});
    PassArgs a;
    if (!pass_args_regular(*d, c, 0, a)) {
      b.irregular.push_back(c);
      b.irregular_dom.push_back(d);
      continue;
    }
    for (int p = 0; p < d->plan.P; p++) {
      pass_args_regular(*d, c, p, a);
      const int la = d->plan.m[p] - 4;
      const bool last = p == d->plan.P - 1;
      auto key = std::make_tuple(p, la, c.inverse ? 1 : 0, last ? 1 : 0);
      auto f = gidx.find(key);
      if (f == gidx.end()) {
        RaggedGroup g;
        g.level = p;
        g.la = la;
        g.inv = c.inverse;
        g.last = last;
        gidx[key] = b.groups.size();
        b.groups.push_back(g);
        f = gidx.find(key);
      }
      RaggedGroup& g = b.groups[f->second];
      const int lgc = kLgTile - d->plan.m[p];
      g.segs.push_back(a);
      g.tiles.push_back((uint32_t)((a.nlines + (1ull << lgc) - 1) >> lgc));
    }
  }
  if (e != hipSuccess) return e;
  // pass order: all groups of depth 0, then depth 1, ...
  std::stable_sort(b.groups.begin(), b.groups.end(),
                   [](const RaggedGroup& x, const RaggedGroup& y) { return x.level < y.level; });
  for (auto& g : b.groups) {
    std::vector<uint32_t> block_seg, seg_block0;
    for (size_t sgi = 0; sgi < g.segs.size(); sgi++) {
      seg_block0.push_back((uint32_t)block_seg.size());
      block_seg.insert(block_seg.end(), g.tiles[sgi], (uint32_t)sgi);
    }
    g.nblocks = (uint32_t)block_seg.size();
    g.d_segs = (PassArgs*)balloc(b, g.segs.size() * sizeof(PassArgs), e);
    g.d_block_seg = (uint32_t*)balloc(b, block_seg.size() * 4, e);
    g.d_seg_block0 = (uint32_t*)balloc(b, seg_block0.size() * 4, e);
    if (e == hipSuccess) e = hipMemcpy(g.d_segs, g.segs.data(), g.segs.size() * sizeof(PassArgs), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(g.d_block_seg, block_seg.data(), block_seg.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(g.d_seg_block0, seg_block0.data(), seg_block0.size() * 4, hipMemcpyHostToDevice);
  }
  if (e != hipSuccess) return e;
  if ((e = hipStreamCreateWithFlags(&b.s_main, hipStreamNonBlocking)) != hipSuccess) return e;
  if ((e = hipStreamCreateWithFlags(&b.s_side, hipStreamNonBlocking)) != hipSuccess) return e;
  if ((e = hipEventCreateWithFlags(&b.ev_fork, hipEventDisableTiming)) != hipSuccess) return e;
  if ((e = hipEventCreateWithFlags(&b.ev_join, hipEventDisableTiming)) != hipSuccess) return e;
  if ((e = hipDeviceSynchronize()) != hipSuccess) return e;
  if ((e = hipStreamBeginCapture(b.s_main, hipStreamCaptureModeThreadLocal)) != hipSuccess) return e;
  hipError_t ee = enqueue_batched(b);
  e = hipStreamEndCapture(b.s_main, &b.graph);
  if (ee != hipSuccess) return ee;
  if (e != hipSuccess) return e;
  size_t nn = 0;
  if ((e = hipGraphGetNodes(b.graph, nullptr, &nn)) != hipSuccess) return e;
  std::vector<hipGraphNode_t> nodes(nn);
  if (nn && (e = hipGraphGetNodes(b.graph, nodes.data(), &nn)) != hipSuccess) return e;
  b.n_kernels = 0;
  for (auto nd : nodes) {
    hipGraphNodeType ty;
    if (hipGraphNodeGetType(nd, &ty) == hipSuccess && ty == hipGraphNodeTypeKernel) b.n_kernels++;
  }
  if (graph) {
    if ((e = hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0)) != hipSuccess) return e;
  } else {
    (void)hipGraphDestroy(b.graph);
    b.graph = nullptr;
  }
  return hipSuccess;
}

hipError_t batch_run(BatchPlan& b, hipStream_t s) {
  if (b.exec) return hipGraphLaunch(b.exec, s);
  // no graph: the same launches, ordered after the caller's stream
  hipError_t e = hipEventRecord(b.ev_join, s);
  if (e == hipSuccess) e = hipStreamWaitEvent(b.s_main, b.ev_join, 0);
  if (e == hipSuccess) e = enqueue_batched(b);
  if (e == hipSuccess) e = hipEventRecord(b.ev_fork, b.s_main);
  if (e == hipSuccess) e = hipStreamWaitEvent(s, b.ev_fork, 0);
  return e;
}

void batch_free(BatchPlan& b) {
  if (b.exec) (void)hipGraphExecDestroy(b.exec);
  if (b.graph) (void)hipGraphDestroy(b.graph);
  if (b.ev_fork) (void)hipEventDestroy(b.ev_fork);
  if (b.ev_join) (void)hipEventDestroy(b.ev_join);
  if (b.s_main) (void)hipStreamDestroy(b.s_main);
  if (b.s_side) (void)hipStreamDestroy(b.s_side);
  for (void* p : b.allocs) (void)hipFree(p);
  b.allocs.clear();
}

}  // namespace pl

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
struct pl_batch {
  pl::BatchPlan plan;
};

static int batch_item_check(const pl_batch_item& it) {
  if (!it.d_in || !it.d_out || ((uintptr_t)it.d_in & 15) || ((uintptr_t)it.d_out & 15)) return PL_ERR_ARG;
  switch (it.op) {
    case PL_BATCH_NTT:
    case PL_BATCH_INTT:
      if (it.lg < 1) return PL_ERR_ARG;
      return it.lg > 28 ? PL_ERR_UNSUPPORTED : PL_OK;
    case PL_BATCH_LDE:
      if (it.lg < 1 || it.lg_blowup < 1) return PL_ERR_ARG;
      return (it.lg_blowup > 4 || it.lg + it.lg_blowup > 28) ? PL_ERR_UNSUPPORTED : PL_OK;
    case PL_BATCH_EVAL_DIV:
      if (it.lg < 0 || !it.d_z || !it.d_y || it.d_in == it.d_out) return PL_ERR_ARG;
      if (((uintptr_t)it.d_z & 15) || ((uintptr_t)it.d_y & 15)) return PL_ERR_ARG;
      return it.lg > 28 ? PL_ERR_UNSUPPORTED : PL_OK;
    default: return PL_ERR_ARG;
  }
}

extern "C" {

int pl_batch_create(const pl_batch_item* items, int n_items, int flags, pl_batch** out) {
  if (!out || !items || n_items < 1 || (flags & ~PL_BATCH_GRAPH)) return PL_ERR_ARG;
  *out = nullptr;
  for (int i = 0; i < n_items; i++) {
    int st = batch_item_check(items[i]);
    if (st != PL_OK) return st;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return PL_ERR_NO_DEVICE;
  pl_batch* b = new (std::nothrow) pl_batch();
  if (!b) return PL_ERR_NOMEM;
  hipGetDevice(&b->plan.device);
  hipError_t e = pl::batch_plan(b->plan, items, n_items, (flags & PL_BATCH_GRAPH) != 0);
  if (e != hipSuccess) {
    pl::batch_free(b->plan);
    delete b;
    return pl::hip_status(e);
  }
  *out = b;
  return PL_OK;
}

int pl_batch_run(pl_batch* b, pl_stream_t stream) {
  if (!b) return PL_ERR_ARG;
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev != b->plan.device) return PL_ERR_ARG;
  hipError_t e = pl::batch_run(b->plan, (hipStream_t)stream);
  return e == hipSuccess ? PL_OK : pl::hip_status(e);
}

int pl_batch_run_plain(pl_batch* b, pl_stream_t stream) {
  if (!b) return PL_ERR_ARG;
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev != b->plan.device) return PL_ERR_ARG;
  hipError_t e = pl::batch_run_plain(b->plan, (hipStream_t)stream);
  return e == hipSuccess ? PL_OK : pl::hip_status(e);
}

int pl_batch_info(const pl_batch* b, int* n_launches, size_t* bytes) {
  if (!b) return PL_ERR_ARG;
  if (n_launches) *n_launches = b->plan.n_kernels;
  if (bytes) *bytes = b->plan.alloc_bytes;
  return PL_OK;
}

void pl_batch_destroy(pl_batch* b) {
  if (!b) return;
  (void)hipDeviceSynchronize();
  pl::batch_free(b->plan);
  delete b;
}

}  // extern "C"
