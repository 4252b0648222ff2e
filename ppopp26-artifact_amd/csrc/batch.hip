// Batched form (C-ABI pl_batch_*): the uneven-size polynomials of one proof
// as one planned phase, replayed as a hipGraph.
//
// The plan (plan.hpp) groups the transforms' passes into ragged launches, so
// small polynomials share the chip instead of each leaving it mostly idle,
// and runs evaluation + division pairs as one suffix scan each on a forked
// stream.  All scratch and launch tables live in one device block sized by
// a dry run of the planner; a run allocates nothing.
#include <new>

#include "plan.hpp"

struct pl_batch {
  pl::PhasePlan plan;
  void* block = nullptr;
  size_t bytes = 0;
  int device = 0;
  int n_kernels = 0;
  hipStream_t s_main = nullptr, s_side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_a = nullptr, ev_b = nullptr;
  hipEvent_t ev_done = nullptr;  // recorded on the caller's stream after each run
  bool ran = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace pl {

static void batch_free(pl_batch& b) {
  // the batch's own work: its streams and the last run on a caller's stream
  // (not the whole device, which other callers' streams share)
  if (b.ran && b.ev_done) (void)hipEventSynchronize(b.ev_done);
  if (b.s_main) (void)hipStreamSynchronize(b.s_main);
  if (b.s_side) (void)hipStreamSynchronize(b.s_side);
  if (b.exec) (void)hipGraphExecDestroy(b.exec);
  if (b.graph) (void)hipGraphDestroy(b.graph);
  for (hipEvent_t ev : {b.ev_fork, b.ev_join, b.ev_a, b.ev_b, b.ev_done})
    if (ev) (void)hipEventDestroy(ev);
  if (b.s_main) (void)hipStreamDestroy(b.s_main);
  if (b.s_side) (void)hipStreamDestroy(b.s_side);
  if (b.block) (void)hipFree(b.block);
}

static hipError_t batch_build(pl_batch& b, const pl_batch_item* items, int n, bool graph) {
  std::vector<Item> its(n);
  for (int i = 0; i < n; i++) {
    its[i].op = items[i].op;  // pl_batch_op values equal the planner's
    its[i].lg = items[i].lg;
    its[i].lg_blowup = items[i].lg_blowup;
    its[i].in = (const fr*)items[i].d_in;
    its[i].out = (fr*)items[i].d_out;
    its[i].z = (const fr*)items[i].d_z;
    its[i].y = (fr*)items[i].d_y;
  }
  Bump dry;
  hipError_t e = phase_plan(b.plan, its, dry);
  if (e != hipSuccess) return e;
  b.bytes = dry.off;
  if ((e = hipMalloc(&b.block, b.bytes)) != hipSuccess) return e;
  Bump real;
  real.base = (char*)b.block;
  real.cap = b.bytes;
  if ((e = phase_plan(b.plan, its, real)) != hipSuccess) return e;
  if ((e = hipStreamCreateWithFlags(&b.s_main, hipStreamNonBlocking)) != hipSuccess) return e;
  if ((e = hipStreamCreateWithFlags(&b.s_side, hipStreamNonBlocking)) != hipSuccess) return e;
  for (hipEvent_t* ev : {&b.ev_fork, &b.ev_join, &b.ev_a, &b.ev_b, &b.ev_done})
    if ((e = hipEventCreateWithFlags(ev, hipEventDisableTiming)) != hipSuccess) return e;
  // the launch tables (and the domain tables the planner built on the null
  // stream) are in place before the first launch
  if ((e = phase_upload(b.plan, b.s_main)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(b.s_main)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(nullptr)) != hipSuccess) return e;
  // capture once: the node count is reported, and kept as the run form when
  // a graph was asked for
  if ((e = hipStreamBeginCapture(b.s_main, hipStreamCaptureModeThreadLocal)) != hipSuccess) return e;
  hipError_t ee = phase_enqueue(b.plan, b.s_main, b.s_side, b.ev_fork, b.ev_join);
  e = hipStreamEndCapture(b.s_main, &b.graph);
  if (ee != hipSuccess) return ee;
  if (e != hipSuccess) return e;
  size_t nn = 0;
  if ((e = hipGraphGetNodes(b.graph, nullptr, &nn)) != hipSuccess) return e;
  std::vector<hipGraphNode_t> nodes(nn);
  if (nn && (e = hipGraphGetNodes(b.graph, nodes.data(), &nn)) != hipSuccess) return e;
  for (auto nd : nodes) {
    hipGraphNodeType ty;
    if (hipGraphNodeGetType(nd, &ty) == hipSuccess && ty == hipGraphNodeTypeKernel) b.n_kernels++;
  }
  if (graph) return hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(b.graph);
  b.graph = nullptr;
  return hipSuccess;
}

static hipError_t batch_run(pl_batch& b, hipStream_t s) {
  hipError_t e;
  if (b.exec) {
    e = hipGraphLaunch(b.exec, s);
  } else {
    // the same launches, ordered after and before the caller's stream
    e = hipEventRecord(b.ev_a, s);
    if (e == hipSuccess) e = hipStreamWaitEvent(b.s_main, b.ev_a, 0);
    if (e == hipSuccess) e = phase_enqueue(b.plan, b.s_main, b.s_side, b.ev_fork, b.ev_join);
    if (e == hipSuccess) e = hipEventRecord(b.ev_b, b.s_main);
    if (e == hipSuccess) e = hipStreamWaitEvent(s, b.ev_b, 0);
  }
  if (e == hipSuccess) e = hipEventRecord(b.ev_done, s);
  if (e == hipSuccess) b.ran = true;
  return e;
}

static hipError_t batch_run_plain(pl_batch& b, hipStream_t s) {
  hipError_t e = phase_enqueue_plain(b.plan, s);
  if (e == hipSuccess) e = hipEventRecord(b.ev_done, s);
  if (e == hipSuccess) b.ran = true;
  return e;
}

}  // namespace pl

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

static int device_matches(const pl_batch* b) {
  int dev = -1;
  return hipGetDevice(&dev) == hipSuccess && dev == b->device;
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
  (void)hipGetDevice(&b->device);
  hipError_t e = pl::batch_build(*b, items, n_items, (flags & PL_BATCH_GRAPH) != 0);
  if (e != hipSuccess) {
    pl::batch_free(*b);
    delete b;
    return pl::hip_status(e);
  }
  *out = b;
  return PL_OK;
}

int pl_batch_run(pl_batch* b, pl_stream_t stream) {
  if (!b || !device_matches(b)) return PL_ERR_ARG;
  hipError_t e = pl::batch_run(*b, (hipStream_t)stream);
  return e == hipSuccess ? PL_OK : pl::hip_status(e);
}

int pl_batch_run_plain(pl_batch* b, pl_stream_t stream) {
  if (!b || !device_matches(b)) return PL_ERR_ARG;
  hipError_t e = pl::batch_run_plain(*b, (hipStream_t)stream);
  return e == hipSuccess ? PL_OK : pl::hip_status(e);
}

int pl_batch_info(const pl_batch* b, int* n_launches, size_t* bytes) {
  if (!b) return PL_ERR_ARG;
  if (n_launches) *n_launches = b->n_kernels;
  if (bytes) *bytes = b->bytes;
  return PL_OK;
}

void pl_batch_destroy(pl_batch* b) {
  if (!b) return;
  pl::batch_free(*b);
  delete b;
}

}  // extern "C"
