// Served form (C-ABI pl_server_*): a request loop for proof jobs with
// continuous batching, admission by HBM footprint and resident domains.
//
// A job is one proof's eight columns of N = 2^lg evaluations e_0..e_7 plus
// beta, gamma and an opening point zeta.  It computes
//   c_j = iNTT(e_j)                       coefficients          (8 x N)
//   L_j = LDE(c_j) over 7^(2^lb) <w_M>    M = N 2^lb             (8 x M)
//   zc  = exclusive grand product of the permutation ratio over the
//         witness columns e_0..e_2 and sigma columns e_3..e_5, k = (1, 7, 13)
//   y_j = c_j(zeta), q_j = (c_j - y_j) / (X - zeta), q_j[N-1] = 0
// into one server-owned output block, kept until the client releases it.
//
// Memory: the server takes one arena of device memory at creation and
// places every job's output block and every running group's work block
// (transform scratch, scan scratch) in it with a first-fit allocator; a job
// is admitted only when its blocks fit, strictly in arrival order.  Domains
// and coset tables are cached per size for the life of the process.
//
// Scheduling (batched mode): a worker thread admits jobs whenever one of
// the slots (a pair of streams) is free; at most one group holding a large
// job (transform work 8 N (1 + 2^lg_blowup) >= large_work) runs at a time,
// and small jobs may start ahead of large ones that wait only for that.  Consecutive small queued jobs are
// merged into one group; a group runs as phases planned by the phase planner:
// phase A = all iNTTs, phase B = all LDEs (ragged launches on the main
// stream) beside the evaluation/division and grand-product scans (side
// stream).  Jobs with large LDEs run their eight LDEs one after another
// sharing one scratch buffer.  Groups in different slots overlap on the
// device, so jobs that arrive while others run start as soon as a slot and
// memory are free.  Plain mode runs one job at a time on one stream as the
// per-call entry points would (the comparison form).
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>

#include "plan.hpp"

namespace pl {

// first-fit allocator over [0, cap), 256-byte granules
struct Arena {
  size_t cap = 0, used = 0, peak = 0;
  std::map<size_t, size_t> free_;  // offset -> length
  void init(size_t c) {
    cap = c & ~(size_t)255;
    free_.clear();
    free_[0] = cap;
    used = peak = 0;
  }
  static size_t round(size_t b) { return (b + 255) & ~(size_t)255; }
  bool alloc(size_t bytes, size_t* off) {
    bytes = round(bytes ? bytes : 1);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      *off = it->first;
      size_t rest = it->second - bytes, o = it->first + bytes;
      free_.erase(it);
      if (rest) free_[o] = rest;
      used += bytes;
      peak = std::max(peak, used);
      return true;
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    bytes = round(bytes ? bytes : 1);
    used -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
};

struct Job {
  uint64_t id = 0;
  pl_job_desc d;
  host::fr64 k[3];
  double t_arr = 0, t_admit = -1, t_done = -1;
  int status = PL_OK;
  bool done = false;
  bool holds_out = false;  // the output block [out_off, +out_bytes) is this job's until released
  size_t out_off = 0, out_bytes = 0;
  fr *coeffs = nullptr, *lde = nullptr, *z = nullptr, *y = nullptr, *q = nullptr;
  uint64_t n() const { return 1ull << d.lg; }
  uint64_t m() const { return 1ull << (d.lg + d.lg_blowup); }
  uint64_t work() const { return 8 * (n() + m()); }  // transform elements
};

static size_t job_out_bytes(int lg, int lb) {
  const uint64_t n = 1ull << lg, m = n << lb;
  return Arena::round(8 * n * 32) * 2 + Arena::round(8 * m * 32) + Arena::round(n * 32) + 256;
}

struct Slot {
  hipStream_t main = nullptr, side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr, done = nullptr;
  char* d_tables = nullptr;
  char* h_staging = nullptr;
  size_t tables_used = 0;
  bool busy = false;
  std::vector<Job*> jobs;
  std::vector<PhasePlan> phases;
  size_t blk_off = 0, blk_bytes = 0;
};

constexpr size_t kTablesCap = 1 << 20;

}  // namespace pl

struct pl_server {
  pl_server_config cfg;
  int device = 0;
  char* arena_base = nullptr;
  pl::Arena arena;
  std::vector<pl::Slot> slots;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<pl::Job*> queue;
  std::unordered_map<uint64_t, pl::Job*> jobs;
  uint64_t next_id = 1;
  uint64_t n_submitted = 0, n_completed = 0, n_groups = 0;
  bool stop = false;
  int fatal = PL_OK;
  std::thread worker;
  std::chrono::steady_clock::time_point t0;
  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace pl {

static constexpr uint64_t kSeqLdeElems = 1ull << 24;  // 8 M above this: LDEs one after another

// The phases of a group of jobs; transform scratch placed in `blk` (each
// phase from offset 0: phases run one after another), tables in the slot's
// table area.  Returns the group block size via *bytes.
static hipError_t group_plan(const pl_server& S, const std::vector<Job*>& jobs, Slot* slot, char* blk,
                             std::vector<PhasePlan>& phases, size_t* bytes) {
  const bool plain = S.cfg.mode == PL_SERVE_PLAIN;
  std::vector<std::vector<Item>> ph_items;
  // phase A: iNTTs
  std::vector<Item> a;
  for (Job* j : jobs)
    for (int c = 0; c < 8; c++) {
      Item it;
      it.op = IT_INTT;
      it.lg = j->d.lg;
      it.in = (const fr*)j->d.d_evals[c];
      it.out = j->coeffs + c * j->n();
      a.push_back(it);
    }
  ph_items.push_back(a);
  // phase B: LDEs beside the scans; jobs whose eight LDEs together are
  // large run LDE c of every such job in phase B_c (c = 1..7) instead, so
  // only one scratch buffer per job is live at a time
  std::vector<Item> b, scans;
  std::vector<std::vector<Item>> seq(7);
  for (Job* j : jobs) {
    for (int c = 0; c < 8; c++) {
      Item e;
      e.op = IT_EVAL_DIV;
      e.lg = j->d.lg;
      e.in = j->coeffs + c * j->n();
      e.out = j->q + c * j->n();
      e.z = (const fr*)j->d.d_zeta;
      e.y = j->y + c;
      scans.push_back(e);
    }
    Item g;
    g.op = IT_PERM_GP;
    g.lg = j->d.lg;
    for (int c = 0; c < 3; c++) {
      g.w[c] = (const fr*)j->d.d_evals[c];
      g.sg[c] = (const fr*)j->d.d_evals[3 + c];
      g.k[c] = j->k[c];
    }
    g.beta = (const fr*)j->d.d_beta;
    g.gamma = (const fr*)j->d.d_gamma;
    g.out = j->z;
    scans.push_back(g);
    const bool sequential = !plain && 8 * j->m() > kSeqLdeElems;
    for (int c = 0; c < 8; c++) {
      Item l;
      l.op = IT_LDE;
      l.lg = j->d.lg;
      l.lg_blowup = j->d.lg_blowup;
      l.in = j->coeffs + c * j->n();
      l.out = j->lde + c * j->m();
      if (sequential && c > 0) seq[c - 1].push_back(l);
      else b.push_back(l);
    }
  }
  b.insert(b.end(), scans.begin(), scans.end());
  ph_items.push_back(b);
  for (auto& s : seq)
    if (!s.empty()) ph_items.push_back(s);

  // Each phase places its scratch from offset 0 of the block (phases run one
  // after another, so their scratch may overlap); the block is the largest
  // phase.  Plain mode runs a phase's transforms one after another, so they
  // share one scratch buffer.
  phases.assign(ph_items.size(), PhasePlan());
  Bump tb;
  tb.base = slot ? slot->d_tables : nullptr;
  tb.cap = kTablesCap;
  size_t need = 0;
  for (size_t p = 0; p < ph_items.size(); p++) {
    Bump bb;
    bb.base = blk;
    auto& its = ph_items[p];
    if (plain) {
      size_t mx = 0;
      for (auto& it : its)
        if (it.op == IT_LDE || it.op == IT_INTT || it.op == IT_NTT)
          mx = std::max(mx, (size_t)32 << (it.lg + (it.op == IT_LDE ? it.lg_blowup : 0)));
      if (mx) {
        fr* t = (fr*)bb.get(mx);
        for (auto& it : its)
          if (it.op == IT_LDE || it.op == IT_INTT || it.op == IT_NTT) {
            it.use_shared_tmp = true;
            it.shared_tmp = t;
          }
      }
    }
    hipError_t e = phase_plan(phases[p], its, bb, &tb);
    if (e != hipSuccess) return e;
    need = std::max(need, bb.off);
  }
  if (slot) slot->tables_used = tb.off;
  *bytes = need;
  return hipSuccess;
}

static void slot_free(Slot& s) {
  for (hipEvent_t ev : {s.fork, s.join, s.done})
    if (ev) (void)hipEventDestroy(ev);
  if (s.main) (void)hipStreamDestroy(s.main);
  if (s.side) (void)hipStreamDestroy(s.side);
  if (s.d_tables) (void)hipFree(s.d_tables);
  if (s.h_staging) (void)hipHostFree(s.h_staging);
}

static hipError_t slot_init(Slot& s) {
  hipError_t e;
  if ((e = hipStreamCreateWithFlags(&s.main, hipStreamNonBlocking)) != hipSuccess) return e;
  if ((e = hipStreamCreateWithFlags(&s.side, hipStreamNonBlocking)) != hipSuccess) return e;
  for (hipEvent_t* ev : {&s.fork, &s.join, &s.done})
    if ((e = hipEventCreateWithFlags(ev, hipEventDisableTiming)) != hipSuccess) return e;
  if ((e = hipMalloc(&s.d_tables, kTablesCap)) != hipSuccess) return e;
  return hipHostMalloc((void**)&s.h_staging, kTablesCap, hipHostMallocDefault);
}

// place a job's outputs at arena offset off
static void job_place(pl_server& S, Job* j, size_t off) {
  j->out_off = off;
  char* p = S.arena_base + off;
  const uint64_t n = j->n(), m = j->m();
  j->coeffs = (fr*)p;
  p += Arena::round(8 * n * 32);
  j->q = (fr*)p;
  p += Arena::round(8 * n * 32);
  j->lde = (fr*)p;
  p += Arena::round(8 * m * 32);
  j->z = (fr*)p;
  p += Arena::round(n * 32);
  j->y = (fr*)p;
}

// copy the planned tables into the slot's staging buffer at their device offsets
static void stage_tables(Slot& s) {
  for (auto& ph : s.phases)
    for (auto& g : ph.groups) {
      memcpy(s.h_staging + ((char*)g.d_segs - s.d_tables), g.segs.data(), g.segs.size() * sizeof(PassArgs));
      memcpy(s.h_staging + ((char*)g.d_seg_block0 - s.d_tables), g.seg_block0.data(),
             g.seg_block0.size() * sizeof(uint32_t));
    }
}

static hipError_t slot_launch(pl_server& S, Slot& s) {
  hipError_t e = hipSuccess;
  if (s.tables_used) e = hipMemcpyAsync(s.d_tables, s.h_staging, s.tables_used, hipMemcpyHostToDevice, s.main);
  const bool plain = S.cfg.mode == PL_SERVE_PLAIN;
  for (auto& ph : s.phases) {
    if (e != hipSuccess) break;
    if (plain) {
      e = phase_enqueue_plain(ph, s.main);
    } else {
      e = phase_enqueue(ph, s.main, s.side, s.fork, s.join);
    }
  }
  if (e == hipSuccess) e = hipEventRecord(s.done, s.main);
  return e;
}

// Try to start the queue head (and, batched, the small jobs behind it) on a
// free slot.  Called with the lock held.  Returns false when nothing started.
static bool try_admit(pl_server& S) {
  if (S.queue.empty()) return false;
  if (S.fatal != PL_OK) {
    // the context reported an error: fail what is queued instead of launching
    for (Job* j : S.queue) {
      j->status = S.fatal;
      j->done = true;
      j->t_done = S.now();
    }
    S.queue.clear();
    S.cv_done.notify_all();
    return false;
  }
  Slot* slot = nullptr;
  for (auto& s : S.slots)
    if (!s.busy) {
      slot = &s;
      break;
    }
  if (!slot) return false;
  const bool plain = S.cfg.mode == PL_SERVE_PLAIN;
  // At most one group with a large job (one whose single transforms fill
  // the chip) runs at a time: two of them side by side only stretch each
  // other.  While one runs, small jobs may start ahead of queued large ones;
  // otherwise admission is in arrival order.
  bool large_running = false;
  for (auto& sl : S.slots)
    if (sl.busy)
      for (Job* j : sl.jobs) large_running |= j->work() >= S.cfg.large_work;
  size_t first = 0;
  if (!plain && large_running) {
    while (first < S.queue.size() && S.queue[first]->work() >= S.cfg.large_work) first++;
    if (first == S.queue.size()) return false;
  }
  // candidate group: that job, then following small jobs
  std::vector<Job*> cand{S.queue[first]};
  auto small = [&](const Job* j) { return 8 * j->n() <= S.cfg.group_elems; };
  if (!plain && small(cand[0])) {
    uint64_t elems = 8 * cand[0]->n();
    for (size_t i = first + 1; i < S.queue.size() && (int)cand.size() < S.cfg.max_group_jobs; i++) {
      Job* j = S.queue[i];
      if (!small(j) || elems + 8 * j->n() > S.cfg.group_elems) break;
      elems += 8 * j->n();
      cand.push_back(j);
    }
  }
  while (!cand.empty()) {
    // place outputs, then size and place the work block
    std::vector<size_t> offs;
    bool fit = true;
    for (Job* j : cand) {
      size_t off;
      if (!S.arena.alloc(j->out_bytes, &off)) {
        fit = false;
        break;
      }
      offs.push_back(off);
      job_place(S, j, off);
    }
    size_t blk_bytes = 0, blk_off = 0;
    std::vector<PhasePlan> dry;
    hipError_t e = hipSuccess;
    if (fit) {
      e = group_plan(S, cand, nullptr, nullptr, dry, &blk_bytes);
      if (e == hipSuccess) fit = S.arena.alloc(blk_bytes, &blk_off);
      else fit = false;
    }
    if (fit) {
      // the placed plan has the dry run's layout; the block keeps its size
      size_t placed = 0;
      e = group_plan(S, cand, slot, S.arena_base + blk_off, slot->phases, &placed);
      if (e == hipSuccess && placed != blk_bytes) e = hipErrorInvalidValue;
      if (e == hipSuccess) {
        stage_tables(*slot);
        const double t = S.now();
        for (Job* j : cand) {
          j->t_admit = t;
          j->holds_out = true;
        }
        e = slot_launch(S, *slot);
      }
      if (e != hipSuccess) {
        // a launch error is not recoverable for these jobs
        S.fatal = hip_status(e);
        for (Job* j : cand) {
          j->status = S.fatal;
          j->done = true;
          j->holds_out = false;  // released just below
          j->t_done = S.now();
        }
        S.arena.release(blk_off, blk_bytes);
        for (size_t i = 0; i < cand.size(); i++) S.arena.release(offs[i], cand[i]->out_bytes);
        S.queue.erase(S.queue.begin() + first, S.queue.begin() + first + cand.size());
        S.cv_done.notify_all();
        return true;
      }
      slot->busy = true;
      slot->jobs = cand;
      slot->blk_off = blk_off;
      slot->blk_bytes = blk_bytes;
      S.queue.erase(S.queue.begin() + first, S.queue.begin() + first + cand.size());
      S.n_groups++;
      return true;
    }
    for (size_t i = 0; i < offs.size(); i++) S.arena.release(offs[i], cand[i]->out_bytes);
    if (cand.size() == 1) break;
    cand.pop_back();  // try a smaller group
  }
  if (first != 0) return false;  // a small job ahead of large ones waits for memory
  // The head does not fit now.  It waits for running groups to finish and
  // for clients to release finished jobs, unless it could not fit even in
  // an empty arena (or cannot be planned at all): then it fails.
  Job* j = S.queue.front();
  std::vector<PhasePlan> dry;
  size_t blk = 0;
  hipError_t e = group_plan(S, {j}, nullptr, nullptr, dry, &blk);
  int st = PL_OK;
  if (e != hipSuccess) st = hip_status(e);
  else if (Arena::round(j->out_bytes) + Arena::round(blk) > S.arena.cap) st = PL_ERR_NOMEM;
  if (st == PL_OK) return false;
  j->status = st;
  j->done = true;
  j->t_done = S.now();
  S.queue.pop_front();
  S.cv_done.notify_all();
  return true;
}

static void worker_main(pl_server* Sp) {
  pl_server& S = *Sp;
  (void)hipSetDevice(S.device);
  std::unique_lock<std::mutex> lk(S.mu);
  while (true) {
    // completions
    bool any_busy = false;
    for (auto& s : S.slots) {
      if (!s.busy) continue;
      hipError_t q = hipEventQuery(s.done);
      if (q == hipErrorNotReady) {
        any_busy = true;
        continue;
      }
      const double t = S.now();
      const int st = q == hipSuccess ? PL_OK : hip_status(q);
      // an error reported at completion leaves the context unusable: admit
      // nothing more (submit and later admissions report it)
      if (st != PL_OK && S.fatal == PL_OK) S.fatal = st;
      for (Job* j : s.jobs) {
        j->t_done = t;
        j->status = st;
        j->done = true;
        S.n_completed++;
      }
      S.arena.release(s.blk_off, s.blk_bytes);
      s.jobs.clear();
      s.phases.clear();
      s.busy = false;
      S.cv_done.notify_all();
    }
    // admissions
    while (try_admit(S)) {
    }
    any_busy = false;
    for (auto& s : S.slots) any_busy |= s.busy;
    if (S.stop && !any_busy && S.queue.empty()) break;
    if (!any_busy && S.queue.empty()) {
      S.cv_work.wait(lk);
      continue;
    }
    // poll: release the lock briefly
    lk.unlock();
    std::this_thread::sleep_for(std::chrono::microseconds(20));
    lk.lock();
  }
}

}  // namespace pl

extern "C" {

size_t pl_job_bytes(int lg, int lg_blowup) {
  if (lg < 1 || lg_blowup < 1 || lg_blowup > 4 || lg + lg_blowup > 28) return 0;
  return pl::job_out_bytes(lg, lg_blowup);
}

int pl_server_create(const pl_server_config* cfg, pl_server** out) {
  if (!out) return PL_ERR_ARG;
  *out = nullptr;
  pl_server_config c{};
  if (cfg) c = *cfg;
  if (c.mode != PL_SERVE_BATCHED && c.mode != PL_SERVE_PLAIN) return PL_ERR_ARG;
  if (c.slots < 0 || c.slots > 8 || c.max_group_jobs < 0) return PL_ERR_ARG;
  if (c.slots == 0) c.slots = 2;  // two stream pairs: the process's four hardware queues
  if (c.mode == PL_SERVE_PLAIN) c.slots = 1;
  if (c.max_group_jobs == 0) c.max_group_jobs = 16;
  if (c.group_elems == 0) c.group_elems = 1ull << 22;
  if (c.large_work == 0) c.large_work = 1ull << 25;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return PL_ERR_NO_DEVICE;
  if (c.hbm_bytes == 0) {
    size_t fr_ = 0, tot = 0;
    if (hipMemGetInfo(&fr_, &tot) != hipSuccess) return PL_ERR_HIP;
    // leave room for the domain caches: coset tables (2^(lg + lg_blowup)
    // elements per size), NTT twiddles (~2 N per size) and windows
    const size_t reserve = (size_t)32 << 30;
    c.hbm_bytes = fr_ > 2 * reserve ? fr_ - reserve : fr_ / 2;
  }
  pl_server* S = new (std::nothrow) pl_server();
  if (!S) return PL_ERR_NOMEM;
  S->cfg = c;
  (void)hipGetDevice(&S->device);
  hipError_t e = hipMalloc((void**)&S->arena_base, c.hbm_bytes);
  if (e == hipSuccess) {
    S->arena.init(c.hbm_bytes);
    S->slots.resize(c.slots);
    for (auto& s : S->slots)
      if (e == hipSuccess) e = pl::slot_init(s);
  }
  if (e != hipSuccess) {
    for (auto& s : S->slots) pl::slot_free(s);
    if (S->arena_base) (void)hipFree(S->arena_base);
    delete S;
    return pl::hip_status(e);
  }
  S->t0 = std::chrono::steady_clock::now();
  S->worker = std::thread(pl::worker_main, S);
  *out = S;
  return PL_OK;
}

double pl_server_now(const pl_server* S) { return S ? S->now() : -1.0; }

int pl_server_submit(pl_server* S, const pl_job_desc* d, uint64_t* job_id) {
  if (!S || !d || !job_id) return PL_ERR_ARG;
  if (d->lg < 1 || d->lg_blowup < 1) return PL_ERR_ARG;
  if (d->lg_blowup > 4 || d->lg + d->lg_blowup > 28) return PL_ERR_UNSUPPORTED;
  for (int c = 0; c < 8; c++)
    if (!d->d_evals[c] || ((uintptr_t)d->d_evals[c] & 15)) return PL_ERR_ARG;
  for (const void* p : {d->d_beta, d->d_gamma, d->d_zeta})
    if (!p || ((uintptr_t)p & 15)) return PL_ERR_ARG;
  pl::Job* j = new (std::nothrow) pl::Job();
  if (!j) return PL_ERR_NOMEM;
  j->d = *d;
  j->k[0] = pl::host::from_u64(1);
  j->k[1] = pl::host::from_u64(7);
  j->k[2] = pl::host::from_u64(13);
  j->out_bytes = pl::job_out_bytes(d->lg, d->lg_blowup);
  std::lock_guard<std::mutex> g(S->mu);
  if (S->fatal != PL_OK) {
    delete j;
    return S->fatal;
  }
  j->id = S->next_id++;
  j->t_arr = S->now();
  S->jobs[j->id] = j;
  S->queue.push_back(j);
  S->n_submitted++;
  *job_id = j->id;
  S->cv_work.notify_one();
  return PL_OK;
}

int pl_server_wait(pl_server* S, uint64_t job_id, pl_job_result* res) {
  if (!S) return PL_ERR_ARG;
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->jobs.find(job_id);
  if (it == S->jobs.end()) return PL_ERR_ARG;
  pl::Job* j = it->second;
  S->cv_done.wait(lk, [&] { return j->done; });
  if (res) {
    res->status = j->status;
    res->arrival_s = j->t_arr;
    res->admit_s = j->t_admit;
    res->done_s = j->t_done;
    const bool ok = j->status == PL_OK;
    res->d_coeffs = ok ? j->coeffs : nullptr;
    res->d_lde = ok ? j->lde : nullptr;
    res->d_z = ok ? j->z : nullptr;
    res->d_y = ok ? j->y : nullptr;
    res->d_q = ok ? j->q : nullptr;
  }
  return j->status;
}

int pl_server_release(pl_server* S, uint64_t job_id) {
  if (!S) return PL_ERR_ARG;
  std::lock_guard<std::mutex> g(S->mu);
  auto it = S->jobs.find(job_id);
  if (it == S->jobs.end() || !it->second->done) return PL_ERR_ARG;
  pl::Job* j = it->second;
  if (j->holds_out) S->arena.release(j->out_off, j->out_bytes);
  S->jobs.erase(it);
  delete j;
  S->cv_work.notify_one();  // memory freed: admission may proceed
  return PL_OK;
}

int pl_server_get_stats(pl_server* S, pl_server_stats* st) {
  if (!S || !st) return PL_ERR_ARG;
  std::lock_guard<std::mutex> g(S->mu);
  st->submitted = S->n_submitted;
  st->completed = S->n_completed;
  st->groups = S->n_groups;
  st->queued = S->queue.size();
  st->arena_bytes = S->arena.cap;
  st->arena_peak = S->arena.peak;
  st->arena_used = S->arena.used;
  st->domain_bytes = pl::cached_domain_bytes();
  return PL_OK;
}

void pl_server_destroy(pl_server* S) {
  if (!S) return;
  {
    std::lock_guard<std::mutex> g(S->mu);
    S->stop = true;
    S->cv_work.notify_all();
  }
  S->worker.join();
  // the worker leaves only when no slot is busy; wait for the slots' own
  // streams, not the whole device
  for (auto& s : S->slots) {
    if (s.main) (void)hipStreamSynchronize(s.main);
    if (s.side) (void)hipStreamSynchronize(s.side);
  }
  for (auto& kv : S->jobs) delete kv.second;
  for (auto& s : S->slots) pl::slot_free(s);
  (void)hipFree(S->arena_base);
  delete S;
}

}  // extern "C"
