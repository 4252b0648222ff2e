// Phase planner shared by the batched and served forms.
//
// A phase is a set of independent work items (transforms, evaluation +
// division pairs, permutation grand products).  Planning a phase splits the
// transforms into their passes and groups passes of the same depth, column
// radix, direction and last-pass flag into ragged launches; the scans go on
// a side stream.  Every buffer a phase needs (launch tables, transform
// scratch, scan scratch) comes from a bump allocator over one caller-provided
// region, so a plan can first be sized with a dry run and then placed in a
// block of a preallocated arena: running a plan allocates nothing.
#pragma once
#include <string>
#include <vector>

#include "ntt_pass.hpp"

namespace pl {

// domains shared by every batch and server of the process, keyed by device
Domain* cached_domain(int lg, hipError_t* err);
// coset factor table of the reference's LDE (shift 7^(2^lg_blowup)); the
// transform itself uses cached_domain(lg_n)
const fr* cached_coset_table(int lg_n, int lg_blowup, hipError_t* err);
size_t cached_domain_bytes();

enum ItemOp { IT_NTT = 0, IT_INTT = 1, IT_LDE = 2, IT_EVAL_DIV = 3, IT_PERM_GP = 4 };

struct Item {
  int op = IT_NTT;
  int lg = 0, lg_blowup = 0;
  const fr* in = nullptr;
  fr* out = nullptr;
  const fr* z = nullptr;  // EVAL_DIV point
  fr* y = nullptr;        // EVAL_DIV value
  // PERM_GP: witness and sigma columns, beta, gamma, coset constants; out = z column
  const fr* w[3] = {};
  const fr* sg[3] = {};
  const fr* beta = nullptr;
  const fr* gamma = nullptr;
  host::fr64 k[3];
  bool use_shared_tmp = false;  // transform scratch provided by the caller (else planned)
  fr* shared_tmp = nullptr;
};

// bump allocation over [base, base + cap); base == nullptr sizes only
struct Bump {
  char* base = nullptr;
  size_t off = 0, cap = ~(size_t)0;
  void* get(size_t bytes) {
    off = (off + 255) & ~(size_t)255;
    void* p = base ? base + off : nullptr;
    off += bytes ? bytes : 32;
    return p;
  }
  bool ok() const { return off <= cap; }
};

struct RaggedGroup {
  int level = 0, la = 0;
  bool inv = false, last = false, scale = false;
  std::vector<PassArgs> segs;
  std::vector<uint32_t> seg_block0;
  PassArgs* d_segs = nullptr;
  uint32_t* d_seg_block0 = nullptr;
  uint32_t nblocks = 0;
};

struct ScanJob {
  int op;   // IT_EVAL_DIV or IT_PERM_GP
  int item;  // index into PhasePlan::items
  const Domain* dom;  // PERM_GP: omega windows
  fr* scratch;
  uint64_t n;
};

struct PhasePlan {
  std::vector<Item> items;
  std::vector<RaggedGroup> groups;
  std::vector<std::pair<const Domain*, NttCall>> irregular;  // tiny / single-pass premultiplied forms
  std::vector<ScanJob> scans;
  std::vector<std::pair<const Domain*, NttCall>> plain_calls;  // per-item form of the transforms
  fr* negz_all = nullptr;  // -z per EVAL_DIV item (plain form; written at run time from the item's z)
};

// Plan one phase; with bump.base == nullptr only sizes (bump.off = bytes
// needed).  Launch tables come from `tables` when given, else from `bump`.
// Domains are looked up (and built) in the shared cache.
hipError_t phase_plan(PhasePlan& ph, const std::vector<Item>& items, Bump& bump, Bump* tables = nullptr);
// Copy the plan's launch tables to the device (stream-ordered on s); the
// host copies stay in the plan.
hipError_t phase_upload(PhasePlan& ph, hipStream_t s);
// Enqueue the batched form: ragged launches on `main`, scans on `side`
// (forked and joined with the two events); no allocation.
hipError_t phase_enqueue(const PhasePlan& ph, hipStream_t main, hipStream_t side, hipEvent_t fork, hipEvent_t join);
// Enqueue the per-item form (one launch sequence per item) on one stream.
hipError_t phase_enqueue_plain(const PhasePlan& ph, hipStream_t s);

}  // namespace pl
