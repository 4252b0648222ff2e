// Operator API on SyncedMemory over libpl's C-ABI (PLONK/utils/function.cu).
#include "function.hpp"

#include <hip/hip_runtime.h>

#include <cstring>
#include <stdexcept>

#include "../../include/pl.h"
#include "../csrc/fr_host.hpp"

namespace {

constexpr size_t kFr = 32;

uint64_t elems(SyncedMemory& m) { return m.size() / kFr; }

int lg_of(uint64_t n) {
  int lg = 0;
  while ((1ull << lg) < n) lg++;
  if ((1ull << lg) != n) throw std::invalid_argument("NTT size must be a power of two");
  return lg;
}

void check(int st, const char* what) {
  if (st != PL_OK)
    throw pl::HipError(std::string(what) + ": " + pl_status_string(st) + " " + pl_last_hip_error());
}

pl::host::fr64 ld(const void* p, uint64_t i) {
  pl::host::fr64 v;
  std::memcpy(v.l, (const char*)p + i * kFr, kFr);
  return v;
}
void st(void* p, uint64_t i, const pl::host::fr64& v) { std::memcpy((char*)p + i * kFr, v.l, kFr); }

// element-wise on the CPU (head at CPU)
void cpu_binary(int op, SyncedMemory& a, SyncedMemory& b, SyncedMemory& out, bool scalar) {
  const uint64_t n = elems(a);
  const void* pa = a.cpu_data();
  const void* pb = b.cpu_data();
  void* po = out.mutable_cpu_data();
  for (uint64_t i = 0; i < n; i++) {
    pl::host::fr64 x = ld(pa, i), y = ld(pb, scalar ? 0 : i), r;
    switch (op) {
      case PL_VEC_ADD: r = pl::host::add(x, y); break;
      case PL_VEC_SUB: r = pl::host::sub(x, y); break;
      case PL_VEC_MUL: r = pl::host::mul(x, y); break;
      default: r = pl::host::is_zero(y) ? y : pl::host::mul(x, pl::host::inv(y)); break;
    }
    st(po, i, r);
  }
}

void binary(int op, SyncedMemory a, SyncedMemory b, SyncedMemory out, bool scalar, hipStream_t s) {
  if (a.head() == SyncedMemory::HEAD_AT_CPU) {
    cpu_binary(op, a, b, out, scalar);
    return;
  }
  const int code = scalar ? op + PL_VEC_ADD_SCALAR : op;
  void* pa = a.mutable_gpu_data_async(s);
  void* pb = b.mutable_gpu_data_async(s);
  void* po = out.mutable_gpu_data_async(s);
  check(pl_vec_op(code, pa, pb, po, elems(a), nullptr, (pl_stream_t)s), "pl_vec_op");
}

SyncedMemory unary(int code, SyncedMemory in, hipStream_t s) {
  SyncedMemory out(in.size());
  void* pi = in.mutable_gpu_data_async(s);
  void* po = out.mutable_gpu_data_async(s);
  check(pl_vec_op(code, pi, nullptr, po, elems(in), nullptr, (pl_stream_t)s), "pl_vec_op");
  return out;
}

}  // namespace

SyncedMemory to_mont(SyncedMemory input) { return unary(PL_VEC_TO_MONT, input, nullptr); }
SyncedMemory to_base(SyncedMemory input, hipStream_t s) { return unary(PL_VEC_FROM_MONT, input, s); }
SyncedMemory neg_mod(SyncedMemory input) { return unary(PL_VEC_NEG, input, nullptr); }
SyncedMemory inv_mod(SyncedMemory input) { return unary(PL_VEC_INV, input, nullptr); }

// exp = 0 -> one, exp = 1 -> copy; CPU when the head is at the CPU
// (function.cu:113-120, cpu/mont_arithmetic.cpp:66-81)
SyncedMemory exp_mod(SyncedMemory input, int exp) {
  if (exp < 0) throw std::invalid_argument("exp_mod: negative exponent");
  SyncedMemory out(input.size());
  if (input.head() == SyncedMemory::HEAD_AT_CPU) {
    const uint64_t n = elems(input);
    const void* pi = input.cpu_data();
    void* po = out.mutable_cpu_data();
    for (uint64_t i = 0; i < n; i++) {
      pl::host::fr64 x = ld(pi, i), r = pl::host::one();
      if (exp == 1) {
        r = x;
      } else {
        for (int b = 31; b >= 0; b--) {
          r = pl::host::mul(r, r);
          if ((exp >> b) & 1) r = pl::host::mul(r, x);
        }
      }
      st(po, i, r);
    }
    return out;
  }
  void* pi = input.mutable_gpu_data();
  void* po = out.mutable_gpu_data();
  check(pl_exp_mod(pi, po, elems(input), exp, nullptr), "pl_exp_mod");
  return out;
}

// out[i] = x^i (function.cu:154-160)
void gen_sequence_(uint64_t N, SyncedMemory x, SyncedMemory y, hipStream_t s) {
  void* px = x.mutable_gpu_data_async(s);
  void* py = y.mutable_gpu_data_async(s);
  check(pl_gen_sequence(N, px, py, (pl_stream_t)s), "pl_gen_sequence");
}
SyncedMemory gen_sequence(uint64_t N, SyncedMemory x, hipStream_t s) {
  SyncedMemory y(N * kFr);
  gen_sequence_(N, x, y, s);
  return y;
}

// out[i] = x (function.cu:162-168)
void repeat_to_poly_(SyncedMemory x, SyncedMemory y, uint64_t N, hipStream_t s) {
  void* px = x.mutable_gpu_data_async(s);
  void* py = y.mutable_gpu_data_async(s);
  check(pl_repeat_to_poly(px, py, N, (pl_stream_t)s), "pl_repeat_to_poly");
}
SyncedMemory repeat_to_poly(SyncedMemory x, uint64_t N, hipStream_t s) {
  SyncedMemory y(N * kFr);
  repeat_to_poly_(x, y, N, s);
  return y;
}

#define PL_BINARY_OPS(NAME, OP, SCALAR)                                                       \
  SyncedMemory NAME(SyncedMemory a, SyncedMemory b, hipStream_t s) {                          \
    SyncedMemory out(a.size());                                                               \
    binary(OP, a, b, out, SCALAR, s);                                                         \
    return out;                                                                               \
  }                                                                                           \
  void NAME##_(SyncedMemory a, SyncedMemory b, hipStream_t s) { binary(OP, a, b, a, SCALAR, s); }

PL_BINARY_OPS(add_mod, PL_VEC_ADD, false)
PL_BINARY_OPS(sub_mod, PL_VEC_SUB, false)
PL_BINARY_OPS(mul_mod, PL_VEC_MUL, false)
PL_BINARY_OPS(div_mod, PL_VEC_DIV, false)
PL_BINARY_OPS(add_mod_scalar, PL_VEC_ADD, true)
PL_BINARY_OPS(sub_mod_scalar, PL_VEC_SUB, true)
PL_BINARY_OPS(mul_mod_scalar, PL_VEC_MUL, true)
PL_BINARY_OPS(div_mod_scalar, PL_VEC_DIV, true)

SyncedMemory evaluate(SyncedMemory poly, SyncedMemory x) {
  SyncedMemory y(kFr);
  void* py = y.mutable_gpu_data();
  void* px = x.mutable_gpu_data();
  const uint64_t n = elems(poly);
  void* pp = n ? poly.mutable_gpu_data() : nullptr;
  check(pl_eval(pp, n, px, py, nullptr, nullptr), "pl_eval");
  return y;
}

SyncedMemory poly_div_poly(SyncedMemory divid, SyncedMemory c) {
  void* pd = divid.mutable_gpu_data();
  void* pc = c.mutable_gpu_data();
  check(pl_div_linear(pd, elems(divid), pc, nullptr, nullptr, nullptr), "pl_div_linear");
  return divid;
}

SyncedMemory accumulate_mul_poly(SyncedMemory product, hipStream_t s) {
  void* p = product.mutable_gpu_data_async(s);
  check(pl_grand_product(p, elems(product), nullptr, nullptr, (pl_stream_t)s), "pl_grand_product");
  return product;
}

void accumulate_mul_poly_(SyncedMemory product, SyncedMemory output, hipStream_t s) {
  void* p = product.mutable_gpu_data_async(s);
  void* o = output.mutable_gpu_data_async(s);
  check(pl_grand_product(p, elems(product), o, nullptr, (pl_stream_t)s), "pl_grand_product");
}

SyncedMemory pad_poly(SyncedMemory x, uint64_t N, hipStream_t s) {
  SyncedMemory out(N * kFr);
  const size_t bytes = std::min<size_t>(x.size(), N * kFr);
  if (x.head() == SyncedMemory::HEAD_AT_CPU) {
    std::memcpy(out.mutable_cpu_data(), x.cpu_data(), bytes);
  } else {
    void* po = out.mutable_gpu_data_async(s);  // zero-filled on first touch
    pl::hip_check(hipMemcpyAsync(po, x.mutable_gpu_data_async(s), bytes, hipMemcpyDeviceToDevice, s), "pad_poly");
  }
  return out;
}

SyncedMemory repeat_zero(uint64_t N) {
  SyncedMemory out(N * kFr);
  std::memset(out.mutable_cpu_data(), 0, out.size());
  return out;
}

SyncedMemory cat(SyncedMemory a, SyncedMemory b) {
  SyncedMemory r(a.size() + b.size());
  if (a.head() == SyncedMemory::HEAD_AT_CPU) {
    char* p = (char*)r.mutable_cpu_data();
    std::memcpy(p, a.cpu_data(), a.size());
    std::memcpy(p + a.size(), b.cpu_data(), b.size());
  } else {
    char* p = (char*)r.mutable_gpu_data();
    pl::hip_check(hipMemcpy(p, a.mutable_gpu_data(), a.size(), hipMemcpyDeviceToDevice), "cat");
    pl::hip_check(hipMemcpy(p + a.size(), b.mutable_gpu_data(), b.size(), hipMemcpyDeviceToDevice), "cat");
  }
  return r;
}

SyncedMemory slice(SyncedMemory a, uint64_t len, bool forward) {
  SyncedMemory r(len * kFr);
  const size_t off = forward ? 0 : a.size() - r.size();
  if (a.head() == SyncedMemory::HEAD_AT_CPU) {
    std::memcpy(r.mutable_cpu_data(), (const char*)a.cpu_data() + off, r.size());
  } else {
    pl::hip_check(hipMemcpy(r.mutable_gpu_data(), (char*)a.mutable_gpu_data() + off, r.size(),
                            hipMemcpyDeviceToDevice), "slice");
  }
  return r;
}

namespace pl {
DomainCache::~DomainCache() {
  for (auto& kv : doms) pl_domain_destroy(kv.second);
}
pl_domain* DomainCache::get(int lg) {
  std::lock_guard<std::mutex> g(mu);
  auto it = doms.find(lg);
  if (it != doms.end()) return it->second;
  pl_domain* d = nullptr;
  check(pl_domain_create(lg, &d), "pl_domain_create");
  doms[lg] = d;
  return d;
}
}  // namespace pl

Ntt::Ntt(int domain_size, hipStream_t stream) : cache_(std::make_shared<pl::DomainCache>()) {
  (void)domain_size;  // domains are built per size on first use
  Params = SyncedMemory(4 * 16384 * kFr);
  check(pl_params_fill(Params.mutable_gpu_data_async(stream), (pl_stream_t)stream), "pl_params_fill");
}

SyncedMemory Ntt::forward(SyncedMemory input, hipStream_t s) {
  SyncedMemory out(input.size());
  forward_(input, out, s);
  return out;
}

void Ntt::forward_(SyncedMemory input, SyncedMemory output, hipStream_t s) {
  const uint64_t n = elems(input);
  const int lg = lg_of(n);
  void* pi = input.mutable_gpu_data_async(s);
  void* po = output.mutable_gpu_data_async(s);
  check(pl_ntt(pi, po, lg, inverse_ ? 1 : 0, domain(lg), nullptr, (pl_stream_t)s), "pl_ntt");
}

void Ntt::stage_no_rotate(SyncedMemory input, int lg_domain_size, int iterations, int stage, int chunk_id,
                          hipStream_t s) {
  if (inverse_) throw std::logic_error("stage_no_rotate: forward transform only (Ntt::stage_no_rotate)");
  void* pi = input.mutable_gpu_data_async(s);
  const void* pp = Params.mutable_gpu_data_async(s);
  check(pl_ntt_stage_no_rotate(pi, elems(input), lg_domain_size, iterations, stage, (uint64_t)chunk_id, pp,
                               (pl_stream_t)s),
        "pl_ntt_stage_no_rotate");
}

Intt::Intt(int domain_size, hipStream_t stream) : Ntt(domain_size, stream) { inverse_ = true; }

void Intt::_forward_(SyncedMemory input, hipStream_t s) { forward_(input, input, s); }

namespace pl {
// coset domains of one Ntt_coset object, by size; LDE domains by (size, blowup)
struct CosetDomains {
  std::mutex mu;
  std::map<int, pl_coset_domain*> coset;
  std::map<std::pair<int, int>, pl_lde_domain*> lde;
  ~CosetDomains() {
    for (auto& kv : coset) pl_coset_domain_destroy(kv.second);
    for (auto& kv : lde) pl_lde_domain_destroy(kv.second);
  }
  pl_coset_domain* get(int lg) {
    std::lock_guard<std::mutex> g(mu);
    auto it = coset.find(lg);
    if (it != coset.end()) return it->second;
    pl_coset_domain* d = nullptr;
    check(pl_coset_domain_create(lg, nullptr, &d), "pl_coset_domain_create");
    coset[lg] = d;
    return d;
  }
  pl_lde_domain* get_lde(int lg, int lb) {
    std::lock_guard<std::mutex> g(mu);
    auto it = lde.find({lg, lb});
    if (it != lde.end()) return it->second;
    const pl::host::fr64 seven = pl::host::from_u64(7);
    pl_lde_domain* d = nullptr;
    check(pl_lde_domain_create(lg, lb, seven.l, &d), "pl_lde_domain_create");
    lde[{lg, lb}] = d;
    return d;
  }
};
}  // namespace pl

Ntt_coset::Ntt_coset(int domain_size, int coset_size)
    : lg_domain_size(coset_size), doms_(std::make_shared<pl::CosetDomains>()) {
  (void)domain_size;
  Params = SyncedMemory(4 * 16384 * kFr);
  check(pl_params_fill(Params.mutable_gpu_data(), nullptr), "pl_params_fill");
}

Intt_coset::Intt_coset(int domain_size, int coset_size, hipStream_t stream) : Ntt_coset(domain_size, coset_size) {
  (void)stream;
  inverse_ = true;
}

void Ntt_coset::init(SyncedMemory input, int lg, int chunk_id, bool is_intt, hipStream_t s) {
  void* p = input.mutable_gpu_data_async(s);
  check(pl_coset_scale(p, elems(input), (uint64_t)chunk_id, is_intt ? 1 : 0, doms_->get(lg), (pl_stream_t)s),
        "pl_coset_scale");
}

void Intt_coset::init(SyncedMemory input, int lg, int chunk_id, bool is_intt, hipStream_t s) {
  Ntt_coset::init(input, lg, chunk_id, is_intt, s);
}

void Ntt_coset::init_and_forward(SyncedMemory input, SyncedMemory output, int lg, int chunk_id, int lambda, int stage,
                                 bool is_intt, hipStream_t s) {
  (void)chunk_id;
  (void)stage;
  if (is_intt) throw std::logic_error("init_and_forward: forward coset LDE only");
  void* pi = input.mutable_gpu_data_async(s);
  void* po = output.mutable_gpu_data_async(s);
  check(pl_lde(pi, po, doms_->get_lde(lg, lambda), nullptr, (pl_stream_t)s), "pl_lde");
}

void Ntt_coset::forward1(SyncedMemory input, int lg, int stage, int chunk_id, hipStream_t s) {
  (void)stage;
  (void)chunk_id;
  void* p = input.mutable_gpu_data_async(s);
  check(pl_coset_ntt_step(p, p, 1, inverse_ ? 1 : 0, doms_->get(lg), (pl_stream_t)s), "pl_coset_ntt_step");
}

void Ntt_coset::forward2(SyncedMemory input, int lg, int stage, int chunk_id, hipStream_t s) {
  (void)stage;
  (void)chunk_id;
  void* p = input.mutable_gpu_data_async(s);
  check(pl_coset_ntt_step(p, nullptr, 2, inverse_ ? 1 : 0, doms_->get(lg), (pl_stream_t)s), "pl_coset_ntt_step");
}

void Ntt_coset::forward3(SyncedMemory input, SyncedMemory output, int lg, int stage, int chunk_id, int lambda,
                         hipStream_t s) {
  (void)stage;
  (void)chunk_id;
  (void)lambda;
  void* pi = input.mutable_gpu_data_async(s);
  void* po = output.mutable_gpu_data_async(s);
  check(pl_coset_ntt_step(pi, po, 3, inverse_ ? 1 : 0, doms_->get(lg), (pl_stream_t)s), "pl_coset_ntt_step");
}

SyncedMemory Ntt_coset::forward(SyncedMemory input, hipStream_t s) {
  SyncedMemory out(input.size());
  const int lg = lg_of(elems(input));
  void* pi = input.mutable_gpu_data_async(s);
  void* po = out.mutable_gpu_data_async(s);
  check(pl_coset_ntt(pi, po, inverse_ ? 1 : 0, doms_->get(lg), nullptr, (pl_stream_t)s), "pl_coset_ntt");
  return out;
}

uint32_t bit_rev(uint32_t i, uint32_t nbits) {
  uint32_t r = 0;
  for (uint32_t b = 0; b < nbits; b++) r |= ((i >> b) & 1u) << (nbits - 1 - b);
  return r;
}

void calculate_map1(uint32_t* map, int lg_N) {
  const uint64_t N = uint64_t{1} << lg_N;
  for (uint64_t i = 0; i < N; i++) map[bit_rev((uint32_t)i, lg_N)] = (uint32_t)i;
}
