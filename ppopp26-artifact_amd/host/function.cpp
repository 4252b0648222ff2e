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

uint32_t bit_rev(uint32_t i, uint32_t nbits) {
  uint32_t r = 0;
  for (uint32_t b = 0; b < nbits; b++) r |= ((i >> b) & 1u) << (nbits - 1 - b);
  return r;
}

void calculate_map1(uint32_t* map, int lg_N) {
  const uint64_t N = uint64_t{1} << lg_N;
  for (uint64_t i = 0; i < N; i++) map[bit_rev((uint32_t)i, lg_N)] = (uint32_t)i;
}
