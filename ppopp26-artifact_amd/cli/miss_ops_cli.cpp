// miss_ops command line (the reference's miss_ops.cu): grand product,
// evaluation and division by (X + c), CPU baselines beside the GPU operators.
//   miss_ops                      reference workload: N = 2^22, ../../input.bin
//   miss_ops --lg <n> [--input <path>] [--perm] [--dump <dir>]
// --perm adds the PLONK permutation grand product (ratio over 3 witness and
// 3 sigma columns, then the exclusive product), seeded inputs.
// --dump writes the input and every operator's output of the last run to
// <dir> as raw 32-byte elements (input.bin, naive_gp.bin, ours_gp.bin,
// ours_eval.bin, naive_div.bin, ours_div.bin; with --perm also perm_cols.bin
// (a, b, c, s1, s2, s3), perm_bg.bin (beta, gamma) and perm_z.bin).
// The CPU baselines restate the reference's serial loops (miss_ops.cu:34-87)
// over this library's host field arithmetic.  Every run reloads its input
// through mutable_cpu_data(); the reference keeps one host pointer across
// runs (miss_ops.cu:95-102), so its second and third GPU runs re-read the
// previous run's output (same timing, different data).
#include <functional>
#include <vector>

#include "../csrc/fr_host.hpp"
#include "common.hpp"

using pl::host::fr64;

static fr64 ld(const void* p, uint64_t i) {
  fr64 v;
  std::memcpy(v.l, (const char*)p + 32 * i, 32);
  return v;
}
static void st(void* p, uint64_t i, const fr64& v) { std::memcpy((char*)p + 32 * i, v.l, 32); }

// one() = Mont(1) (miss_ops.cu:26-32)
static SyncedMemory one() {
  SyncedMemory o(32);
  st(o.mutable_cpu_data(), 0, pl::host::one());
  return o;
}

// naive_gp (miss_ops.cu:34-47): out[i] = prod_{j<=i} in[j] for i <= N-2, out[N-1] = 0
static SyncedMemory naive_gp(int lg_N, SyncedMemory input) {
  const uint64_t N = 1ull << lg_N;
  SyncedMemory output(input.size());
  const void* in = input.mutable_cpu_data();
  void* out = output.mutable_cpu_data();
  fr64 state = pl::host::one();
  st(out, 0, state);
  for (uint64_t i = 0; i + 1 < N; i++) {
    state = pl::host::mul(state, ld(in, i));
    st(out, i, state);
  }
  return output;
}

// naive_div (miss_ops.cu:49-87): long division by [c, d1], N-1 quotient coefficients
static SyncedMemory naive_div(int lg_N, SyncedMemory input, SyncedMemory divisor) {
  const uint64_t N = 1ull << lg_N;
  SyncedMemory quotient((N - 1) * 32);
  void* q = quotient.mutable_cpu_data();
  std::vector<fr64> rem(N);
  std::memcpy(rem.data(), input.mutable_cpu_data(), N * 32);
  const void* dv = divisor.mutable_cpu_data();
  const fr64 c = ld(dv, 0), d1 = ld(dv, 1);
  const fr64 lead_inv = pl::host::inv(d1);
  uint64_t len = N;
  while (len >= 2) {
    fr64 cur = pl::host::mul(rem[len - 1], lead_inv);
    st(q, len - 2, cur);
    rem[len - 2] = pl::host::sub(rem[len - 2], pl::host::mul(cur, c));
    rem[len - 1] = pl::host::sub(rem[len - 1], pl::host::mul(cur, d1));
    while (len > 0 && pl::host::is_zero(rem[len - 1])) len--;
  }
  return quotient;
}

static void dump(const std::string& dir, const char* name, const void* p, size_t bytes) {
  if (dir.empty()) return;
  const std::string path = dir + "/" + name;
  std::ofstream f(path, std::ios::binary);
  f.write((const char*)p, (std::streamsize)bytes);
  if (!f) throw std::runtime_error("cannot write " + path);
}

static float gpu_time(const std::function<void()>& fn) {
  cli::EventTimer t;
  t.start();
  fn();
  return t.stop_ms();
}

int main(int argc, char** argv) {
  try {
    int lg_N = 22;
    std::string file_path = "../../input.bin";
    bool perm = false;
    std::string dump_dir;
    for (int i = 1; i < argc; i++) {
      std::string a(argv[i]);
      if (a == "--lg" && i + 1 < argc) lg_N = std::atoi(argv[++i]);
      else if (a == "--input" && i + 1 < argc) file_path = argv[++i];
      else if (a == "--perm") perm = true;
      else if (a == "--dump" && i + 1 < argc) dump_dir = argv[++i];
      else {
        std::cerr << "Usage: " << argv[0] << " [--lg <n>] [--input <path>] [--perm] [--dump <dir>]\n";
        return 1;
      }
    }
    if (lg_N < 1 || lg_N > 28) throw std::runtime_error("lg_N must be in [1, 28]");
    const uint64_t N = 1ull << lg_N;
    SyncedMemory input(N * 32);
    void* inp_ = input.mutable_cpu_data();
    const uint64_t seed = cli::seed_of(4, lg_N, 0);

    std::cout << "test grand product" << std::endl;
    for (int rep = 0; rep < 3; rep++) {
      inp_ = input.mutable_cpu_data();  // head back at the CPU: the GPU run reads this rep's input
      cli::read_file(file_path.c_str(), inp_, N * 32, seed);
      std::cout << "[naive] run " << (rep + 1) << "/3" << std::endl;
      auto t0 = std::chrono::high_resolution_clock::now();
      SyncedMemory out = naive_gp(lg_N, input);
      auto t1 = std::chrono::high_resolution_clock::now();
      std::cout << "[naive_gp] execution time: "
                << std::chrono::duration_cast<std::chrono::milliseconds>(t1 - t0).count() << " ms" << std::endl;
      if (rep == 2) {
        dump(dump_dir, "input.bin", input.cpu_data(), N * 32);
        dump(dump_dir, "naive_gp.bin", out.cpu_data(), N * 32);
      }
    }
    for (int rep = 0; rep < 3; rep++) {
      inp_ = input.mutable_cpu_data();  // head back at the CPU: the GPU run reads this rep's input
      cli::read_file(file_path.c_str(), inp_, N * 32, seed);
      input.mutable_gpu_data();
      std::cout << "[ours] run " << (rep + 1) << "/3" << std::endl;
      float ms = gpu_time([&]() { accumulate_mul_poly(input); });
      std::cout << "[ours_gp] execution time: " << ms << " ms" << std::endl;
      if (rep == 2) dump(dump_dir, "ours_gp.bin", input.cpu_data(), N * 32);
    }

    if (perm) {
      // the permutation workload: ratio over a,b,c / s1,s2,s3 with k = (1, 7, 13), then z
      std::vector<void*> cols(6);
      for (int j = 0; j < 6; j++) {
        std::vector<uint64_t> h(4 * N);
        cli::fill_seeded(h.data(), N, cli::seed_of(4, lg_N, 10 + j));
        cli::hipc(hipMalloc(&cols[j], N * 32), "hipMalloc");
        cli::hipc(hipMemcpy(cols[j], h.data(), N * 32, hipMemcpyHostToDevice), "H2D");
      }
      uint64_t bg[8];
      cli::fill_seeded(bg, 2, cli::seed_of(4, lg_N, 20));
      void *d_beta, *d_gamma, *d_z, *d_work;
      cli::hipc(hipMalloc(&d_beta, 32), "hipMalloc");
      cli::hipc(hipMalloc(&d_gamma, 32), "hipMalloc");
      cli::hipc(hipMalloc(&d_z, N * 32), "hipMalloc");
      cli::hipc(hipMalloc(&d_work, pl_vec_work_bytes(N)), "hipMalloc");
      cli::hipc(hipMemcpy(d_beta, bg, 32, hipMemcpyHostToDevice), "H2D");
      cli::hipc(hipMemcpy(d_gamma, bg + 4, 32, hipMemcpyHostToDevice), "H2D");
      fr64 k[3] = {pl::host::from_u64(1), pl::host::from_u64(7), pl::host::from_u64(13)};
      pl_domain* dom = nullptr;
      cli::check(pl_domain_create(lg_N, &dom), "pl_domain_create");
      const void* w[3] = {cols[0], cols[1], cols[2]};
      const void* s[3] = {cols[3], cols[4], cols[5]};
      std::cout << "test permutation grand product" << std::endl;
      for (int rep = 0; rep < 3; rep++) {
        std::cout << "[ours] run " << (rep + 1) << "/3" << std::endl;
        float ms = gpu_time([&]() {
          cli::check(pl_perm_grand_product(w, s, d_beta, d_gamma, k, lg_N, dom, d_z, d_work, nullptr),
                     "pl_perm_grand_product");
        });
        std::cout << "[ours_perm_gp] execution time: " << ms << " ms" << std::endl;
      }
      if (!dump_dir.empty()) {
        std::vector<uint64_t> h(4 * N * 6);
        for (int j = 0; j < 6; j++)
          cli::hipc(hipMemcpy(h.data() + 4 * N * j, cols[j], N * 32, hipMemcpyDeviceToHost), "D2H");
        dump(dump_dir, "perm_cols.bin", h.data(), 6 * N * 32);
        dump(dump_dir, "perm_bg.bin", bg, 64);
        cli::hipc(hipMemcpy(h.data(), d_z, N * 32, hipMemcpyDeviceToHost), "D2H");
        dump(dump_dir, "perm_z.bin", h.data(), N * 32);
      }
      pl_domain_destroy(dom);
      for (void* p : cols) (void)hipFree(p);
      (void)hipFree(d_beta);
      (void)hipFree(d_gamma);
      (void)hipFree(d_z);
      (void)hipFree(d_work);
    }

    // evaluation at the reference's point (raw Montgomery limbs, miss_ops.cu:133-135)
    SyncedMemory point(32);
    const uint64_t pointt[4] = {123456789UL, 987654321UL, 1122334455UL, 6677889900UL};
    std::memcpy(point.mutable_cpu_data(), pointt, 32);
    std::cout << "test poly evaluation" << std::endl;
    for (int rep = 0; rep < 3; rep++) {
      inp_ = input.mutable_cpu_data();  // head back at the CPU: the GPU run reads this rep's input
      cli::read_file(file_path.c_str(), inp_, N * 32, seed);
      input.mutable_gpu_data();
      point.mutable_gpu_data();
      std::cout << "[ours] run " << (rep + 1) << "/3" << std::endl;
      SyncedMemory out(32);
      float ms = gpu_time([&]() { out = evaluate(input, point); });
      std::cout << "[ours_eval] execution time: " << ms << " ms" << std::endl;
      if (rep == 2) dump(dump_dir, "ours_eval.bin", out.cpu_data(), 32);
    }

    // division by X + c (miss_ops.cu:161-200)
    std::cout << "test poly division" << std::endl;
    SyncedMemory divisor(64);
    const uint64_t divisorr[8] = {123456789123456789UL, 987654321987654321UL, 1122334455667788UL, 1122334455667788UL,
                                  8589934590UL, 6378425256633387010UL, 11064306276430008309UL, 1739710354780652911UL};
    std::memcpy(divisor.mutable_cpu_data(), divisorr, 64);
    for (int rep = 0; rep < 3; rep++) {
      inp_ = input.mutable_cpu_data();  // head back at the CPU: the GPU run reads this rep's input
      cli::read_file(file_path.c_str(), inp_, N * 32, seed);
      std::cout << "[naive] run " << (rep + 1) << "/3" << std::endl;
      auto t0 = std::chrono::high_resolution_clock::now();
      SyncedMemory out = naive_div(lg_N, input, divisor);
      auto t1 = std::chrono::high_resolution_clock::now();
      std::cout << "[naive_div] execution time: "
                << std::chrono::duration_cast<std::chrono::milliseconds>(t1 - t0).count() << " ms" << std::endl;
      if (rep == 2) dump(dump_dir, "naive_div.bin", out.cpu_data(), (N - 1) * 32);
    }
    SyncedMemory divisor_gpu(32);
    std::memcpy(divisor_gpu.mutable_cpu_data(), divisorr, 32);
    divisor_gpu.mutable_gpu_data();
    for (int rep = 0; rep < 3; rep++) {
      inp_ = input.mutable_cpu_data();  // head back at the CPU: the GPU run reads this rep's input
      cli::read_file(file_path.c_str(), inp_, N * 32, seed);
      input.mutable_gpu_data();
      std::cout << "[ours] run " << (rep + 1) << "/3" << std::endl;
      SyncedMemory out(32);
      float ms = gpu_time([&]() { out = poly_div_poly(input, divisor_gpu); });
      std::cout << "[ours_div] execution time: " << ms << " ms" << std::endl;
      if (rep == 2) dump(dump_dir, "ours_div.bin", out.cpu_data(), N * 32);
    }
    return 0;
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 1;
  }
}
