// LDE command line (the reference's lde.cu driver, lde.cu:1492-1854):
//   LDE                              suite: lg_N in {20,22,24} x lambda in {2,3}
//   LDE --bench <lg_N> [lambda]      our_lde
//   LDE --bench-naive <lg_N> [lambda]
//   LDE --bench-sweep <lg_N> <lambda> [reps]
// lambda is log2 of the blowup; our_lde evaluates the N coefficients on the
// coset 7^(2^lambda) * <w_M>, M = N * 2^lambda (lde.cu:642-645).  The input
// is a host buffer; the timed region covers H2D, the LDE and the D2H of the
// M outputs, as the reference's does (its result is discarded too).
#include <vector>

#include "common.hpp"

struct LdeScaleConfig {
  int lg_N, lambda, ours_lg_k, reps;
  const char* input_file;
};

static LdeScaleConfig lde_scale_config(int lg_N, int lambda) {
  if (lambda == 2 || lambda == 3) {
    switch (lg_N) {
      case 20: return {20, lambda, 4, 3, "../../input.bin"};
      case 22: return {22, lambda, lambda == 2 ? 3 : 2, 3, "../../input.bin"};
      case 24: return {24, lambda, 1, 3, "../../input-24.bin"};
      default: break;
    }
  }
  throw std::runtime_error("unsupported LDE test config; supported lg_N={20,22,24}, lambda={2,3}");
}

static bool env_enabled(const char* name) {
  const char* v = std::getenv(name);
  return v && std::atoi(v) != 0;
}

struct LdeRun {
  int lg_N, lambda;
  pl_lde_domain* dom = nullptr;
  pl_domain* ndom = nullptr;
  void* d_in = nullptr;
  void* d_out = nullptr;
  void* d_work = nullptr;
  void* h_out = nullptr;
  LdeRun(int lg, int lb, bool naive) : lg_N(lg), lambda(lb) {
    const size_t n_bytes = (size_t)32 << lg, m_bytes = n_bytes << lb;
    if (naive)
      cli::check(pl_domain_create(lg, &ndom), "pl_domain_create");
    else
      cli::check(pl_lde_domain_create(lg, lb, nullptr, &dom), "pl_lde_domain_create");
    cli::hipc(hipMalloc(&d_in, n_bytes), "hipMalloc");
    cli::hipc(hipMalloc(&d_out, m_bytes), "hipMalloc");
    if (!naive) cli::hipc(hipMalloc(&d_work, pl_lde_work_bytes(lg, lb)), "hipMalloc");
    cli::hipc(hipHostMalloc(&h_out, m_bytes, hipHostMallocDefault), "hipHostMalloc");
  }
  ~LdeRun() {
    if (dom) pl_lde_domain_destroy(dom);
    if (ndom) pl_domain_destroy(ndom);
    (void)hipFree(d_in);
    (void)hipFree(d_out);
    if (d_work) (void)hipFree(d_work);
    (void)hipHostFree(h_out);
  }
  // one timed LDE: H2D, transform, D2H
  float run(const void* h_in, int lg_k, bool naive, bool profile) {
    const size_t n_bytes = (size_t)32 << lg_N, m_bytes = n_bytes << lambda;
    cli::EventTimer all, h2d, dev, d2h;
    cli::hipc(hipDeviceSynchronize(), "sync");
    all.start();
    h2d.start();
    cli::copy_chunked(d_in, h_in, n_bytes, 1 << lg_k, hipMemcpyHostToDevice, nullptr);
    float h2d_ms = h2d.stop_ms();
    dev.start();
    if (naive)
      cli::check(pl_lde_naive(d_in, d_out, lg_N, lambda, ndom, nullptr), "pl_lde_naive");
    else
      cli::check(pl_lde(d_in, d_out, dom, d_work, nullptr), "pl_lde");
    float dev_ms = dev.stop_ms();
    d2h.start();
    cli::copy_chunked(h_out, d_out, m_bytes, 1 << lg_k, hipMemcpyDeviceToHost, nullptr);
    float d2h_ms = d2h.stop_ms();
    float ms = all.stop_ms();
    if (naive) {
      std::cout << "naive-lde execution time: " << ms << " ms" << std::endl;
      if (profile)
        std::cout << "[naive-lde-profile] h2d=" << h2d_ms << " ms, device=" << dev_ms << " ms, d2h=" << d2h_ms
                  << " ms" << std::endl;
    } else {
      std::cout << "our-lde execution time: " << ms << " ms" << std::endl;
      if (profile)
        std::cout << "[lde-profile] h2d=" << h2d_ms << " ms, device=" << dev_ms << " ms, d2h=" << d2h_ms << " ms"
                  << std::endl;
    }
    return ms;
  }
};

static void print_workspace(int lg_N, int lambda) {
  const double mib = 1024.0 * 1024.0;
  const double in_mb = ((size_t)32 << lg_N) / mib, out_mb = ((size_t)32 << (lg_N + lambda)) / mib;
  std::cout << std::fixed << std::setprecision(2) << "[workspace] device-resident input=" << in_mb
            << " MiB, output=" << out_mb << " MiB, coset NTTs=" << (1 << lambda) << " x 2^" << lg_N
            << std::defaultfloat << std::setprecision(6) << std::endl;
}

static int run_benchmark(const LdeScaleConfig& cfg, int k_begin, int k_end, int reps, bool naive) {
  const uint64_t N = 1ull << cfg.lg_N;
  std::vector<uint64_t> input(4 * N);
  cli::read_file(cfg.input_file, input.data(), N * 32, cli::seed_of(3, cfg.lg_N, cfg.lambda));
  const bool profile = env_enabled("LDE_PROFILE");
  const char* tag = naive ? "naive_lde" : "our_lde";
  std::cout << (naive ? "[naive-lde] lg_N=" : "[lde] lg_N=") << cfg.lg_N << ", lambda=" << cfg.lambda
            << ", input=" << cfg.input_file << ", reps=" << reps << std::endl;
  LdeRun run(cfg.lg_N, cfg.lambda, naive);
  for (int lg_k = k_begin; lg_k <= k_end; lg_k++) {
    std::cout << "=== " << (naive ? "naive lde" : "lde") << " lg_N = " << cfg.lg_N << ", lambda = " << cfg.lambda
              << ", copy_chunks = " << (1 << lg_k) << " ===" << std::endl;
    print_workspace(cfg.lg_N, cfg.lambda);
    double sum = 0;
    for (int rep = 0; rep < reps; rep++) {
      std::cout << "[" << tag << "] run " << (rep + 1) << "/" << reps << std::endl;
      sum += run.run(input.data(), lg_k, naive, profile);
    }
    std::cout << "[" << tag << "] avg lg_N=" << cfg.lg_N << ", lambda=" << cfg.lambda
              << ", copy_chunks=" << (1 << lg_k) << ": " << (sum / reps) << " ms" << std::endl;
  }
  return 0;
}

int main(int argc, char** argv) {
  try {
    if (argc == 1) {
      for (int lg : {20, 22, 24})
        for (int lb : {2, 3}) {
          LdeScaleConfig c = lde_scale_config(lg, lb);
          int rc = run_benchmark(c, c.ours_lg_k, c.ours_lg_k, c.reps, false);
          if (rc) return rc;
        }
      return 0;
    }
    const std::string cmd(argv[1]);
    if (cmd == "--bench" || cmd == "--bench-naive") {
      if (argc < 3 || argc > 4) {
        std::cerr << "Usage: " << argv[0] << " " << cmd << " <lg_N> [lambda]\n";
        return 1;
      }
      const bool naive = cmd == "--bench-naive";
      const int lg_N = std::atoi(argv[2]);
      std::vector<int> lambdas = argc == 4 ? std::vector<int>{std::atoi(argv[3])} : std::vector<int>{2, 3};
      for (int lb : lambdas) {
        LdeScaleConfig c = lde_scale_config(lg_N, lb);
        int rc = naive ? run_benchmark(c, 1, 8, c.reps, true) : run_benchmark(c, c.ours_lg_k, c.ours_lg_k, c.reps, false);
        if (rc) return rc;
      }
      return 0;
    }
    if (cmd == "--bench-sweep") {
      if (argc < 4 || argc > 5) {
        std::cerr << "Usage: " << argv[0] << " --bench-sweep <lg_N> <lambda> [reps]\n";
        return 1;
      }
      LdeScaleConfig c = lde_scale_config(std::atoi(argv[2]), std::atoi(argv[3]));
      return run_benchmark(c, 1, 8, argc == 5 ? std::atoi(argv[4]) : 1, false);
    }
    std::cerr << "Unknown command: " << cmd << "\n";
    return 1;
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 1;
  }
}
