// NTT command line (the reference's ntt.cu driver, ntt.cu:1182-1505):
//   NTT --bench-split <lg_N> | --bench-baseline <lg_N> | --check-split <lg_N>
//       | --check-baseline <lg_N> | --bench-sweep <lg_N> [reps]
// lg_N in {22, 24, 26}.  The input lives in pinned host memory as in the
// reference.  `our_ntt` is the host-to-host pipelined transform
// (pl_ntt_host: 2^lg_k copy groups overlapping the first and last passes);
// --check-split compares it with the reference's own chunked flow run on the
// device (pl_ntt_segmented: bit reversal, stage_no_rotate steps, transition,
// gathers and scatters), an independent set of kernels.
#include <vector>

#include "common.hpp"

struct NttScaleConfig {
  int lg_N, lg_i, ours_lg_k, baseline_lg_k_begin, baseline_lg_k_end, reps;
  const char* input_file;
};

static NttScaleConfig ntt_scale_config(int lg_N) {
  switch (lg_N) {
    case 22: return {22, 12, 4, 1, 8, 3, "../../input.bin"};
    case 24: return {24, 14, 4, 1, 8, 3, "../../input-24.bin"};
    case 26: return {26, 16, 2, 1, 8, 3, "../../input-26.bin"};
    default:
      std::cerr << "Unsupported lg_N = " << lg_N << ". Supported NTT test scales are 22, 24, and 26." << std::endl;
      std::exit(1);
  }
}

static void print_workspace() {
  std::cout << "[workspace] mode=host-to-host pipeline, two device slots, grouped 2-D copies" << std::endl;
}

// our_ntt (ntt.cu:979-1037): natural order in and out, in place in the host
// buffer of `input`; the printed time covers copies in, transform and copies
// out, the scope of the reference's timer.  lg_i / lg_j only shape the
// printed split; lg_k sets the copy groups.
static float our_ntt(SyncedMemory input, int lg_N, int lg_i, int lg_j, int lg_k, Ntt& ntt, void* d_work,
                     hipStream_t s) {
  (void)lg_i;
  (void)lg_j;
  void* h = input.mutable_cpu_data();
  cli::hipc(hipStreamSynchronize(s), "sync");
  const auto t0 = std::chrono::steady_clock::now();
  cli::check(pl_ntt_host(h, h, lg_N, 1, 0, ntt.domain(lg_N), lg_k, d_work, (pl_stream_t)s), "pl_ntt_host");
  const float ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
  std::cout << "our-ntt execution time: " << ms << " ms" << std::endl;
  return ms;
}

// the reference's chunked flow on the device (pl_ntt_segmented), host to
// host: the expected output of --check-split
static void segmented_reference(SyncedMemory input, int lg_N, int lg_i, int lg_k, hipStream_t s) {
  const size_t bytes = (size_t)32 << lg_N;
  void* h = input.mutable_cpu_data();
  void *d = nullptr, *params = nullptr;
  cli::hipc(hipMalloc(&d, bytes), "hipMalloc");
  cli::hipc(hipMalloc(&params, 4 * 16384 * 32), "hipMalloc");
  cli::check(pl_params_fill(params, (pl_stream_t)s), "pl_params_fill");
  cli::hipc(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s), "hipMemcpyAsync");
  cli::check(pl_ntt_segmented(d, lg_N, lg_i, lg_k, params, nullptr, (pl_stream_t)s), "pl_ntt_segmented");
  cli::hipc(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
  cli::hipc(hipStreamSynchronize(s), "sync");
  (void)hipFree(d);
  (void)hipFree(params);
  std::cout << "[check] expected: the reference's chunked flow on the device, split " << lg_i << "+"
            << (lg_N - lg_i) << ", 2^" << lg_k << " segments" << std::endl;
}

// device-resident NTT of the same size (HBM in, HBM out), for comparison
static float device_resident_ms(int lg_N, Ntt& ntt, void* d_work, hipStream_t s) {
  const size_t bytes = (size_t)32 << lg_N;
  void* d = nullptr;
  cli::hipc(hipMalloc(&d, bytes), "hipMalloc");
  cli::check(pl_fill_uniform(d, 1ull << lg_N, cli::seed_of(1, lg_N, 0), nullptr, (pl_stream_t)s), "fill");
  cli::check(pl_ntt(d, d, lg_N, 0, ntt.domain(lg_N), d_work, (pl_stream_t)s), "pl_ntt");
  cli::EventTimer t(s);
  t.start();
  cli::check(pl_ntt(d, d, lg_N, 0, ntt.domain(lg_N), d_work, (pl_stream_t)s), "pl_ntt");
  const float ms = t.stop_ms();
  (void)hipFree(d);
  return ms;
}

// baseline (standard4_ntt's role, ntt.cu:1039-1138): bit reversal plus one
// radix-2 launch per stage, host-to-host
static float baseline_ntt(SyncedMemory input, int lg_N, int lg_k, Ntt& ntt, void* d_a, void* d_b, hipStream_t s) {
  const size_t bytes = (size_t)32 << lg_N;
  void* h = input.mutable_cpu_data();
  cli::EventTimer all(s);
  cli::hipc(hipDeviceSynchronize(), "sync");
  all.start();
  cli::copy_chunked(d_a, h, bytes, 1 << lg_k, hipMemcpyHostToDevice, s);
  cli::check(pl_ntt_naive(d_a, d_b, lg_N, ntt.domain(lg_N), (pl_stream_t)s), "pl_ntt_naive");
  cli::copy_chunked(h, d_b, bytes, 1 << lg_k, hipMemcpyDeviceToHost, s);
  float ms = all.stop_ms();
  std::cout << "standard4-ntt execution time: " << ms << " ms" << std::endl;
  return ms;
}

struct DevBufs {
  void* a = nullptr;
  void* b = nullptr;
  explicit DevBufs(size_t bytes) {
    cli::hipc(hipMalloc(&a, bytes), "hipMalloc");
    cli::hipc(hipMalloc(&b, bytes), "hipMalloc");
  }
  ~DevBufs() {
    (void)hipFree(a);
    (void)hipFree(b);
  }
};

static int run_check(const NttScaleConfig& cfg, bool baseline) {
  const uint64_t N = 1ull << cfg.lg_N;
  Ntt ntt(32);
  SyncedMemory reference_input(N * 32), ours_input(N * 32);
  cli::read_file(cfg.input_file, reference_input.mutable_cpu_data(), N * 32, cli::seed_of(1, cfg.lg_N, 0));
  std::memcpy(ours_input.mutable_cpu_data(), reference_input.cpu_data(), N * 32);
  SyncedMemory expected(N * 32);
  if (baseline) {
    // baseline check: the library's NTT against the naive baseline
    expected = ntt.forward(reference_input);
    cli::hipc(hipDeviceSynchronize(), "sync");
    expected.mutable_cpu_data();
    DevBufs buf(N * 32);
    std::cout << "[workspace] mode=naive, bit reversal + one radix-2 launch per stage" << std::endl;
    baseline_ntt(ours_input, cfg.lg_N, cfg.ours_lg_k, ntt, buf.a, buf.b, nullptr);
  } else {
    // split check: the reference's chunked flow (expected) against our_ntt
    std::memcpy(expected.mutable_cpu_data(), reference_input.cpu_data(), N * 32);
    segmented_reference(expected, cfg.lg_N, cfg.lg_i, cfg.ours_lg_k, nullptr);
    print_workspace();
    void* w = nullptr;
    cli::hipc(hipMalloc(&w, pl_ntt_host_work_bytes(cfg.lg_N, 1)), "hipMalloc");
    our_ntt(ours_input, cfg.lg_N, cfg.lg_i, cfg.lg_N - cfg.lg_i, cfg.ours_lg_k, ntt, w, nullptr);
    (void)hipFree(w);
  }
  cli::hipc(hipDeviceSynchronize(), "sync");
  return cli::compare_outputs(expected.cpu_data(), ours_input.cpu_data(), N) ? 0 : 1;
}

static int run_bench(const NttScaleConfig& cfg, int k_begin, int k_end, int reps, bool baseline) {
  const uint64_t N = 1ull << cfg.lg_N;
  const int lg_j = cfg.lg_N - cfg.lg_i;
  Ntt ntt(32);
  SyncedMemory input(N * 32);
  void* h = input.mutable_cpu_data();
  DevBufs buf(N * 32);
  void* w = nullptr;  // the pipeline's device slot (data + scratch)
  if (baseline) {
    std::cout << "[workspace] mode=naive, bit reversal + one radix-2 launch per stage" << std::endl;
  } else {
    print_workspace();
    cli::hipc(hipMalloc(&w, pl_ntt_host_work_bytes(cfg.lg_N, 1)), "hipMalloc");
  }
  for (int lg_k = k_begin; lg_k <= k_end; lg_k++) {
    const size_t chunk_bytes = ((size_t)1 << (cfg.lg_i + lg_k)) * 32;
    std::cout << "=== lg_N = " << cfg.lg_N << ", split = " << cfg.lg_i << "+" << lg_j
              << ", mode = " << (baseline ? "baseline-naive" : "host-to-host pipeline")
              << ", chunk_size = " << cli::format_bytes(chunk_bytes) << " ===" << std::endl;
    for (int rep = 0; rep < reps; rep++) {
      cli::read_file(cfg.input_file, h, N * 32, cli::seed_of(1, cfg.lg_N, 0));
      if (baseline) {
        std::cout << "[standard4_ntt] run " << (rep + 1) << "/" << reps << std::endl;
        baseline_ntt(input, cfg.lg_N, lg_k, ntt, buf.a, buf.b, nullptr);
      } else {
        std::cout << "[our_ntt] run " << (rep + 1) << "/" << reps << std::endl;
        our_ntt(input, cfg.lg_N, cfg.lg_i, lg_j, lg_k, ntt, w, nullptr);
      }
    }
  }
  if (!baseline) {
    (void)hipFree(w);
    std::cout << "[our_ntt] device-resident NTT (HBM in, HBM out): " << device_resident_ms(cfg.lg_N, ntt, buf.b, nullptr)
              << " ms" << std::endl;
  }
  return 0;
}

static void usage(const char* exe) {
  std::cerr << "Usage:\n"
            << "  " << exe << " --bench-split <lg_N>\n"
            << "  " << exe << " --bench-baseline <lg_N>\n"
            << "  " << exe << " --check-split <lg_N>\n"
            << "  " << exe << " --check-baseline <lg_N>\n"
            << "  " << exe << " --bench-sweep <lg_N> [reps]\n"
            << "  " << exe << " --autotune <lg_lo> <lg_hi> [cache_file [host]]\n";
}

int main(int argc, char** argv) {
  if (argc < 2) {
    usage(argv[0]);
    return 1;
  }
  try {
    const std::string cmd(argv[1]);
    const int lg_N = argc >= 3 ? std::atoi(argv[2]) : 22;
    if (cmd == "--bench-baseline") {
      NttScaleConfig c = ntt_scale_config(lg_N);
      return run_bench(c, c.baseline_lg_k_begin, c.baseline_lg_k_end, c.reps, true);
    }
    if (cmd == "--check-baseline") return run_check(ntt_scale_config(lg_N), true);
    if (cmd == "--check-split") return run_check(ntt_scale_config(lg_N), false);
    if (cmd == "--bench-split") {
      NttScaleConfig c = ntt_scale_config(lg_N);
      return run_bench(c, c.ours_lg_k, c.ours_lg_k, c.reps, false);
    }
    if (cmd == "--autotune") {
      // pass-split autotuner, and with "host" the copy groups (replaces the
      // manual segment sweep)
      if (argc < 4) {
        usage(argv[0]);
        return 1;
      }
      const int lo = std::atoi(argv[2]), hi = std::atoi(argv[3]);
      const char* cache = argc >= 5 ? argv[4] : nullptr;
      // "host": also the copy groups of our_ntt's host-to-host scope
      const bool host = argc >= 6 && std::string(argv[5]) == "host";
      cli::check(pl_autotune(lo, hi, cache, 5, host ? PL_TUNE_HOST : 0), "pl_autotune");
      for (int lg = lo; lg <= hi; lg++) {
        int m[8], P = 0;
        cli::check(pl_plan_get(lg, m, &P), "pl_plan_get");
        std::cout << "[autotune] lg_N = " << lg << ": " << P << " passes of";
        for (int i = 0; i < P; i++) std::cout << " 2^" << m[i];
        if (host && lg >= 12 && lg <= 26) std::cout << "; host-to-host: 2^" << pl_ntt_host_groups(lg) << " copy groups";
        std::cout << std::endl;
      }
      return 0;
    }
    if (cmd == "--bench-sweep") {
      NttScaleConfig c = ntt_scale_config(lg_N);
      const int reps = argc >= 4 ? std::atoi(argv[3]) : 1;
      return run_bench(c, 1, 8, reps, false);
    }
    std::cerr << "Unknown command: " << cmd << "\n";
    usage(argv[0]);
    return 1;
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 1;
  }
}
