// Exercises the C++ operator API (SyncedMemory + function.hpp) on seeded
// inputs and dumps raw results, so tests can check them against the oracle:
//   api_check <out_dir> <lg>
//   api_check --host-helpers <out_dir> <lg>   (no GPU call)
// Files (raw 32-byte elements): x.bin y.bin (inputs), ntt.bin intt.bin,
// gp.bin gp_incl.bin, eval.bin, div.bin, add.bin mul.bin div_ew.bin,
// add_cpu.bin (CPU-dispatched add), lde.bin (our_lde semantics, lambda=2),
// stage0.bin stage6.bin (Ntt::stage_no_rotate on seeded chunks), coset.bin
// icoset.bin coset_steps.bin icoset_steps.bin coset_init0.bin coset_lde3.bin
// (Ntt_coset / Intt_coset), exp5.bin exp5_cpu.bin (exp_mod(x, 5) on the GPU
// and the CPU), seq.bin (gen_sequence(n + 5, pt)), rep.bin (repeat_to_poly(pt, 33)).
#include <string>

#include "common.hpp"

static void dump(const std::string& path, SyncedMemory m) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot write " + path);
  std::fwrite(m.cpu_data(), 1, m.size(), f);
  std::fclose(f);
}

static SyncedMemory seeded(uint64_t n, uint64_t seed) {
  SyncedMemory m(n * 32);
  cli::fill_seeded(m.mutable_cpu_data(), n, seed);
  return m;
}

// calculate_map1 / bit_rev_step1_parallel (function.cuh:157-244) on the host
// only: in.bin (2^lg seeded elements), map.bin (the 2^lg map, u32), brev_t<T>.bin for T = 1, 3, 8 threads,
// and brev_idx.bin (the 2^lg inputs through the upper half of the 2^(lg+1)
// map, idx = 2^lg, into 2^(lg+1) zeroed outputs)
struct el32 {
  uint64_t w[4];
};
static int host_helpers(const std::string& dir, int lg) {
  const uint64_t n = 1ull << lg;
  std::vector<el32> in(n), out(n), out2(2 * n);
  cli::fill_seeded(in.data(), n, 1001);
  std::vector<uint32_t> map(n), map2(2 * n);
  calculate_map1(map.data(), lg);
  calculate_map1(map2.data(), lg + 1);
  auto write = [&](const std::string& name, const void* p, size_t bytes) {
    FILE* f = std::fopen((dir + "/" + name).c_str(), "wb");
    if (!f) throw std::runtime_error("cannot write " + name);
    std::fwrite(p, 1, bytes, f);
    std::fclose(f);
  };
  write("in.bin", in.data(), n * 32);
  write("map.bin", map.data(), n * 4);
  for (int t : {1, 3, 8}) {
    std::fill(out.begin(), out.end(), el32{});
    bit_rev_step1_parallel(out.data(), in.data(), map.data(), lg, t, 0);
    write("brev_t" + std::to_string(t) + ".bin", out.data(), n * 32);
  }
  bit_rev_step1_parallel(out2.data(), in.data(), map2.data(), lg, 4, n);
  write("brev_idx.bin", out2.data(), 2 * n * 32);
  std::cout << "host helpers ok" << std::endl;
  return 0;
}

int main(int argc, char** argv) {
  if (argc == 4 && std::string(argv[1]) == "--host-helpers") return host_helpers(argv[2], std::atoi(argv[3]));
  if (argc != 3) {
    std::cerr << "Usage: " << argv[0] << " <out_dir> <lg>\n";
    return 1;
  }
  try {
    const std::string dir(argv[1]);
    const int lg = std::atoi(argv[2]);
    const uint64_t n = 1ull << lg;
    SyncedMemory x = seeded(n, 1001), y = seeded(n, 1002), pt = seeded(1, 1003), c = seeded(1, 1004);
    dump(dir + "/x.bin", x);
    dump(dir + "/y.bin", y);
    dump(dir + "/pt.bin", pt);
    dump(dir + "/c.bin", c);
    Ntt ntt(32);
    Intt intt(32);
    x.mutable_gpu_data();
    SyncedMemory X = ntt.forward(x);
    dump(dir + "/ntt.bin", X);
    SyncedMemory xi = intt.forward(X);
    dump(dir + "/intt.bin", xi);
    // accumulate_mul_poly in place + inclusive output
    SyncedMemory g = seeded(n, 1001), gi(n * 32);
    g.mutable_gpu_data();
    accumulate_mul_poly_(g, gi);
    dump(dir + "/gp.bin", g);
    dump(dir + "/gp_incl.bin", gi);
    SyncedMemory e = evaluate(x, pt);
    dump(dir + "/eval.bin", e);
    SyncedMemory d = seeded(n, 1001);
    d.mutable_gpu_data();
    SyncedMemory q = poly_div_poly(d, c);
    dump(dir + "/div.bin", q);
    y.mutable_gpu_data();
    dump(dir + "/add.bin", add_mod(x, y));
    dump(dir + "/mul.bin", mul_mod(x, y));
    dump(dir + "/div_ew.bin", div_mod(x, y));
    SyncedMemory xc = seeded(n, 1001), yc = seeded(n, 1002);  // heads at CPU -> CPU path
    dump(dir + "/add_cpu.bin", add_mod(xc, yc));
    dump(dir + "/mul_cpu.bin", mul_mod(xc, yc));
    dump(dir + "/exp5.bin", exp_mod(x, 5));
    dump(dir + "/exp5_cpu.bin", exp_mod(seeded(n, 1001), 5));
    dump(dir + "/seq.bin", gen_sequence(n + 5, pt));
    dump(dir + "/rep.bin", repeat_to_poly(pt, 33));
    // our_lde with lambda = 2 through the C-ABI
    pl_lde_domain* ld = nullptr;
    cli::check(pl_lde_domain_create(lg, 2, nullptr, &ld), "pl_lde_domain_create");
    SyncedMemory L(n * 4 * 32);
    cli::check(pl_lde(x.mutable_gpu_data(), L.mutable_gpu_data(), ld, nullptr, nullptr), "pl_lde");
    pl_lde_domain_destroy(ld);
    dump(dir + "/lde.bin", L);
    // Ntt::stage_no_rotate: the whole of x as one chunk at stage 0, and the
    // first 1024 elements as chunk 1 of stages [6, 12)
    SyncedMemory s0 = seeded(n, 1001), s1 = seeded(1024, 1001);
    ntt.stage_no_rotate(s0, lg, 6, 0, 0);
    ntt.stage_no_rotate(s1, lg, 6, 6, 1);
    dump(dir + "/stage0.bin", s0);
    dump(dir + "/stage6.bin", s1);
    bool threw = false;
    try {
      intt.stage_no_rotate(s1, lg, 6, 6, 1);
    } catch (const std::logic_error&) {
      threw = true;
    }
    if (!threw) throw std::runtime_error("Intt::stage_no_rotate did not refuse");
    // Ntt_coset / Intt_coset: one call, three steps, chunked init, coset LDE
    Ntt_coset nc(32, lg);
    Intt_coset ic(32, lg);
    SyncedMemory xs = seeded(n, 1001);
    xs.mutable_gpu_data();
    dump(dir + "/coset.bin", nc.forward(xs));
    dump(dir + "/icoset.bin", ic.forward(xs));
    SyncedMemory st(n * 32), sti(n * 32);
    SyncedMemory a1 = seeded(n, 1001), a2 = seeded(n, 1001);
    nc.forward1(a1, lg, 0, 0);
    nc.forward2(a1, lg, 0, 0);
    nc.forward3(a1, st, lg, 0, 0, 0);
    ic.forward1(a2, lg, 0, 0);
    ic.forward2(a2, lg, 0, 0);
    ic.forward3(a2, sti, lg, 0, 0, 0);
    dump(dir + "/coset_steps.bin", st);
    dump(dir + "/icoset_steps.bin", sti);
    SyncedMemory ch = seeded(n / 4, 1001);  // chunk 0 of x, then chunk 0 scaled back by Intt_coset::init
    nc.init(ch, lg, 0);
    dump(dir + "/coset_init0.bin", ch);
    SyncedMemory lc(n * 8 * 32);
    nc.init_and_forward(xs, lc, lg, 0, 3, 0, false);
    dump(dir + "/coset_lde3.bin", lc);
    std::cout << "api_check ok" << std::endl;
    return 0;
  } catch (const std::exception& ex) {
    std::cerr << "error: " << ex.what() << std::endl;
    return 1;
  }
}
