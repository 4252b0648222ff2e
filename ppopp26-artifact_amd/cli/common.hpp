// Helpers shared by the NTT, LDE and miss_ops command lines: raw input files,
// seeded synthetic inputs, timing, output comparison (ntt.cu:25-41,
// 1250-1274; lde.cu:222-231; miss_ops.cu:8-24).
#pragma once
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>

#include "../../include/pl.h"
#include "../host/function.hpp"

namespace cli {

inline void check(int st, const char* what) {
  if (st != PL_OK)
    throw std::runtime_error(std::string(what) + ": " + pl_status_string(st) + " " + pl_last_hip_error());
}

inline void hipc(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Uniform Fr in Montgomery form from SplitMix64: 4 draws per element, top
// bit cleared, values >= r redrawn (the seeded-input convention of the tests).
inline void fill_seeded(void* out, uint64_t n, uint64_t seed) {
  static const uint64_t P[4] = {0xffffffff00000001ull, 0x53bda402fffe5bfeull, 0x3339d80809a1d805ull,
                                0x73eda753299d7d48ull};
  uint64_t s = seed;
  auto next = [&]() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  uint64_t* o = (uint64_t*)out;
  for (uint64_t i = 0; i < n; i++) {
    uint64_t l[4];
    for (;;) {
      for (int j = 0; j < 4; j++) l[j] = next();
      l[3] &= 0x7fffffffffffffffull;
      bool ge = true;
      for (int j = 3; j >= 0; j--) {
        if (l[j] != P[j]) {
          ge = l[j] > P[j];
          break;
        }
      }
      if (!ge) break;
    }
    std::memcpy(o + 4 * i, l, 32);
  }
}

inline uint64_t seed_of(int op, int lg, int aux) {
  return 0x504C4E4B00000000ull ^ ((uint64_t)op << 40) ^ ((uint64_t)lg << 8) ^ (uint64_t)aux;
}

// read_file (ntt.cu:25-41): a missing file is reported and the buffer is
// filled with seeded uniform data instead of being left zero.
inline void read_file(const char* filename, void* data, size_t bytes, uint64_t seed) {
  std::ifstream file(filename, std::ios::binary);
  if (!file) {
    std::cerr << "Error opening file: " << filename << std::endl;
    std::cerr << "[input] using seeded synthetic data instead" << std::endl;
    fill_seeded(data, bytes / 32, seed);
    return;
  }
  file.seekg(0, std::ios::end);
  size_t size = file.tellg();
  file.seekg(0, std::ios::beg);
  // the reference reads the whole file into the buffer whatever its size;
  // here the read stops at the buffer and a short file is zero-padded
  if (size != bytes)
    std::cerr << "[input] " << filename << ": " << size << " bytes, buffer " << bytes
              << (size > bytes ? " (extra bytes ignored)" : " (rest zero-filled)") << std::endl;
  file.read(reinterpret_cast<char*>(data), std::min(size, bytes));
  if (!file) std::cerr << "Error reading file: " << filename << std::endl;
  if (size < bytes) std::memset((char*)data + size, 0, bytes - size);
}

inline std::string format_bytes(size_t bytes) {
  const double kb = bytes / 1024.0, mb = kb / 1024.0;
  std::ostringstream out;
  if (mb >= 1.0)
    out << std::fixed << std::setprecision(2) << mb << " MB";
  else
    out << std::fixed << std::setprecision(2) << kb << " KB";
  return out.str();
}

// compare_outputs (ntt.cu:1250-1274)
inline bool compare_outputs(const void* expected, const void* actual, size_t count) {
  const uint64_t* e = (const uint64_t*)expected;
  const uint64_t* a = (const uint64_t*)actual;
  for (size_t i = 0; i < count; i++) {
    if (std::memcmp(e + 4 * i, a + 4 * i, 32) != 0) {
      std::cout << "mismatch at element " << i << "\nexpected:";
      for (int j = 0; j < 4; j++) std::cout << " 0x" << std::hex << e[4 * i + j];
      std::cout << "\nactual:  ";
      for (int j = 0; j < 4; j++) std::cout << " 0x" << std::hex << a[4 * i + j];
      std::cout << std::dec << std::endl;
      return false;
    }
  }
  std::cout << "outputs match for " << count << " elements" << std::endl;
  return true;
}

struct EventTimer {
  hipEvent_t a, b;
  hipStream_t s;
  explicit EventTimer(hipStream_t st = nullptr) : s(st) {
    hipc(hipEventCreate(&a), "hipEventCreate");
    hipc(hipEventCreate(&b), "hipEventCreate");
  }
  ~EventTimer() {
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
  }
  void start() { hipc(hipEventRecord(a, s), "hipEventRecord"); }
  float stop_ms() {
    hipc(hipEventRecord(b, s), "hipEventRecord");
    hipc(hipEventSynchronize(b), "hipEventSynchronize");
    float ms = 0;
    hipc(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
    return ms;
  }
};

// chunked copies between a pinned host buffer and the device (2^lg_k chunks)
inline void copy_chunked(void* dst, const void* src, size_t bytes, int chunks, hipMemcpyKind kind, hipStream_t s) {
  if (chunks < 1) chunks = 1;
  size_t step = (bytes + chunks - 1) / chunks;
  step = (step + 255) & ~size_t(255);
  for (size_t off = 0; off < bytes; off += step) {
    size_t len = std::min(step, bytes - off);
    hipc(hipMemcpyAsync((char*)dst + off, (const char*)src + off, len, kind, s), "hipMemcpyAsync");
  }
}

}  // namespace cli
