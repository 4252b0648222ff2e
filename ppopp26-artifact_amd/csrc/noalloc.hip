// The grand product with library-held scratch: pl_grand_product_noalloc.
//
// For callers that cannot allocate (code under HIP graph capture, tight
// request loops): the prefix product's block partials live in a file-scope
// device array instead of caller workspace or stream-ordered allocations.
// Each GPU the library runs on holds its own copy of the array until the
// process ends.  The kernels are pl_grand_product's; only the scratch differs.
#include <mutex>

#include "pl_internal.hpp"

namespace {

// 2^28 elements (the library's largest size) need 2 (2^16 + 2^4) + 1 block
// partials; the array holds that much, 4 MiB
constexpr uint64_t kMaxN = 1ull << 28;
constexpr uint64_t kScratchElems = 2 * ((kMaxN >> 12) + (kMaxN >> 24)) + 1;

static __device__ __align__(256) uint4 scratch[2 * kScratchElems];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
pl::fr* g_scratch[kMaxDevices] = {};

// The array's device address on the current device, looked up once per
// device with that device current (the host-side name is not an address).
hipError_t scratch_address(pl::fr** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> g(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[dev] = (pl::fr*)p;
  }
  *out = g_scratch[dev];
  return hipSuccess;
}

}  // namespace

// No counterpart in the reference: pl_grand_product (accumulate_mul_poly,
// function.cu:256-262) for callers that cannot allocate.
extern "C" int pl_grand_product_noalloc(void* d_inout, uint64_t n, void* d_incl_out, pl_stream_t stream) {
  if (!d_inout || n == 0) return PL_ERR_ARG;
  if (((uintptr_t)d_inout & 15) || ((uintptr_t)d_incl_out & 15)) return PL_ERR_ARG;
  if (n > kMaxN || pl::grand_product_scratch_elems(n) > kScratchElems) return PL_ERR_UNSUPPORTED;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return PL_ERR_NO_DEVICE;
  pl::fr* sc = nullptr;
  hipError_t e = scratch_address(&sc);
  if (e == hipSuccess) e = pl::grand_product((pl::fr*)d_inout, (pl::fr*)d_incl_out, n, sc, (hipStream_t)stream);
  return e == hipSuccess ? PL_OK : pl::hip_status(e);
}
