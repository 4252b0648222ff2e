// Pass kernels for columns of 2^(LGPER+4) points when LGPER >= 4 (none with 8 elements per lane).
#include "ntt_pass.hpp"

namespace pl {
#if PL_NTT_LGPER >= 4
PL_DEFINE_PASS_LAUNCHER(4)
#else
void launch_pass_la4(bool, bool, const PassArgs&, hipStream_t) {}
void launch_pass_ragged_la4(bool, bool, bool, const PassArgs*, const uint32_t*, uint32_t, uint32_t, hipStream_t) {}
#endif
}  // namespace pl
