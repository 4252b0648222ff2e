// Pass kernels for columns of 2^5 points (forward/inverse, intermediate/last).
#include "ntt_pass.hpp"

namespace pl {
PL_DEFINE_PASS_LAUNCHER(1)
}  // namespace pl
