// Pass kernels for columns of 2^7 points (forward/inverse, intermediate/last).
#include "ntt_pass.hpp"

namespace pl {
PL_DEFINE_PASS_LAUNCHER(3)
}  // namespace pl
