// Pass kernels for columns of 2^6 points (forward/inverse, intermediate/last).
#include "ntt_pass.hpp"

namespace pl {
PL_DEFINE_PASS_LAUNCHER(2)
}  // namespace pl
