// Fast pass kernels of the general path for float stage input and double output (FP32 solves; see
// rowpass_kernels.hpp).
#include "rowpass_kernels.hpp"

namespace flups {

FLUPS_ROWPASS_IO(rowpass_fd, float, double)

}  // namespace flups
