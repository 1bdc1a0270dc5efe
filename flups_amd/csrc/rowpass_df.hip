// Fast pass kernels of the general path for double stage input and float output (FP32 solves; see
// rowpass_kernels.hpp).
#include "rowpass_kernels.hpp"

namespace flups {

FLUPS_ROWPASS_IO(rowpass_df, double, float)

}  // namespace flups
