// FP32-MFMA right-hand side (placeholder until the MFMA kernels land).
#include <hip/hip_runtime.h>

#include "launch.h"

namespace spoc {
int rhs_f32(int, int, int, const RhsArgs &, void *) { return -1; }
}  // namespace spoc
