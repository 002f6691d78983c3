// far_mfma.hip -- placeholder for the MFMA far-field contraction (not yet enabled).
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace dbie {

bool far_field_mfma_supported(int, int) { return false; }

int far_field_mfma(hipStream_t, const DevGeom &, int, int, const double *, int64_t, double *, int64_t) { return 0; }

}  // namespace dbie
