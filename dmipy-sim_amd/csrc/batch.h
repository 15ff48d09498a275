// batch.h -- error helpers shared by the batched launcher and abi.cpp.
#pragma once

#include <hip/hip_runtime.h>

namespace dmw {

// Record a failure for dmw_last_error() on this thread and return its code.
int batch_fail(int code, const char* msg);
// Map a HIP error to a dmw_status ("out of memory" for allocation failures).
int batch_hip_fail(hipError_t e, const char* what);

}  // namespace dmw
