// reduce.hip -- deterministic float64 reduction of per-wave partial sums.
//
// Replaces the float32 walker sum of _ens (dmipy_sim/core.py:426-429): each
// column is summed by one workgroup in a fixed order (strided per thread, then a
// fixed LDS tree), so repeated runs give identical bits.
#include <hip/hip_runtime.h>

#include "walk_launch.h"

namespace dmw {

__global__ __launch_bounds__(256) void reduce_partials_kernel(const double* __restrict__ in,
                                                              long long n_rows, int n_cols,
                                                              double scale, double* __restrict__ out) {
  __shared__ double buf[256];
  const int c = blockIdx.x;
  double acc = 0.0;
  for (long long r = threadIdx.x; r < n_rows; r += 256) acc += in[r * n_cols + c];
  buf[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) buf[threadIdx.x] += buf[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = buf[0] * scale;
}

hipError_t launch_reduce_partials(const double* in, long long n_rows, int n_cols, double scale,
                                  double* out, hipStream_t st) {
  if (n_cols <= 0) return hipSuccess;
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(n_cols), dim3(256), 0, st, in, n_rows, n_cols,
                     scale, out);
  return hipGetLastError();
}

}  // namespace dmw
