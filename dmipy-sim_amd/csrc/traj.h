// traj.h -- device-resident trajectory set of the walk-once producer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

struct dmw_traj {
  long long N = 0;           // walkers
  int n_t = 0;               // saved samples
  uint16_t* traj = nullptr;  // (3 n_t, N) float16 bits: row 3t + c
  uint16_t* dlog = nullptr;  // (n_t, N) float16 bits, or null
  uint16_t* comp = nullptr;  // (n_t, N) float16 bits, or null
  double walk_ms = 0.0;      // producer kernel time
  ~dmw_traj() {
    if (traj) (void)hipFree(traj);
    if (dlog) (void)hipFree(dlog);
    if (comp) (void)hipFree(comp);
  }
};

namespace dmw {
// Transpose a (rows, cols) matrix of 16-bit words into (cols, rows) on the device.
hipError_t launch_transpose16(const uint16_t* in, uint16_t* out, long long rows, long long cols, hipStream_t st);
}  // namespace dmw
