// walk_launch.h -- host-side launchers of the walk kernel instantiations.
#pragma once

#include <hip/hip_runtime.h>

namespace dmw {

struct WalkArgs;

// Measurement-block sizes compiled for every (geometry, wall) pair (96 and 128
// without compensation only); the parked compensated phase runs blocks of 128.
constexpr int kMeasBlocks[] = {4, 16, 32, 64, 96, 128};
constexpr int kParkBlock = 128;

hipError_t launch_walk_free(int wall, int mb, bool kahan, bool park, dim3 grid, size_t lds,
                           hipStream_t st, const WalkArgs& a);
hipError_t launch_walk_box1d(int wall, int mb, bool kahan, bool park, dim3 grid, size_t lds,
                           hipStream_t st, const WalkArgs& a);
hipError_t launch_walk_cyl(int wall, int mb, bool kahan, bool park, dim3 grid, size_t lds,
                           hipStream_t st, const WalkArgs& a);
hipError_t launch_walk_sph(int wall, int mb, bool kahan, bool park, dim3 grid, size_t lds,
                           hipStream_t st, const WalkArgs& a);
hipError_t launch_walk_pcyl(int wall, int mb, bool kahan, bool park, dim3 grid, size_t lds,
                           hipStream_t st, const WalkArgs& a);
hipError_t launch_walk_psph(int wall, int mb, bool kahan, bool park, dim3 grid, size_t lds,
                           hipStream_t st, const WalkArgs& a);
hipError_t launch_walk_ell(int wall, int mb, bool kahan, bool park, dim3 grid, size_t lds,
                           hipStream_t st, const WalkArgs& a);
hipError_t launch_walk_slab(int wall, int mb, bool kahan, bool park, dim3 grid, size_t lds,
                           hipStream_t st, const WalkArgs& a);
hipError_t launch_walk_shell(int wall, int mb, bool kahan, bool park, dim3 grid, size_t lds,
                           hipStream_t st, const WalkArgs& a);
hipError_t launch_walk_shellc(int wall, int mb, bool kahan, bool park, dim3 grid, size_t lds,
                           hipStream_t st, const WalkArgs& a);
hipError_t launch_walk_myel(int wall, int mb, bool kahan, bool park, dim3 grid, size_t lds,
                           hipStream_t st, const WalkArgs& a);
hipError_t launch_walk_pmyel(int wall, int mb, bool kahan, bool park, dim3 grid, size_t lds,
                           hipStream_t st, const WalkArgs& a);

// Fixed-order float64 sum of per-wave partials: out[c] = scale * sum_w in[w*n_cols + c].
hipError_t launch_reduce_partials(const double* in, long long n_rows, int n_cols, double scale,
                                  double* out, hipStream_t st);

}  // namespace dmw

// Instantiates launch_walk_<NAME> for geometry GEOM with the listed walls.
#define DMW_DEFINE_LAUNCHER(NAME, GEOM, ALLOW_SURFACE, ALLOW_PERMEATE)                          \
  namespace dmw {                                                                               \
  template <int W, int MB, bool K, bool P = false>                                              \
  static hipError_t go_##NAME(dim3 grid, size_t lds, hipStream_t st, const WalkArgs& a) {       \
    hipLaunchKernelGGL((walk_kernel<GEOM, W, MB, K, P>), grid, dim3(kBlock), lds, st, a);       \
    return hipGetLastError();                                                                   \
  }                                                                                             \
  template <int W>                                                                              \
  static hipError_t mb_##NAME(int mb, bool k, bool park, dim3 g, size_t lds, hipStream_t st,    \
                              const WalkArgs& a) {                                              \
    if (park) return mb == kParkBlock ? go_##NAME<W, kParkBlock, true, true>(g, lds, st, a)      \
                                      : hipErrorInvalidValue;                                   \
    switch (mb) {                                                                               \
      case 4: return k ? go_##NAME<W, 4, true>(g, lds, st, a) : go_##NAME<W, 4, false>(g, lds, st, a); \
      case 16: return k ? go_##NAME<W, 16, true>(g, lds, st, a) : go_##NAME<W, 16, false>(g, lds, st, a); \
      case 32: return k ? go_##NAME<W, 32, true>(g, lds, st, a) : go_##NAME<W, 32, false>(g, lds, st, a); \
      case 64: return k ? go_##NAME<W, 64, true>(g, lds, st, a) : go_##NAME<W, 64, false>(g, lds, st, a); \
      case 96: return k ? hipErrorInvalidValue : go_##NAME<W, 96, false>(g, lds, st, a);            \
      case 128: return k ? hipErrorInvalidValue : go_##NAME<W, 128, false>(g, lds, st, a);          \
      default: return hipErrorInvalidValue;                                                     \
    }                                                                                           \
  }                                                                                             \
  hipError_t launch_walk_##NAME(int wall, int mb, bool k, bool park, dim3 g, size_t lds,        \
                                hipStream_t st, const WalkArgs& a) {                            \
    if (wall == W_REFLECT) return mb_##NAME<W_REFLECT>(mb, k, park, g, lds, st, a);             \
    if (ALLOW_SURFACE && wall == W_SURFACE) return mb_##NAME<(ALLOW_SURFACE ? W_SURFACE : W_REFLECT)>(mb, k, park, g, lds, st, a); \
    if (ALLOW_PERMEATE && wall == W_PERMEATE) return mb_##NAME<(ALLOW_PERMEATE ? W_PERMEATE : W_REFLECT)>(mb, k, park, g, lds, st, a); \
    return hipErrorInvalidValue;                                                                \
  }                                                                                             \
  }
