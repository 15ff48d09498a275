// arith_check.hip -- test infrastructure: checks the walk's unscaled correctly
// rounded square root and quotients (dmipy-sim_amd/csrc/rn_arith.h) against the
// compiler's IEEE sqrtf and '/' on the GPU, bit for bit.
//   arith_check_sqrt: every float bit pattern in [lo, hi) (grid-stride);
//   arith_check_div:  n operand pairs from a counter hash, each operand with a
//                     random sign, mantissa and an exponent in [-emax, emax]; pairs
//                     whose quotient exponent leaves [-100, 100] are skipped; one
//                     pair in 16 uses an all-ones mantissa divisor, one in 64 a zero
//                     numerator.  known_divisor = 1 checks div_rn_r with the
//                     reciprocal RN(1/y) taken from IEEE division;
//   arith_check_recip (host only): recip_rn(y) against IEEE 1.0f / y;
//   arith_check_sqrt_cut (host only): sqrt_cut(r) is the least float whose IEEE
//                     square root reaches r.
// Mismatches are counted; the first one found is reported.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../dmipy-sim_amd/csrc/rn_arith.h"

namespace {

__device__ __forceinline__ uint32_t mix(uint64_t v) {
  v ^= v >> 33; v *= 0xff51afd7ed558ccdull; v ^= v >> 33; v *= 0xc4ceb9fe1a85ec53ull; v ^= v >> 33;
  return (uint32_t)v;
}

__device__ __forceinline__ float make_float(uint32_t h1, uint32_t h2, int emax, bool ones) {
  const uint32_t sign = h1 & 0x80000000u;
  const int e = (int)(h1 % (uint32_t)(2 * emax + 1)) - emax;
  const uint32_t mant = ones ? 0x7FFFFFu : (h2 & 0x7FFFFFu);
  return __builtin_bit_cast(float, sign | ((uint32_t)(e + 127) << 23) | mant);
}

__global__ void sqrt_kernel(uint32_t lo, uint32_t hi, unsigned long long* bad, uint32_t* first) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t b = lo + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; b < hi; b += stride) {
    const float x = __builtin_bit_cast(float, (uint32_t)b);
    const float ref = sqrtf(x);
    const float got = dmw::sqrt_rn(x);
    if (__builtin_bit_cast(uint32_t, ref) != __builtin_bit_cast(uint32_t, got)) {
      if (atomicAdd(bad, 1ull) == 0ull) first[0] = (uint32_t)b;
    }
  }
}

__global__ void div_kernel(uint64_t n, uint32_t seed, int emax, int known_divisor, unsigned long long* bad,
                           unsigned long long* skipped, uint32_t* first) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint64_t c = (i << 2) ^ ((uint64_t)seed << 40);
    const uint32_t h0 = mix(c), h1 = mix(c + 1), h2 = mix(c + 2), h3 = mix(c + 3);
    float x = make_float(h0, h1, emax, false);
    const float y = make_float(h2, h3, emax, (h3 >> 28) == 0u);
    if ((h1 >> 26) == 0u) x = (h0 & 1u) ? -0.0f : 0.0f;
    const float ref = x / y;
    const uint32_t rb = __builtin_bit_cast(uint32_t, ref);
    const int qe = (int)((rb >> 23) & 0xFFu) - 127;
    if (x != 0.0f && (qe < -100 || qe > 100)) {
      atomicAdd(skipped, 1ull);
      continue;
    }
    const float got = known_divisor ? dmw::div_rn_r(x, y, 1.0f / y) : dmw::div_rn(x, y);
    if (rb != __builtin_bit_cast(uint32_t, got)) {
      if (atomicAdd(bad, 1ull) == 0ull) {
        first[0] = __builtin_bit_cast(uint32_t, x);
        first[1] = __builtin_bit_cast(uint32_t, y);
      }
    }
  }
}

int run(unsigned long long* out_bad, unsigned long long* out_skipped, uint32_t* out_first,
        bool is_div, uint32_t lo, uint32_t hi, uint64_t n, uint32_t seed, int emax, int known_divisor) {
  unsigned long long* d = nullptr;
  uint32_t* f = nullptr;
  if (hipMalloc(&d, 16) != hipSuccess || hipMalloc(&f, 8) != hipSuccess) return 1;
  if (hipMemset(d, 0, 16) != hipSuccess || hipMemset(f, 0, 8) != hipSuccess) return 1;
  if (is_div)
    hipLaunchKernelGGL(div_kernel, dim3(8192), dim3(256), 0, 0, n, seed, emax, known_divisor, d, d + 1, f);
  else
    hipLaunchKernelGGL(sqrt_kernel, dim3(8192), dim3(256), 0, 0, lo, hi, d, f);
  if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) return 1;
  unsigned long long h[2];
  if (hipMemcpy(h, d, 16, hipMemcpyDeviceToHost) != hipSuccess) return 1;
  if (hipMemcpy(out_first, f, 8, hipMemcpyDeviceToHost) != hipSuccess) return 1;
  *out_bad = h[0];
  if (out_skipped) *out_skipped = h[1];
  (void)hipFree(d);
  (void)hipFree(f);
  return 0;
}

}  // namespace

extern "C" int arith_check_sqrt(uint32_t lo, uint32_t hi, unsigned long long* bad, uint32_t* first) {
  return run(bad, nullptr, first, false, lo, hi, 0, 0, 0, 0);
}

extern "C" int arith_check_div(uint64_t n, uint32_t seed, int emax, int known_divisor, unsigned long long* bad,
                               unsigned long long* skipped, uint32_t* first) {
  return run(bad, skipped, first, true, 0, 0, n, seed, emax, known_divisor);
}

// Host: recip_rn(y) == 1.0f / y (IEEE float32 division of the host compiler) for n
// random normal y with exponents in [-emax, emax], plus every mantissa of [1, 2)
// when exhaustive is set.
extern "C" int arith_check_recip(uint64_t n, uint32_t seed, int emax, int exhaustive, unsigned long long* bad,
                                 uint32_t* first) {
  unsigned long long nb = 0;
  auto check = [&](float y) {
    volatile float one = 1.0f;
    const float ref = one / y;
    const float got = dmw::recip_rn(y);
    if (__builtin_bit_cast(uint32_t, ref) != __builtin_bit_cast(uint32_t, got)) {
      if (nb++ == 0) *first = __builtin_bit_cast(uint32_t, y);
    }
  };
  uint64_t v = 0x9E3779B97F4A7C15ull ^ ((uint64_t)seed << 32);
  for (uint64_t i = 0; i < n; ++i) {
    v ^= v << 13; v ^= v >> 7; v ^= v << 17;
    const uint32_t h = (uint32_t)v, g = (uint32_t)(v >> 32);
    const int e = (int)(g % (uint32_t)(2 * emax + 1)) - emax;
    check(__builtin_bit_cast(float, (h & 0x807FFFFFu) | ((uint32_t)(e + 127) << 23)));
  }
  if (exhaustive)
    for (uint32_t m = 0; m < (1u << 23); ++m) check(__builtin_bit_cast(float, 0x3F800000u | m));
  *bad = nb;
  return 0;
}

// Host: for n random r (exponents in [-emax, emax]), sqrtf(cut) >= r > sqrtf(cut-),
// with cut = sqrt_cut(r) and cut- its predecessor.
extern "C" int arith_check_sqrt_cut(uint64_t n, uint32_t seed, int emax, unsigned long long* bad,
                                    uint32_t* first) {
  unsigned long long nb = 0;
  uint64_t v = 0xD1B54A32D192ED03ull ^ ((uint64_t)seed << 32);
  for (uint64_t i = 0; i < n; ++i) {
    v ^= v << 13; v ^= v >> 7; v ^= v << 17;
    const uint32_t h = (uint32_t)v, g = (uint32_t)(v >> 32);
    const int e = (int)(g % (uint32_t)(2 * emax + 1)) - emax;
    const float r = __builtin_bit_cast(float, (h & 0x007FFFFFu) | ((uint32_t)(e + 127) << 23));
    const float cut = dmw::sqrt_cut(r);
    const float below = std::nextafter(cut, 0.0f);
    if (!(std::sqrt(cut) >= r && std::sqrt(below) < r)) {
      if (nb++ == 0) *first = __builtin_bit_cast(uint32_t, r);
    }
  }
  *bad = nb;
  return 0;
}
