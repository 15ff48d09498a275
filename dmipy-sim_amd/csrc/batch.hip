// batch.hip -- one persistent launch walking thousands of substrate instances.
//
// The reference runs one simulate() call per substrate (core.py:193), e.g. the
// serial loop of simulate_mixture (core.py:821-858).  Here every instance (any
// analytic geometry with its own wall, D, sizes, rho, kappa, walker count,
// sub-step count and waveform set) is cut into 64-walker tiles; a persistent
// grid of wavefronts pops runs of consecutive tiles of one instance (the
// walkers a wavefront takes per pop: 64, 128, 256 or 512) from a queue ordered
// by decreasing cost (n_sub), so the long instances start first and the short
// ones fill the tail.  The sub-step loop is compiled plain and unrolled by two;
// both knobs, and the number of resident wavefronts, are chosen by measurement
// (autotune.py).  Packed instances keep their object tables and periodic neighbour grids
// in one concatenated device array (read through the L1/L2 caches); permeable
// packings walk the single walk's wall screen in a second persistent kernel with
// its own queue, their phase parked in device memory between samples.  Each unit
// writes float64 partial sums; a second kernel reduces every instance's units
// in a fixed order.  The three launches (queue reset, walk, reduce) are
// captured once in a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/dmipy_walk.h"
#include "batch.h"
#include "envelope.h"
#include "grid.h"
#include "phase16.h"
#include "walk_kernel.h"

// Occupancy budget of the screened kernel (wavefronts per SIMD).
#ifndef DMW_SCREENED_WAVES
#define DMW_SCREENED_WAVES 4
#endif

namespace dmw {

struct InstDev {
  float size, eps, nudge, rhoD, kappaD, step_len, dt_sub, inv_T2, gdt;
  float inv_size;            // RN(1 / size)
  float size_sq_cut;         // least float s with sqrtf(s) >= size
  float size_sq_cut_gt;      // least float s with sqrtf(s) > size
  float cell, inv_cell, grid_inv_h, grid_h;
  // screened permeable packings (walk_kernel.h PackedScreen): cache flag, box
  // width, margin terms
  int scr_cache;
  float scr_w, scr_m2, scr_cout, scr_cin;
  float ax, ay, az, aux;
  int kind, wall, n_sub, has_t2;
  int grid_n, exact_minimage, n_obj;
  int obj_off, start_off, items_off;
  long long n_walkers;
  long long g_off;          // offset (floats) of the instance's waveform set
  long long gh_off, gsum_off;  // offsets of its float16 phase operands (uint2) and column sums (floats)
  float p_scale, inv_pg_scale;  // matrix-core phase scales (powers of two)
  int unit_begin, unit_end;  // the instance's tiles: partial rows [unit_begin, unit_end)
  unsigned long long seed_offset;
  unsigned int stream;       // Philox instance word
  int pad_;
};

struct BatchArgs {
  const InstDev* inst;
  const int2* units;        // queue entries (instance, first tile) in queue order
  int n_units;              // queue entries
  int tiles_per_pop;        // tiles an entry covers (fewer at an instance's end)
  const float* G;           // sets x (n_t, n_meas_pad, 3) T/m, not scaled
  const uint2* GH;          // sets x (n_t, 2, n_meas_pad) float16 x4 hi/lo parts of G * g_scale(set)
  const float* Gsum;        // sets x (3, n_meas_pad): sum over t of G
  const float* chi;
  int n_t, n_meas, n_meas_pad;
  const float4* objs;       // concatenated object tables of packed instances
  const int* starts;        // concatenated grid cell starts
  const int* items;         // concatenated grid cell lists
  const unsigned long long* seed;  // device, read at kernel start
  unsigned int* queue;      // work counter
  double* partials;         // (tiles, n_meas_pad, 2), row unit_begin + tile of the instance
  float4* park;             // screened kernel: parked phase rows of every resident wavefront
};

template <int GEOM, int WALL, int MB, int UNROLL, bool SCR = false>
__device__ __forceinline__ void walk_unit(const BatchArgs& b, const InstDev& I, int tile, int unit,
                                          unsigned long long base_seed) {
  const unsigned long long seed = base_seed + I.seed_offset;
  const uint32_t seed_lo = (uint32_t)seed, seed_hi = (uint32_t)(seed >> 32);
  const uint32_t inst = I.stream;
  constexpr bool PACKED = (GEOM == G_PCYL || GEOM == G_PSPH);
  constexpr int DIM = (GEOM == G_PSPH) ? 3 : 2;
  // permeable packings walk the wall screen, as the single walk does (in the
  // screened kernel, which takes exactly those instances)
  constexpr bool SCREEN = SCR && PACKED && WALL == W_PERMEATE;
  const int lane = threadIdx.x & (kWave - 1);
  const long long widx = (long long)tile * kWave + lane;
  const bool active = widx < I.n_walkers;
  const uint32_t w = (uint32_t)(active ? widx : I.n_walkers - 1);
  WalkArgs a{};
  a.size = I.size; a.inv_size = I.inv_size; a.size_sq_cut = I.size_sq_cut;
  a.size_sq_cut_gt = I.size_sq_cut_gt; a.eps = I.eps; a.nudge = I.nudge; a.rhoD = I.rhoD; a.kappaD = I.kappaD;
  a.cell = I.cell; a.inv_cell = I.inv_cell; a.grid_n = I.grid_n; a.grid_inv_h = I.grid_inv_h;
  a.grid_h = I.grid_h;
  a.scr_cache = I.scr_cache; a.scr_w = I.scr_w; a.scr_m2 = I.scr_m2; a.scr_cout = I.scr_cout;
  a.scr_cin = I.scr_cin;
  a.exact_minimage = I.exact_minimage; a.n_obj = I.n_obj;
  a.ax = I.ax; a.ay = I.ay; a.az = I.az; a.aux = I.aux;
  a.instance = inst; a.seed_lo = seed_lo; a.seed_hi = seed_hi;
  a.Gs = b.G + I.g_off;
  a.GsH = b.GH + I.gh_off;
  a.Gsum = b.Gsum + I.gsum_off;
  a.p_scale = I.p_scale;
  a.inv_pg_scale = I.inv_pg_scale;
  a.n_meas_pad = b.n_meas_pad;
  a.n_meas = b.n_meas;
  a.park = b.park;
  Tables tb{nullptr, nullptr, nullptr};
  if (PACKED) tb = Tables{b.objs + I.obj_off, b.starts + I.start_off, b.items + I.items_off};
  Pos p = sample_start<GEOM>(w, a, tb);
  // the screened kernel's 64- and 128-measurement variants park the compensated
  // phase in device memory between samples (the sub-step loop then holds no phase
  // registers)
  PhaseAcc<MB, false, SCREEN && MB >= 32> ph;
  ph.init();
  float logw = 0.0f;
  uint32_t hits = 0, ncross = 0, ncand = 0;
  const int n_sub = I.n_sub;
  const float fsub = (float)n_sub;
  const float L = I.cell;
  // start point of the matrix-core phase, in the units of the update's operand (x gdt)
  if (PACKED)
    ph.set_ref(a, 0, lane, (fsub * (p.x + L * (float)p.cx)) * I.gdt, (fsub * (p.y + L * (float)p.cy)) * I.gdt,
               (fsub * (p.z + L * (float)p.cz)) * I.gdt);
  else
    ph.set_ref(a, 0, lane, (fsub * p.x) * I.gdt, (fsub * p.y) * I.gdt, (fsub * p.z) * I.gdt);
  PackedScreen<DIM> scr;
  if (SCREEN) scr.refresh(p, a, tb);
  u32x4 rnd{0u, 0u, 0u, 0u};
  for (int t = 0; t < b.n_t; ++t) {
    ph.prefetch(a, t, 0, lane);
    const float chi = b.chi ? b.chi[t] : 1.0f;
    const float q2chi = I.has_t2 ? (I.dt_sub * I.inv_T2) * chi : 0.0f;  // zero when off (x - 0 == x)
    const float x0 = p.x, y0 = p.y, z0 = p.z;
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
#pragma unroll UNROLL
    for (int s = 0; s < n_sub; ++s) {
      const uint32_t gs = (uint32_t)t * (uint32_t)n_sub + (uint32_t)s;
      if (!(gs & 1u) || (t == 0 && s == 0))
        rnd = philox4x32_10({w, inst, gs >> 1, TAG_STEP}, seed_lo, seed_hi);
      float ux, uy, uz;
      iso_direction((gs & 1u) ? rnd.z : rnd.x, (gs & 1u) ? rnd.w : rnd.y, &ux, &uy, &uz);
      const float sx = ux * I.step_len, sy = uy * I.step_len, sz = uz * I.step_len;
      const PowlesSrc ps{w, inst, gs, seed_lo, seed_hi, 0.0f, false};
      float dlog = SCREEN ? screened_permeate<DIM>(p, sx, sy, sz, a, tb, scr, ps, hits, ncross, ncand)
                          : wall_step<GEOM, WALL>(p, sx, sy, sz, a, tb, ps, hits, ncross, ncand);
      if (WALL != W_REFLECT) dlog = dlog * chi;
      dlog = dlog - q2chi;
      logw = logw + dlog;
      ax = ax + (p.x - x0);   // read only when n_sub > 1
      ay = ay + (p.y - y0);
      az = az + (p.z - z0);
    }
    float ox = x0, oy = y0, oz = z0, ex = p.x, ey = p.y, ez = p.z;
    if (PACKED) {
      ox = ox + L * (float)p.cx; oy = oy + L * (float)p.cy; oz = oz + L * (float)p.cz;
      ex = ex + L * (float)p.cx; ey = ey + L * (float)p.cy; ez = ez + L * (float)p.cz;
    }
    float Px, Py, Pz;
    if (n_sub > 1) {
      Px = fmaf(fsub, ox, ax); Py = fmaf(fsub, oy, ay); Pz = fmaf(fsub, oz, az);
    } else {
      Px = ex; Py = ey; Pz = ez;
    }
    ph.update(a, t, 0, lane, Px * I.gdt, Py * I.gdt, Pz * I.gdt);
    if (PACKED) fold<DIM>(p, a);
    if (SCREEN) scr.refresh(p, a, tb);   // the folded frame: cell, box and images anew
  }
  const float wgt = expf(logw);
  ph.emit(a, 0, wgt, active, lane, b.partials + (size_t)unit * b.n_meas_pad * 2);
  (void)hits; (void)ncross; (void)ncand;
}

// Four wavefronts per SIMD: the 64-measurement variant fits 128 VGPRs with 280 bytes of
// scratch; with a persistent grid of four workgroups per CU the 4096-instance
// Sphere/Cylinder batch runs in 1281 ms against 1354 ms at three (168 VGPRs, 68 bytes of
// scratch) and 1652 ms at two (187 VGPRs): the per-unit instance loads and dispatch are
// latency that more resident wavefronts hide.  Five (96 VGPRs, 372 bytes of scratch) is
// slower again, 1314 ms.  Permeable packings walk in a kernel of their own
// (SCREENED): the wall screen's per-walker state would raise every other geometry's
// register pressure in a shared kernel.
template <int MB, int UNROLL, bool SCREENED>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(SCREENED ? DMW_SCREENED_WAVES : 4)))
void batch_kernel(const BatchArgs b) {
  const unsigned long long seed = *b.seed;
  const unsigned int inc = (threadIdx.x & (kWave - 1)) == 0 ? 1u : 0u;
  for (;;) {
    // Pop one unit per wave.  Every lane executes the atomic (lane 0 adds 1, the
    // others 0) so the pop sits in uniform control flow, and the position is
    // taken from lane 0 with readfirstlane: a scalar, so the exit test, the
    // instance loads and the dispatch below are uniform branches.  (A pop under
    // `if (lane == 0)` followed by a lane shuffle lets the optimiser thread the
    // loop back edge through the lane test, and lanes 1..63 then shuffle a
    // position lane 0 has not fetched.)
    const unsigned int old = atomicAdd(b.queue, inc);
    const int pos = (int)__builtin_amdgcn_readfirstlane(old);
    if (pos >= b.n_units) break;  // uniform: every wave reaches this exit
    const int2 u = b.units[pos];
    const int tile_end = min(u.y + b.tiles_per_pop, b.inst[u.x].unit_end - b.inst[u.x].unit_begin);
    for (int tile = u.y; tile < tile_end; ++tile) {
    // re-read per tile (scalar loads) rather than held across the walk
    const InstDev I = b.inst[u.x];
    if (SCREENED) {
      if (I.kind == G_PCYL)
        walk_unit<G_PCYL, W_PERMEATE, MB, UNROLL, true>(b, I, tile, I.unit_begin + tile, seed);
      else
        walk_unit<G_PSPH, W_PERMEATE, MB, UNROLL, true>(b, I, tile, I.unit_begin + tile, seed);
      continue;
    }
    switch (I.kind * 3 + I.wall) {
#define DMW_UNIT(G, W) \
      case G * 3 + W: walk_unit<G, W, MB, UNROLL>(b, I, tile, I.unit_begin + tile, seed); break;
      DMW_UNIT(G_FREE, W_REFLECT)
      DMW_UNIT(G_BOX1D, W_REFLECT)
      DMW_UNIT(G_BOX1D, W_SURFACE)
      DMW_UNIT(G_CYL, W_REFLECT)
      DMW_UNIT(G_CYL, W_SURFACE)
      DMW_UNIT(G_CYL, W_PERMEATE)
      DMW_UNIT(G_SPH, W_REFLECT)
      DMW_UNIT(G_SPH, W_SURFACE)
      DMW_UNIT(G_SPH, W_PERMEATE)
      DMW_UNIT(G_PCYL, W_REFLECT)
      DMW_UNIT(G_PCYL, W_SURFACE)
      DMW_UNIT(G_PSPH, W_REFLECT)
      DMW_UNIT(G_PSPH, W_SURFACE)
      DMW_UNIT(G_ELL, W_REFLECT)
      DMW_UNIT(G_ELL, W_SURFACE)
      DMW_UNIT(G_ELL, W_PERMEATE)
      DMW_UNIT(G_SLAB, W_REFLECT)
      DMW_UNIT(G_SLAB, W_SURFACE)
      DMW_UNIT(G_SLAB, W_PERMEATE)
      DMW_UNIT(G_SHELL, W_REFLECT)
      DMW_UNIT(G_SHELL, W_PERMEATE)
      DMW_UNIT(G_SHELLC, W_REFLECT)
      DMW_UNIT(G_SHELLC, W_PERMEATE)
#undef DMW_UNIT
      default: break;
    }
    }
  }
}

// The two queue counters (general kernel, screened kernel).
__global__ void batch_reset_kernel(unsigned int* queue) { queue[0] = 0u; queue[1] = 0u; }

// Fixed-order per-instance sum: one workgroup per instance, one thread per column.
__global__ __launch_bounds__(256) void batch_reduce_kernel(const InstDev* __restrict__ inst,
                                                           const double* __restrict__ partials,
                                                           int n_cols, double* __restrict__ out) {
  const InstDev I = inst[blockIdx.x];
  const double scale = 1.0 / (double)I.n_walkers;
  for (int c = threadIdx.x; c < n_cols; c += blockDim.x) {
    double acc = 0.0;
    for (int u = I.unit_begin; u < I.unit_end; ++u) acc += partials[(size_t)u * n_cols + c];
    out[(size_t)blockIdx.x * n_cols + c] = acc * scale;
  }
}

}  // namespace dmw

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
struct dmw_batch {
  int n_inst = 0, n_units = 0, n_meas = 0, n_meas_pad = 0, mb = 64, grid = 0, unroll = 1;
  int grid_scr = 0;               // persistent grid of the screened kernel (0: no such instances)
  uint64_t walker_steps = 0;
  std::vector<void*> allocs;
  void *d_inst = nullptr, *d_seed = nullptr, *d_queue = nullptr, *d_part = nullptr, *d_out = nullptr;
  unsigned long long* h_seed = nullptr;
  hipStream_t st = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool use_graph = true;
  dmw::BatchArgs args{}, args_scr{};   // general kernel, screened kernel
  ~dmw_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (void* p : allocs) (void)hipFree(p);
    if (h_seed) (void)hipHostFree(h_seed);
    for (hipEvent_t e : {ev0, ev1})
      if (e) (void)hipEventDestroy(e);
    if (st) (void)hipStreamDestroy(st);
  }
  hipError_t upload(void** dst, const void* src, size_t bytes) {
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, std::max<size_t>(bytes, 16));
    if (e != hipSuccess) return e;
    allocs.push_back(p);
    *dst = p;
    if (bytes && src) e = hipMemcpy(p, src, bytes, hipMemcpyHostToDevice);
    return e;
  }
};

namespace {

hipError_t enqueue_batch(dmw_batch* B) {
  hipLaunchKernelGGL(dmw::batch_reset_kernel, dim3(1), dim3(1), 0, B->st, (unsigned int*)B->d_queue);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const dim3 blk(dmw::kBlock);
  if (B->grid > 0) {
    const dim3 g(B->grid);
    if (B->mb == 16 && B->unroll == 2)
      hipLaunchKernelGGL((dmw::batch_kernel<16, 2, false>), g, blk, 0, B->st, B->args);
    else if (B->mb == 16)
      hipLaunchKernelGGL((dmw::batch_kernel<16, 1, false>), g, blk, 0, B->st, B->args);
    else if (B->unroll == 2)
      hipLaunchKernelGGL((dmw::batch_kernel<64, 2, false>), g, blk, 0, B->st, B->args);
    else
      hipLaunchKernelGGL((dmw::batch_kernel<64, 1, false>), g, blk, 0, B->st, B->args);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (B->grid_scr > 0) {
    const dim3 g(B->grid_scr);
    if (B->mb == 128 && B->unroll == 2)
      hipLaunchKernelGGL((dmw::batch_kernel<128, 2, true>), g, blk, 0, B->st, B->args_scr);
    else if (B->mb == 128)
      hipLaunchKernelGGL((dmw::batch_kernel<128, 1, true>), g, blk, 0, B->st, B->args_scr);
    else if (B->mb == 16 && B->unroll == 2)
      hipLaunchKernelGGL((dmw::batch_kernel<16, 2, true>), g, blk, 0, B->st, B->args_scr);
    else if (B->mb == 16)
      hipLaunchKernelGGL((dmw::batch_kernel<16, 1, true>), g, blk, 0, B->st, B->args_scr);
    else if (B->unroll == 2)
      hipLaunchKernelGGL((dmw::batch_kernel<64, 2, true>), g, blk, 0, B->st, B->args_scr);
    else
      hipLaunchKernelGGL((dmw::batch_kernel<64, 1, true>), g, blk, 0, B->st, B->args_scr);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(dmw::batch_reduce_kernel, dim3(B->n_inst), dim3(256), 0, B->st,
                     (const dmw::InstDev*)B->d_inst, (const double*)B->d_part, B->n_meas_pad * 2,
                     (double*)B->d_out);
  return hipGetLastError();
}

}  // namespace

extern "C" {

int dmw_batch_create(dmw_batch** out, int32_t n_inst, const dmw_instance* inst, int32_t n_sets,
                     const dmw_waveform* sets, const dmw_batch_tuning* tuning, int32_t use_graph) {
  if (!out || !inst || !sets || n_inst < 1 || n_sets < 1) return dmw::batch_fail(DMW_EINVAL, "null or empty batch");
  *out = nullptr;
  const dmw_batch_tuning tn = tuning ? *tuning : dmw_batch_tuning{0, 0, 0};
  if (tn.waves_per_simd < 0 || tn.waves_per_simd > 8)
    return dmw::batch_fail(DMW_EINVAL, "waves_per_simd must be 0 (automatic) or 1..8");
  const int wpw = tn.walkers_per_wave == 0 ? dmw::kWave : tn.walkers_per_wave;
  if (wpw != 64 && wpw != 128 && wpw != 256 && wpw != 512)
    return dmw::batch_fail(DMW_EINVAL, "walkers_per_wave must be 0 (64), 64, 128, 256 or 512");
  const int unroll = tn.unroll == 0 ? 1 : tn.unroll;
  if (unroll != 1 && unroll != 2) return dmw::batch_fail(DMW_EINVAL, "unroll must be 0 (1), 1 or 2");
  const int tpp = wpw / dmw::kWave;
  const int n_meas = sets[0].n_meas, n_t = sets[0].n_t;
  // more than 64 measurements: only the screened kernel (its phase parked in device
  // memory) takes them, so every instance must be a permeable packing
  bool all_screened = true;
  for (int i = 0; i < n_inst; ++i)
    all_screened = all_screened &&
                   (inst[i].geom.kind == DMW_PACKED_CYLINDERS || inst[i].geom.kind == DMW_PACKED_SPHERES) &&
                   inst[i].geom.wall == DMW_WALL_PERMEATE;
  if (n_meas < 1 || n_t < 1 || n_meas > 128)
    return dmw::batch_fail(DMW_EINVAL, "batched launch needs 1..128 measurements");
  if (n_meas > 64 && !all_screened)
    return dmw::batch_fail(DMW_EINVAL,
                           "batched launch of more than 64 measurements needs every instance to be a permeable packing");
  for (int s = 0; s < n_sets; ++s)
    if (sets[s].n_meas != n_meas || sets[s].n_t != n_t || !sets[s].G)
      return dmw::batch_fail(DMW_EINVAL, "every waveform set needs the same n_meas and n_t");
  const bool have_chi = sets[0].chi != nullptr;
  for (int i = 0; i < n_inst; ++i) {
    const dmw_instance& I = inst[i];
    if (I.geom.kind == DMW_MYELINATED_CYLINDER || I.geom.kind == DMW_PACKED_MYELINATED)
      return dmw::batch_fail(DMW_EUNSUPPORTED, "myelinated geometries run on the single-walk path, not the batched launch");
    if (I.geom.kind < DMW_FREE || I.geom.kind > DMW_PERMEABLE_SHELL_CYL) return dmw::batch_fail(DMW_EINVAL, "unknown geometry kind");
    if (I.geom.wall < 0 || I.geom.wall > 2 || I.n_walkers < 1 || I.n_sub < 1 || I.wave_set < 0 || I.wave_set >= n_sets)
      return dmw::batch_fail(DMW_EINVAL, "bad instance");
    if ((I.geom.kind == DMW_FREE && I.geom.wall != DMW_WALL_REFLECT) ||
        (I.geom.kind == DMW_BOX1D && I.geom.wall == DMW_WALL_PERMEATE) ||
        ((I.geom.kind == DMW_PERMEABLE_SHELL || I.geom.kind == DMW_PERMEABLE_SHELL_CYL) &&
         I.geom.wall == DMW_WALL_SURFACE))
      return dmw::batch_fail(DMW_EUNSUPPORTED, "wall mode not available for this geometry");
    if (const char* why = dmw::outside_float_envelope(I.geom, I.step_len)) return dmw::batch_fail(DMW_EINVAL, why);
    const bool packed = I.geom.kind == DMW_PACKED_CYLINDERS || I.geom.kind == DMW_PACKED_SPHERES;
    if (packed && (I.geom.n_obj < 1 || !I.geom.centers || !I.geom.radii || !(I.geom.cell > 0.0f)))
      return dmw::batch_fail(DMW_EINVAL, "packed instance needs objects and a positive cell side");
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return dmw::batch_fail(DMW_ENODEV, "no HIP device");
  dmw_batch* B = new dmw_batch();
  B->n_inst = n_inst;
  B->n_meas = n_meas;
  B->mb = n_meas <= 16 ? 16 : (n_meas <= 64 ? 64 : 128);
  B->unroll = unroll;
  B->n_meas_pad = B->mb;
  B->use_graph = use_graph != 0;
  const int Mp = B->n_meas_pad;

  std::vector<dmw::InstDev> hi(n_inst);
  std::vector<int2> units;
  std::vector<double> cost;
  std::vector<char> screened_unit;   // the entry's instance walks in the screened kernel
  std::vector<float4> objs;
  std::vector<int> starts, items;
  int n_tiles = 0;
  for (int i = 0; i < n_inst; ++i) {
    const dmw_instance& I = inst[i];
    dmw::InstDev& d = hi[i];
    d.size = I.geom.size; d.eps = I.geom.eps; d.nudge = I.geom.nudge;
    d.inv_size = I.geom.size > 0.0f ? dmw::recip_rn(I.geom.size) : 0.0f;
    d.size_sq_cut = I.geom.size > 0.0f ? dmw::sqrt_cut(I.geom.size) : 0.0f;
    d.size_sq_cut_gt = I.geom.size > 0.0f ? dmw::sqrt_cut(std::nextafter(I.geom.size, INFINITY)) : 0.0f;
    d.rhoD = I.geom.rho_over_D; d.kappaD = I.geom.kappa_over_D;
    d.step_len = I.step_len; d.dt_sub = I.dt_sub; d.inv_T2 = I.inv_T2; d.gdt = I.gamma_dt_sub;
    d.kind = I.geom.kind;
    d.ax = I.geom.axes[0]; d.ay = I.geom.axes[1]; d.az = I.geom.axes[2]; d.aux = I.geom.aux;
    d.wall = I.geom.wall; d.n_sub = I.n_sub; d.has_t2 = I.has_t2;
    d.n_walkers = I.n_walkers;
    d.seed_offset = I.seed_offset;
    d.stream = I.stream < 0 ? (unsigned int)i : (unsigned int)I.stream;
    d.pad_ = 0;
    d.cell = I.geom.cell;
    d.inv_cell = I.geom.cell > 0.0f ? 1.0f / I.geom.cell : 0.0f;
    d.g_off = (long long)I.wave_set * n_t * Mp * 3;
    d.gh_off = (long long)I.wave_set * n_t * 2 * ((Mp + 31) / 32 * 32);  // phase_half_table's size per set
    d.gsum_off = (long long)I.wave_set * 3 * Mp;
    d.p_scale = (float)dmw::phase_p_scale(n_t, I.n_sub, I.step_len, I.gamma_dt_sub);
    d.grid_n = 1; d.grid_inv_h = 0.0f; d.grid_h = 0.0f; d.exact_minimage = 0; d.n_obj = 0;
    d.scr_cache = 0; d.scr_w = d.scr_m2 = d.scr_cout = d.scr_cin = 0.0f;
    d.obj_off = d.start_off = d.items_off = 0;
    if (I.geom.kind == DMW_PACKED_CYLINDERS || I.geom.kind == DMW_PACKED_SPHERES) {
      const int dim = I.geom.kind == DMW_PACKED_CYLINDERS ? 2 : 3;
      // permeable packings: cells expanded by half a side for the screen, less where
      // a list would exceed its object cache (the box level rides in the list start)
      const bool screened = I.geom.wall == DMW_WALL_PERMEATE;
      Grid G = build_grid(I.geom, I.step_len, dim, 60 * 1024, 16, screened ? 0.5 : 0.0,
                          screened ? dmw::SCR_MAXC : 0);
      d.obj_off = (int)objs.size();
      d.start_off = (int)starts.size();
      d.items_off = (int)items.size();
      for (int k = 0; k < I.geom.n_obj; ++k) {
        float4 o = make_float4(0.0f, 0.0f, 0.0f, I.geom.radii[k]);
        o.x = I.geom.centers[k * dim];
        o.y = I.geom.centers[k * dim + 1];
        if (dim == 3) o.z = I.geom.centers[k * dim + 2];
        objs.push_back(o);
      }
      const std::vector<int> enc = G.encoded_start();
      starts.insert(starts.end(), enc.begin(), enc.end());
      items.insert(items.end(), G.items.begin(), G.items.end());
      d.grid_n = G.n;
      d.grid_inv_h = 1.0f / G.h;
      d.grid_h = G.h;
      d.exact_minimage = G.exact_minimage ? 1 : 0;
      d.n_obj = I.geom.n_obj;
      if (screened) {
        double rmax = 0.0;
        for (int k = 0; k < I.geom.n_obj; ++k) rmax = std::max(rmax, (double)I.geom.radii[k]);
        const double m = dmw::screen_margin(I.geom.cell, rmax, I.step_len);
        int max_len = 0;
        for (size_t c = 0; c + 1 < G.start.size(); ++c) max_len = std::max(max_len, G.start[c + 1] - G.start[c]);
        d.scr_cache = (G.exact_minimage || max_len > 32) ? 0 : 1;
        d.scr_w = (float)(0.5 * G.h);
        d.scr_m2 = (float)(2.0 * m);
        d.scr_cin = (float)(m * m);
        d.scr_cout = (float)(m * m + 0.25 * (double)I.step_len * (double)I.step_len * (1.0 + 0x1p-20));
      }
    }
    d.unit_begin = n_tiles;
    const int tiles = (int)((I.n_walkers + dmw::kWave - 1) / dmw::kWave);
    const double packed_cost = (I.geom.kind >= DMW_PACKED_CYLINDERS) ? 2.0 : 1.0;
    const bool scr_inst = (I.geom.kind == DMW_PACKED_CYLINDERS || I.geom.kind == DMW_PACKED_SPHERES) &&
                          I.geom.wall == DMW_WALL_PERMEATE;
    for (int j = 0; j < tiles; j += tpp) {
      units.push_back(make_int2(i, j));
      screened_unit.push_back(scr_inst ? 1 : 0);
      cost.push_back((double)std::min(tpp, tiles - j) * I.n_sub * packed_cost *
                     (I.geom.wall == DMW_WALL_PERMEATE ? 1.3 : 1.0));
    }
    n_tiles += tiles;
    d.unit_end = n_tiles;
    B->walker_steps += (uint64_t)I.n_walkers * (uint64_t)n_t * (uint64_t)I.n_sub;
  }
  B->n_units = n_tiles;
  const int n_entries = (int)units.size();
  std::vector<int> order(n_entries);
  std::iota(order.begin(), order.end(), 0);
  // two queues, each in decreasing cost: the general kernel's entries, then the
  // screened kernel's
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    if (screened_unit[x] != screened_unit[y]) return screened_unit[x] < screened_unit[y];
    return cost[x] > cost[y];
  });
  std::vector<int2> q_units(n_entries);
  for (int k = 0; k < n_entries; ++k) q_units[k] = units[order[k]];
  int n_scr = 0;
  for (char c : screened_unit) n_scr += c;
  const int n_gen = n_entries - n_scr;

  std::vector<float> G((size_t)n_sets * n_t * Mp * 3, 0.0f);
  for (int s = 0; s < n_sets; ++s)
    for (int m = 0; m < n_meas; ++m)
      for (int t = 0; t < n_t; ++t)
        for (int c = 0; c < 3; ++c)
          G[(((size_t)s * n_t + t) * Mp + m) * 3 + c] = sets[s].G[((size_t)m * n_t + t) * 3 + c];
  // float16 hi/lo phase operands and column sums per waveform set
  std::vector<uint2> GH;
  std::vector<float> Gsum;
  std::vector<double> g_scale(n_sets);
  for (int s = 0; s < n_sets; ++s) {
    const float* gs = G.data() + (size_t)s * n_t * Mp * 3;
    g_scale[s] = dmw::phase_g_scale(gs, (size_t)n_t * Mp * 3);
    std::vector<uint2> gh;
    std::vector<float> sums;
    dmw::phase_half_table(gs, n_t, Mp, g_scale[s], gh);
    dmw::phase_column_sums(gs, n_t, Mp, nullptr, 0, sums);
    GH.insert(GH.end(), gh.begin(), gh.end());
    Gsum.insert(Gsum.end(), sums.begin(), sums.end());
  }
  for (int i = 0; i < n_inst; ++i)
    hi[i].inv_pg_scale = (float)(1.0 / ((double)hi[i].p_scale * g_scale[inst[i].wave_set]));

  int n_cu = 256;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) == hipSuccess) n_cu = prop.multiProcessorCount;
  const int wps = tn.waves_per_simd > 0 ? tn.waves_per_simd : 4;  // the kernel's occupancy budget
  B->grid = n_gen > 0 ? std::max(1, std::min(n_cu * wps, (n_gen + 3) / 4)) : 0;
  // (the launch shape's wavefronts per SIMD size both persistent grids)
  B->grid_scr = n_scr > 0 ? std::max(1, std::min(n_cu * wps, (n_scr + 3) / 4)) : 0;

  auto fail_free = [&](hipError_t err, const char* what) {
    delete B;
    return dmw::batch_hip_fail(err, what);
  };
  void *d_units, *d_G, *d_GH, *d_Gsum, *d_chi = nullptr, *d_objs, *d_starts, *d_items;
#define BT(x, what)                                       \
  do {                                                    \
    hipError_t err_ = (x);                                \
    if (err_ != hipSuccess) return fail_free(err_, what); \
  } while (0)
  BT(hipStreamCreateWithFlags(&B->st, hipStreamNonBlocking), "hipStreamCreate");
  BT(B->upload(&B->d_inst, hi.data(), hi.size() * sizeof(dmw::InstDev)), "batch upload");
  BT(B->upload(&d_units, q_units.data(), q_units.size() * sizeof(int2)), "batch upload");
  BT(B->upload(&d_G, G.data(), G.size() * 4), "batch upload");
  BT(B->upload(&d_GH, GH.data(), GH.size() * sizeof(uint2)), "batch upload");
  BT(B->upload(&d_Gsum, Gsum.data(), Gsum.size() * 4), "batch upload");
  if (have_chi) BT(B->upload(&d_chi, sets[0].chi, (size_t)n_t * 4), "batch upload");
  BT(B->upload(&d_objs, objs.data(), objs.size() * sizeof(float4)), "batch upload");
  BT(B->upload(&d_starts, starts.data(), starts.size() * 4), "batch upload");
  BT(B->upload(&d_items, items.data(), items.size() * 4), "batch upload");
  BT(B->upload(&B->d_seed, nullptr, 8), "batch alloc");
  BT(B->upload(&B->d_queue, nullptr, 8), "batch alloc");
  BT(B->upload(&B->d_part, nullptr, (size_t)B->n_units * Mp * 2 * 8), "batch alloc");
  // the parked phase writes partials only up to n_meas rounded to 16: the padding
  // columns stay the zeros written here
  BT(hipMemsetAsync(B->d_part, 0, (size_t)B->n_units * Mp * 2 * 8, B->st), "hipMemset");
  BT(hipStreamSynchronize(B->st), "hipMemset");
  void* d_park = nullptr;
  if (B->grid_scr > 0)
    BT(B->upload(&d_park, nullptr, (size_t)B->grid_scr * (dmw::kBlock / dmw::kWave) * Mp * 32 * sizeof(float4)),
       "batch alloc");
  BT(B->upload(&B->d_out, nullptr, (size_t)n_inst * Mp * 2 * 8), "batch alloc");
  BT(hipHostMalloc((void**)&B->h_seed, 8, hipHostMallocDefault), "hipHostMalloc");
  BT(hipEventCreate(&B->ev0), "hipEventCreate");
  BT(hipEventCreate(&B->ev1), "hipEventCreate");

  dmw::BatchArgs& a = B->args;
  a.inst = (const dmw::InstDev*)B->d_inst;
  a.units = (const int2*)d_units;
  a.n_units = n_gen;
  a.tiles_per_pop = tpp;
  a.G = (const float*)d_G;
  a.GH = (const uint2*)d_GH;
  a.Gsum = (const float*)d_Gsum;
  a.chi = have_chi ? (const float*)d_chi : nullptr;
  a.n_t = n_t;
  a.n_meas = n_meas;
  a.n_meas_pad = Mp;
  a.objs = (const float4*)d_objs;
  a.starts = (const int*)d_starts;
  a.items = (const int*)d_items;
  a.seed = (const unsigned long long*)B->d_seed;
  a.queue = (unsigned int*)B->d_queue;
  a.partials = (double*)B->d_part;
  B->args_scr = a;
  B->args_scr.park = (float4*)d_park;
  B->args_scr.units = (const int2*)d_units + n_gen;
  B->args_scr.n_units = n_scr;
  B->args_scr.queue = (unsigned int*)B->d_queue + 1;

  if (B->use_graph) {
    BT(hipStreamBeginCapture(B->st, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    hipError_t ce = enqueue_batch(B);
    hipGraph_t g = nullptr;
    hipError_t ee = hipStreamEndCapture(B->st, &g);
    B->graph = g;
    if (ce != hipSuccess) return fail_free(ce, "batch capture");
    BT(ee, "hipStreamEndCapture");
    BT(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
  }
#undef BT
  *out = B;
  return DMW_OK;
}

int dmw_batch_execute(dmw_batch* B, uint64_t seed, double* signals, dmw_batch_stats* stats) {
  if (!B || !signals) return dmw::batch_fail(DMW_EINVAL, "null argument");
  *B->h_seed = seed;
  hipError_t e = hipMemcpyAsync(B->d_seed, B->h_seed, 8, hipMemcpyHostToDevice, B->st);
  if (e != hipSuccess) return dmw::batch_hip_fail(e, "seed upload");
  e = hipEventRecord(B->ev0, B->st);
  if (e != hipSuccess) return dmw::batch_hip_fail(e, "hipEventRecord");
  e = B->use_graph ? hipGraphLaunch(B->exec, B->st) : enqueue_batch(B);
  if (e != hipSuccess) return dmw::batch_hip_fail(e, "batch launch");
  e = hipEventRecord(B->ev1, B->st);
  if (e != hipSuccess) return dmw::batch_hip_fail(e, "hipEventRecord");
  std::vector<double> h((size_t)B->n_inst * B->n_meas_pad * 2);
  e = hipMemcpyAsync(h.data(), B->d_out, h.size() * 8, hipMemcpyDeviceToHost, B->st);
  if (e != hipSuccess) return dmw::batch_hip_fail(e, "D2H");
  e = hipStreamSynchronize(B->st);
  if (e != hipSuccess) return dmw::batch_hip_fail(e, "batch walk");
  for (double v : h)
    if (std::isnan(v))
      return dmw::batch_fail(DMW_EHIP, "batch produced a NaN signal (non-finite input, or a phase operand beyond the float16 range)");
  for (int i = 0; i < B->n_inst; ++i)
    for (int m = 0; m < B->n_meas; ++m) signals[(size_t)i * B->n_meas + m] = h[((size_t)i * B->n_meas_pad + m) * 2];
  if (stats) {
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, B->ev0, B->ev1);
    stats->total_ms = ms;
    stats->kernel_ms = ms;
    stats->walker_steps = B->walker_steps;
    stats->n_units = B->n_units;
    stats->grid = std::max(B->grid, B->grid_scr);
  }
  return DMW_OK;
}

int dmw_batch_destroy(dmw_batch* B) {
  delete B;
  return DMW_OK;
}

}  // extern "C"
