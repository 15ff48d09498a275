// batch.hip -- one persistent launch walking thousands of substrate instances.
//
// The reference runs one simulate() call per substrate (core.py:193), e.g. the
// serial loop of simulate_mixture (core.py:821-858).  Here every instance
// (Cylinder or Sphere, reflecting / relaxing / permeable wall, its own D, radius,
// rho, kappa, walker count and sub-step count) is cut into 64-walker work units;
// a persistent grid of wavefronts pops units from a queue ordered by decreasing
// cost (walkers x n_t x n_sub), so the long permeable instances start first and
// the short ones fill the tail.  Each unit writes float64 partial sums; a second
// kernel reduces every instance's units in a fixed order.  The three launches
// (queue reset, walk, reduce) are captured once in a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/dmipy_walk.h"
#include "batch.h"
#include "walk_kernel.h"

namespace dmw {

struct InstDev {
  float size, eps, nudge, rhoD, kappaD, step_len, dt_sub, inv_T2, gdt;
  int kind, wall, n_sub, has_t2;
  long long n_walkers;
  int unit_begin, unit_end;
};

struct BatchArgs {
  const InstDev* inst;
  const int2* units;        // (instance, tile) in queue order
  int n_units;
  const float* G;           // (n_t, n_meas_pad, 3) T/m, not scaled
  const float* GT;          // (n_t, 4, n_meas_pad) component-major, row 3 = 0
  const float* chi;
  int n_t, n_meas, n_meas_pad;
  const unsigned long long* seed;  // device, read at kernel start
  unsigned int* queue;      // work counter
  double* partials;         // (n_units, n_meas_pad, 2) indexed by unit id
  const int* unit_id;       // queue position -> unit id
};

template <int GEOM, int WALL, int MB>
__device__ __forceinline__ void walk_unit(const BatchArgs& b, const InstDev& I, int inst, int tile,
                                          int unit, uint32_t seed_lo, uint32_t seed_hi) {
  const int lane = threadIdx.x & (kWave - 1);
  const long long widx = (long long)tile * kWave + lane;
  const bool active = widx < I.n_walkers;
  const uint32_t w = (uint32_t)(active ? widx : I.n_walkers - 1);
  WalkArgs a{};
  a.size = I.size; a.eps = I.eps; a.nudge = I.nudge; a.rhoD = I.rhoD; a.kappaD = I.kappaD;
  a.instance = (uint32_t)inst; a.seed_lo = seed_lo; a.seed_hi = seed_hi;
  Tables tb{nullptr, nullptr, nullptr};
  a.GsT = b.GT;
  a.Gs = b.G;
  a.n_meas_pad = b.n_meas_pad;
  Pos p = sample_start<GEOM>(w, a, tb);
  PhaseAcc<MB, false> ph;
  ph.init();
  float logw = 0.0f;
  uint32_t hits = 0, ncross = 0;
  const int n_sub = I.n_sub;
  const float fsub = (float)n_sub;
  u32x4 rnd{0u, 0u, 0u, 0u};
  for (int t = 0; t < b.n_t; ++t) {
    ph.prefetch(a, t, 0, lane);
    const float chi = b.chi ? b.chi[t] : 1.0f;
    const float x0 = p.x, y0 = p.y, z0 = p.z;
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    for (int s = 0; s < n_sub; ++s) {
      const uint32_t gs = (uint32_t)t * (uint32_t)n_sub + (uint32_t)s;
      if (!(gs & 1u) || (t == 0 && s == 0))
        rnd = philox4x32_10({w, (uint32_t)inst, gs >> 1, TAG_STEP}, seed_lo, seed_hi);
      float ux, uy, uz;
      iso_direction((gs & 1u) ? rnd.z : rnd.x, (gs & 1u) ? rnd.w : rnd.y, &ux, &uy, &uz);
      const float sx = ux * I.step_len, sy = uy * I.step_len, sz = uz * I.step_len;
      const PowlesSrc ps{w, (uint32_t)inst, gs, seed_lo, seed_hi, 0.0f, false};
      float dlog = 0.0f;
      if (GEOM == G_CYL) {
        if (WALL == W_PERMEATE) {
          dlog = cyl_permeate(p.x, p.y, sx, sy, a, ps, hits, ncross);
        } else {
          const float ds = cyl_reflect(p.x, p.y, sx, sy, a, hits);
          if (WALL == W_SURFACE) dlog = (-2.0f * a.rhoD) * ds;
        }
        p.z = p.z + sz;
      } else {
        if (WALL == W_PERMEATE) {
          dlog = sph_permeate(p, sx, sy, sz, a, ps, hits, ncross);
        } else {
          const float ds = sph_reflect(p, sx, sy, sz, a, hits);
          if (WALL == W_SURFACE) dlog = (-2.0f * a.rhoD) * ds;
        }
      }
      if (WALL != W_REFLECT) dlog = dlog * chi;
      if (I.has_t2) dlog = dlog - (I.dt_sub * I.inv_T2) * chi;
      logw = logw + dlog;
      if (n_sub > 1) {
        ax = ax + (p.x - x0);
        ay = ay + (p.y - y0);
        az = az + (p.z - z0);
      }
    }
    float Px, Py, Pz;
    if (n_sub > 1) {
      Px = fmaf(fsub, x0, ax); Py = fmaf(fsub, y0, ay); Pz = fmaf(fsub, z0, az);
    } else {
      Px = p.x; Py = p.y; Pz = p.z;
    }
    ph.update(a, t, 0, lane, Px * I.gdt, Py * I.gdt, Pz * I.gdt);
  }
  const float wgt = expf(logw);
  ph.emit(wgt, active, lane, b.partials + (size_t)unit * b.n_meas_pad * 2);
  (void)hits; (void)ncross;
}

template <int MB>
__global__ __launch_bounds__(kBlock) void batch_kernel(const BatchArgs b) {
  const unsigned long long seed = *b.seed;
  const uint32_t seed_lo = (uint32_t)seed, seed_hi = (uint32_t)(seed >> 32);
  for (;;) {
    int pos = 0;
    if ((threadIdx.x & (kWave - 1)) == 0) pos = (int)atomicAdd(b.queue, 1u);
    pos = __shfl(pos, 0, kWave);
    if (pos >= b.n_units) break;  // wave-uniform: every wave reaches this exit
    const int2 u = b.units[pos];
    const int unit = b.unit_id[pos];
    const InstDev I = b.inst[u.x];
    const int key = I.kind * 3 + I.wall;
    switch (key) {
      case G_CYL * 3 + W_REFLECT: walk_unit<G_CYL, W_REFLECT, MB>(b, I, u.x, u.y, unit, seed_lo, seed_hi); break;
      case G_CYL * 3 + W_SURFACE: walk_unit<G_CYL, W_SURFACE, MB>(b, I, u.x, u.y, unit, seed_lo, seed_hi); break;
      case G_CYL * 3 + W_PERMEATE: walk_unit<G_CYL, W_PERMEATE, MB>(b, I, u.x, u.y, unit, seed_lo, seed_hi); break;
      case G_SPH * 3 + W_REFLECT: walk_unit<G_SPH, W_REFLECT, MB>(b, I, u.x, u.y, unit, seed_lo, seed_hi); break;
      case G_SPH * 3 + W_SURFACE: walk_unit<G_SPH, W_SURFACE, MB>(b, I, u.x, u.y, unit, seed_lo, seed_hi); break;
      default: walk_unit<G_SPH, W_PERMEATE, MB>(b, I, u.x, u.y, unit, seed_lo, seed_hi); break;
    }
  }
}

__global__ void batch_reset_kernel(unsigned int* queue) { *queue = 0u; }

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
  int n_inst = 0, n_units = 0, n_meas = 0, n_meas_pad = 0, mb = 64, grid = 0;
  uint64_t walker_steps = 0;
  void *d_inst = nullptr, *d_units = nullptr, *d_uid = nullptr, *d_G = nullptr, *d_GT = nullptr,
       *d_chi = nullptr;
  void *d_seed = nullptr, *d_queue = nullptr, *d_part = nullptr, *d_out = nullptr;
  unsigned long long* h_seed = nullptr;
  hipStream_t st = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
  bool use_graph = true;
  dmw::BatchArgs args{};
  ~dmw_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (void* p : {d_inst, d_units, d_uid, d_G, d_GT, d_chi, d_seed, d_queue, d_part, d_out})
      if (p) (void)hipFree(p);
    if (h_seed) (void)hipHostFree(h_seed);
    for (hipEvent_t e : {ev0, ev1, ev2})
      if (e) (void)hipEventDestroy(e);
    if (st) (void)hipStreamDestroy(st);
  }
};

namespace {

hipError_t enqueue_batch(dmw_batch* B) {
  hipLaunchKernelGGL(dmw::batch_reset_kernel, dim3(1), dim3(1), 0, B->st, (unsigned int*)B->d_queue);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (B->mb == 16)
    hipLaunchKernelGGL((dmw::batch_kernel<16>), dim3(B->grid), dim3(dmw::kBlock), 0, B->st, B->args);
  else
    hipLaunchKernelGGL((dmw::batch_kernel<64>), dim3(B->grid), dim3(dmw::kBlock), 0, B->st, B->args);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(dmw::batch_reduce_kernel, dim3(B->n_inst), dim3(256), 0, B->st,
                     (const dmw::InstDev*)B->d_inst, (const double*)B->d_part, B->n_meas_pad * 2,
                     (double*)B->d_out);
  return hipGetLastError();
}

}  // namespace

extern "C" {

int dmw_batch_create(dmw_batch** out, int32_t n_inst, const dmw_instance* inst,
                     const dmw_waveform* wave, int32_t waves_per_simd, int32_t use_graph) {
  if (!out || !inst || !wave || n_inst < 1) return dmw::batch_fail(DMW_EINVAL, "null or empty batch");
  *out = nullptr;
  if (wave->n_meas < 1 || wave->n_meas > 64 || wave->n_t < 1 || !wave->G)
    return dmw::batch_fail(DMW_EINVAL, "batched launch needs 1..64 measurements");
  for (int i = 0; i < n_inst; ++i) {
    const dmw_instance& I = inst[i];
    if (I.geom.kind != DMW_CYLINDER && I.geom.kind != DMW_SPHERE)
      return dmw::batch_fail(DMW_EUNSUPPORTED, "batched instances must be Cylinder or Sphere");
    if (I.geom.wall < 0 || I.geom.wall > 2 || I.n_walkers < 1 || I.n_sub < 1)
      return dmw::batch_fail(DMW_EINVAL, "bad instance");
  }
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return dmw::batch_fail(DMW_ENODEV, "no HIP device");
  dmw_batch* B = new dmw_batch();
  B->n_inst = n_inst;
  B->n_meas = wave->n_meas;
  B->mb = wave->n_meas <= 16 ? 16 : 64;
  B->n_meas_pad = B->mb;
  B->use_graph = use_graph != 0;

  std::vector<dmw::InstDev> hi(n_inst);
  std::vector<int2> units;
  std::vector<double> cost;
  for (int i = 0; i < n_inst; ++i) {
    const dmw_instance& I = inst[i];
    dmw::InstDev& d = hi[i];
    d.size = I.geom.size; d.eps = I.geom.eps; d.nudge = I.geom.nudge;
    d.rhoD = I.geom.rho_over_D; d.kappaD = I.geom.kappa_over_D;
    d.step_len = I.step_len; d.dt_sub = I.dt_sub; d.inv_T2 = I.inv_T2; d.gdt = I.gamma_dt_sub;
    d.kind = I.geom.kind == DMW_CYLINDER ? dmw::G_CYL : dmw::G_SPH;
    d.wall = I.geom.wall; d.n_sub = I.n_sub; d.has_t2 = I.has_t2;
    d.n_walkers = I.n_walkers;
    d.unit_begin = (int)units.size();
    const int tiles = (int)((I.n_walkers + dmw::kWave - 1) / dmw::kWave);
    for (int j = 0; j < tiles; ++j) {
      units.push_back(make_int2(i, j));
      const long long nw = std::min<long long>(dmw::kWave, I.n_walkers - (long long)j * dmw::kWave);
      // Lanes of a partial tile still run, so the cost is the full tile.
      (void)nw;
      cost.push_back((double)I.n_sub * (1.0 + (I.geom.wall == DMW_WALL_PERMEATE ? 0.3 : 0.0)));
    }
    d.unit_end = (int)units.size();
    B->walker_steps += (uint64_t)I.n_walkers * (uint64_t)wave->n_t * (uint64_t)I.n_sub;
  }
  B->n_units = (int)units.size();
  std::vector<int> order(B->n_units);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost[x] > cost[y]; });
  std::vector<int2> q_units(B->n_units);
  for (int k = 0; k < B->n_units; ++k) q_units[k] = units[order[k]];

  std::vector<float> G((size_t)wave->n_t * B->n_meas_pad * 3, 0.0f);
  for (int m = 0; m < wave->n_meas; ++m)
    for (int t = 0; t < wave->n_t; ++t)
      for (int c = 0; c < 3; ++c)
        G[((size_t)t * B->n_meas_pad + m) * 3 + c] = wave->G[((size_t)m * wave->n_t + t) * 3 + c];

  int n_cu = 256;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) == hipSuccess) n_cu = prop.multiProcessorCount;
  const int wps = waves_per_simd > 0 ? waves_per_simd : 2;
  B->grid = std::max(1, std::min(n_cu * wps, (B->n_units + 3) / 4));

  auto fail_free = [&](hipError_t err, const char* what) {
    delete B;
    return dmw::batch_hip_fail(err, what);
  };
#define BT(x, what)                              \
  do {                                           \
    hipError_t err_ = (x);                       \
    if (err_ != hipSuccess) return fail_free(err_, what); \
  } while (0)
  BT(hipStreamCreateWithFlags(&B->st, hipStreamNonBlocking), "hipStreamCreate");
  BT(hipMalloc(&B->d_inst, hi.size() * sizeof(dmw::InstDev)), "hipMalloc");
  BT(hipMalloc(&B->d_units, q_units.size() * sizeof(int2)), "hipMalloc");
  BT(hipMalloc(&B->d_uid, order.size() * sizeof(int)), "hipMalloc");
  BT(hipMalloc(&B->d_G, G.size() * 4), "hipMalloc");
  std::vector<float> GT((size_t)wave->n_t * 4 * B->n_meas_pad, 0.0f);
  for (int t = 0; t < wave->n_t; ++t)
    for (int c = 0; c < 3; ++c)
      for (int m = 0; m < B->n_meas_pad; ++m)
        GT[((size_t)t * 4 + c) * B->n_meas_pad + m] = G[((size_t)t * B->n_meas_pad + m) * 3 + c];
  BT(hipMalloc(&B->d_GT, GT.size() * 4), "hipMalloc");
  BT(hipMemcpy(B->d_GT, GT.data(), GT.size() * 4, hipMemcpyHostToDevice), "H2D");
  if (wave->chi) BT(hipMalloc(&B->d_chi, (size_t)wave->n_t * 4), "hipMalloc");
  BT(hipMalloc(&B->d_seed, 8), "hipMalloc");
  BT(hipMalloc(&B->d_queue, 4), "hipMalloc");
  BT(hipMalloc(&B->d_part, (size_t)B->n_units * B->n_meas_pad * 2 * 8), "hipMalloc");
  BT(hipMalloc(&B->d_out, (size_t)n_inst * B->n_meas_pad * 2 * 8), "hipMalloc");
  BT(hipHostMalloc((void**)&B->h_seed, 8, hipHostMallocDefault), "hipHostMalloc");
  BT(hipMemcpy(B->d_inst, hi.data(), hi.size() * sizeof(dmw::InstDev), hipMemcpyHostToDevice), "H2D");
  BT(hipMemcpy(B->d_units, q_units.data(), q_units.size() * sizeof(int2), hipMemcpyHostToDevice), "H2D");
  BT(hipMemcpy(B->d_uid, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice), "H2D");
  BT(hipMemcpy(B->d_G, G.data(), G.size() * 4, hipMemcpyHostToDevice), "H2D");
  if (wave->chi) BT(hipMemcpy(B->d_chi, wave->chi, (size_t)wave->n_t * 4, hipMemcpyHostToDevice), "H2D");
  BT(hipEventCreate(&B->ev0), "hipEventCreate");
  BT(hipEventCreate(&B->ev1), "hipEventCreate");
  BT(hipEventCreate(&B->ev2), "hipEventCreate");

  dmw::BatchArgs& a = B->args;
  a.inst = (const dmw::InstDev*)B->d_inst;
  a.units = (const int2*)B->d_units;
  a.unit_id = (const int*)B->d_uid;
  a.n_units = B->n_units;
  a.G = (const float*)B->d_G;
  a.GT = (const float*)B->d_GT;
  a.chi = wave->chi ? (const float*)B->d_chi : nullptr;
  a.n_t = wave->n_t;
  a.n_meas = wave->n_meas;
  a.n_meas_pad = B->n_meas_pad;
  a.seed = (const unsigned long long*)B->d_seed;
  a.queue = (unsigned int*)B->d_queue;
  a.partials = (double*)B->d_part;

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
  for (int i = 0; i < B->n_inst; ++i)
    for (int m = 0; m < B->n_meas; ++m) signals[(size_t)i * B->n_meas + m] = h[((size_t)i * B->n_meas_pad + m) * 2];
  if (stats) {
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, B->ev0, B->ev1);
    stats->total_ms = ms;
    stats->kernel_ms = ms;
    stats->walker_steps = B->walker_steps;
    stats->n_units = B->n_units;
    stats->grid = B->grid;
  }
  return DMW_OK;
}

int dmw_batch_destroy(dmw_batch* B) {
  delete B;
  return DMW_OK;
}

}  // extern "C"
