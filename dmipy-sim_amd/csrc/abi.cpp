// abi.cpp -- host side of the C ABI declared in include/dmipy_walk.h.
//
// Uploads a walk (waveform, geometry tables, optional start positions and
// injected increments), launches the walk kernel over walker tiles and
// measurement blocks, reduces the per-wave partial sums in float64 and copies
// the requested outputs back.  One HIP stream per calling thread and device; no
// other state is kept between calls except inside a dmw_plan.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/dmipy_walk.h"
#include "envelope.h"
#include "walk_kernel.h"
#include "walk_launch.h"
#include "batch.h"
#include "grid.h"
#include "phase16.h"
#include "traj.h"

namespace {

thread_local std::string g_err;
thread_local int g_device = 0;

// One non-blocking stream per (thread, device), created on first use and kept
// for the life of the thread (at most one per device: switching devices back and
// forth reuses them).  They are not destroyed at thread exit, where the HIP
// runtime may already be shutting down.
thread_local std::vector<hipStream_t> g_streams;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) {
    m += " (out of memory)";
    return fail(DMW_ENOMEM, m);
  }
  if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return fail(DMW_ENODEV, m);
  return fail(DMW_EHIP, m);
}

#define HIP_TRY(expr, what)                        \
  do {                                             \
    hipError_t e_ = (expr);                        \
    if (e_ != hipSuccess) return hip_fail(e_, what); \
  } while (0)

// Makes `device` current on this thread and returns its stream.
int ensure_stream_on(int device, hipStream_t* out) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(DMW_ENODEV, "no HIP device visible");
  if (device < 0 || device >= n) return fail(DMW_ENODEV, "device index out of range");
  HIP_TRY(hipSetDevice(device), "hipSetDevice");
  std::vector<hipStream_t>& v = g_streams;
  if ((int)v.size() < n) v.resize(n, nullptr);
  if (v[device] == nullptr) HIP_TRY(hipStreamCreateWithFlags(&v[device], hipStreamNonBlocking), "hipStreamCreate");
  *out = v[device];
  return DMW_OK;
}

int ensure_stream(hipStream_t* out) { return ensure_stream_on(g_device, out); }

// ---------------------------------------------------------------------------
// Allocation-free walks (dmw_walk_noalloc): every device buffer of the call is
// carved from `scratch`, a static device array of kScratchBytes per GPU, instead
// of hipMalloc.  Host code never names the array itself (that is a host-side
// stand-in, not a device address): its device address comes from
// hipGetSymbolAddress, once per device, with that device current.
// ---------------------------------------------------------------------------
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::once_flag g_scratch_once[kMaxDevices];
void* g_scratch_ptr[kMaxDevices];
hipError_t g_scratch_err[kMaxDevices];
std::mutex g_scratch_mutex[kMaxDevices];

// Bump allocator over one call's share of the scratch array.
struct Arena {
  char* base;
  size_t cap, used;
};
thread_local Arena* g_arena = nullptr;

// Device allocation owned by a plan (from the arena while one is installed).
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  bool in_arena = false;
  int alloc(size_t n) {
    if (n <= bytes && p) return DMW_OK;
    release();
    if (n == 0) return DMW_OK;
    if (g_arena) {
      const size_t off = (g_arena->used + 255) & ~(size_t)255;
      if (off > g_arena->cap || n > g_arena->cap - off)
        return fail(DMW_EINVAL, "dmw_walk_noalloc: the walk's device buffers need more than the " +
                                    std::to_string(kScratchBytes >> 20) +
                                    " MiB scratch array (fewer walkers, measurements or per-walker "
                                    "outputs, or dmw_walk)");
      p = g_arena->base + off;
      g_arena->used = off + n;
      bytes = n;
      in_arena = true;
      return DMW_OK;
    }
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
      p = nullptr;
      return hip_fail(e, "hipMalloc");
    }
    bytes = n;
    return DMW_OK;
  }
  void release() {
    if (p && !in_arena) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    in_arena = false;
  }
  template <class T>
  T* as() const { return reinterpret_cast<T*>(p); }
};

// Copies n bytes from host memory and waits for the copy: callers pass
// temporaries that are freed on return.
int upload(DevBuf& b, const void* src, size_t n, hipStream_t st) {
  int rc = b.alloc(n);
  if (rc) return rc;
  if (n) {
    HIP_TRY(hipMemcpyAsync(b.p, src, n, hipMemcpyHostToDevice, st), "hipMemcpyAsync H2D");
    HIP_TRY(hipStreamSynchronize(st), "hipMemcpyAsync H2D");
  }
  return DMW_OK;
}

}  // namespace

namespace dmw {
int batch_fail(int code, const char* msg) { return fail(code, msg); }
int batch_hip_fail(hipError_t e, const char* what) { return hip_fail(e, what); }
}  // namespace dmw

// ---------------------------------------------------------------------------
struct dmw_plan {
  int device = 0;            // device holding the plan's buffers (current at dmw_plan_create)
  dmw_geometry geom{};
  dmw_myelin myelin{};
  double inv_norm = 1.0;     // 1 / (sum of spin weights): 1/N except for the myelinated cylinder
  dmw_run run{};
  int n_meas = 0, n_t = 0, n_echo = 0;
  int mb = 4, n_chunks = 1, n_meas_pad = 4;
  bool kahan = false;
  bool park = false;         // compensated phase parked in device memory between samples
  long long n_waves = 0;
  unsigned grid_x = 0;
  size_t lds = 0;
  dmw::WalkArgs args{};
  DevBuf Gs, GsH, Gsum, chi, echo, objs, cstart, citems, r0, inj_n, inj_u, pm_start;
  DevBuf partials, echo_partials, sig, park_buf, wstate;
  // walk in progress (dmw_plan_advance): next sample, its seed, the outputs it
  // was started with, which half of wstate holds the walkers
  int t_next = 0;
  uint64_t walk_seed = 0;
  unsigned out_mask = 0;
  int state_half = 0;
  DevBuf o_pos, o_c0, o_cf, o_cfull, o_pfull, o_ncross, o_logw, o_phi, o_counters;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  ~dmw_plan() {
    for (DevBuf* b : {&Gs, &GsH, &Gsum, &chi, &echo, &objs, &cstart, &citems, &r0, &inj_n, &inj_u, &pm_start, &partials,
                      &echo_partials, &sig, &park_buf, &wstate, &o_pos, &o_c0, &o_cf, &o_cfull, &o_pfull, &o_ncross,
                      &o_logw, &o_phi, &o_counters})
      b->release();
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
  }
};

extern "C" {

int dmw_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int dmw_set_device(int device) {
  g_device = device;
  return DMW_OK;
}

int dmw_device_arch(int32_t device, char* buf, int32_t len) {
  if (!buf || len < 1) return fail(DMW_EINVAL, "null or empty buffer");
  buf[0] = '\0';
  hipDeviceProp_t prop;
  const hipError_t e = hipGetDeviceProperties(&prop, device);
  if (e != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
  std::snprintf(buf, (size_t)len, "%s", prop.gcnArchName);
  return DMW_OK;
}

const char* dmw_last_error(void) { return g_err.c_str(); }

// Sub-steps per sample from which the phase is parked in device memory.
constexpr int kParkMinSub = 32;

// Measurements per kernel pass.  Blocks of 32 or more without compensation use
// the matrix-core phase update; the compensated (Kahan) form keeps phase and
// compensation in VGPRs and is limited to 64 measurements per pass.
static int pick_meas_block(int n_meas, int requested, bool kahan) {
  if (requested > 0) {
    for (int mb : dmw::kMeasBlocks)
      if (mb == requested && !(kahan && mb > 64)) return mb;
    return -1;
  }
  if (n_meas <= 4) return 4;
  if (n_meas <= 16) return 16;
  if (n_meas <= 32) return 32;
  if (n_meas <= 64) return 64;
  if (kahan) return 64;
  return n_meas <= 96 ? 96 : 128;
}

int dmw_plan_create(dmw_plan** out, const dmw_geometry* geom, const dmw_waveform* wave,
                    const dmw_run* run) {
  if (!out || !geom || !wave || !run) return fail(DMW_EINVAL, "null argument");
  *out = nullptr;
  if (run->n_walkers < 1 || run->n_walkers > (1LL << 31)) return fail(DMW_EINVAL, "n_walkers must be in [1, 2^31]");
  if (run->first_walker < 0 || run->first_walker + run->n_walkers > (1LL << 32))
    return fail(DMW_EINVAL, "first_walker must be >= 0 with first_walker + n_walkers <= 2^32");
  if (wave->n_meas < 1 || wave->n_t < 1 || !wave->G) return fail(DMW_EINVAL, "empty waveform");
  if (run->n_sub < 1) return fail(DMW_EINVAL, "n_sub must be >= 1");
  if ((long long)wave->n_t * run->n_sub >= (1LL << 32)) return fail(DMW_EINVAL, "n_t * n_sub too large");
  if (geom->kind < DMW_FREE || geom->kind > DMW_PACKED_MYELINATED) return fail(DMW_EINVAL, "unknown geometry kind");
  if (geom->wall < DMW_WALL_REFLECT || geom->wall > DMW_WALL_PERMEATE) return fail(DMW_EINVAL, "unknown wall mode");
  if (geom->kind == DMW_FREE && geom->wall != DMW_WALL_REFLECT)
    return fail(DMW_EUNSUPPORTED, "FreeDiffusion has no wall");
  if (geom->kind == DMW_BOX1D && geom->wall == DMW_WALL_PERMEATE)
    return fail(DMW_EUNSUPPORTED, "Box1D has no permeable wall");
  if ((geom->kind == DMW_PERMEABLE_SHELL || geom->kind == DMW_PERMEABLE_SHELL_CYL) &&
      geom->wall == DMW_WALL_SURFACE)
    return fail(DMW_EUNSUPPORTED, "PermeableShell relaxes only on its permeable membrane");
  if ((geom->kind == DMW_PERMEABLE_SHELL || geom->kind == DMW_PERMEABLE_SHELL_CYL) &&
      !(geom->aux > geom->size && geom->size > 0.0f))
    return fail(DMW_EINVAL, "PermeableShell needs 0 < r_inner < r_outer");
  if (const char* why = dmw::outside_float_envelope(*geom, run->step_len)) return fail(DMW_EINVAL, why);
  const bool myel = geom->kind == DMW_MYELINATED_CYLINDER;
  if (myel) {
    if (!geom->myelin) return fail(DMW_EINVAL, "MyelinatedCylinder needs its compartment constants");
    if (geom->wall != DMW_WALL_REFLECT)
      return fail(DMW_EUNSUPPORTED, "MyelinatedCylinder handles its membranes itself (wall must be reflect)");
    if (run->n_sub != 1) return fail(DMW_EINVAL, "MyelinatedCylinder takes one step per sample (n_sub = 1)");
    if (!(geom->aux > geom->size && geom->size > 0.0f))
      return fail(DMW_EINVAL, "outer_radius must be > inner_radius");
    const int64_t* ns = geom->myelin->n_start;
    if (!run->r0 && !(0 <= ns[0] && ns[0] <= ns[1] && ns[1] <= run->first_walker + run->n_walkers))
      return fail(DMW_EINVAL, "MyelinatedCylinder start counts must satisfy 0 <= n0 <= n01 <= n_walkers");
  }
  const bool pmyel = geom->kind == DMW_PACKED_MYELINATED;
  if (pmyel) {
    const dmw_packed_myelin* pm = geom->packed_myelin;
    if (!pm || !pm->inner_radii || !pm->step_len || !pm->two_kappa_over_D || !pm->start)
      return fail(DMW_EINVAL, "PackedMyelinatedCylinders needs its per-axon tables");
    if (geom->n_obj < 1 || !geom->centers || !geom->radii || !(geom->cell > 0.0f))
      return fail(DMW_EINVAL, "PackedMyelinatedCylinders needs axons and a positive cell side");
    if (geom->wall != DMW_WALL_REFLECT)
      return fail(DMW_EUNSUPPORTED, "PackedMyelinatedCylinders handles its walls itself (wall must be reflect)");
    if (run->r0)
      return fail(DMW_EUNSUPPORTED, "simulate(r0=...) is unsupported for PackedMyelinatedCylinders; it "
                                    "initialises walkers (and their compartments) from `seed`");
    for (int k = 0; k < geom->n_obj; ++k)
      if (!(pm->inner_radii[k] > 0.0f && geom->radii[k] > pm->inner_radii[k]))
        return fail(DMW_EINVAL, "PackedMyelinatedCylinders needs 0 < inner < outer radius per axon");
    const int64_t* st0 = pm->start;
    if (st0[0] != 0) return fail(DMW_EINVAL, "start boundaries must begin at 0");
    for (int j = 0; j < 2 * geom->n_obj; ++j)
      if (st0[j + 1] < st0[j]) return fail(DMW_EINVAL, "start boundaries must be ascending");
    if (st0[2 * geom->n_obj] > run->first_walker + run->n_walkers)
      return fail(DMW_EINVAL, "start boundaries exceed n_walkers");
  }
  const bool packed = geom->kind == DMW_PACKED_CYLINDERS || geom->kind == DMW_PACKED_SPHERES;
  if (packed && (geom->n_obj < 1 || !geom->centers || !geom->radii || !(geom->cell > 0.0f)))
    return fail(DMW_EINVAL, "packed substrate needs objects and a positive cell side");
  if (wave->n_echo < 0 || (wave->n_echo > 0 && !wave->echo_idx)) return fail(DMW_EINVAL, "bad echo list");
  for (int e = 0; e < wave->n_echo; ++e) {
    const int v = wave->echo_idx[e];
    if (v < 0 || v >= wave->n_t || (e > 0 && v <= wave->echo_idx[e - 1]))
      return fail(DMW_EINVAL, "echo indices must be ascending and inside [0, n_t)");
  }
  const long long N = run->n_walkers;
  const long long S = (long long)wave->n_t * run->n_sub;
  if (run->inj_normals && (geom->wall == DMW_WALL_PERMEATE || myel || pmyel) && !run->inj_uniforms)
    return fail(DMW_EINVAL, "injected walk on a permeable wall needs injected uniforms");

  hipStream_t st;
  int rc = ensure_stream(&st);
  if (rc) return rc;

  dmw_plan* P = new dmw_plan();
  P->device = g_device;
  P->geom = *geom;
  P->geom.myelin = nullptr;
  P->geom.packed_myelin = nullptr;
  P->run = *run;
  P->inv_norm = 1.0 / (double)N;
  P->n_meas = wave->n_meas;
  P->n_t = wave->n_t;
  P->n_echo = wave->n_echo;
  P->kahan = run->kahan != 0;
  P->mb = pick_meas_block(wave->n_meas, run->meas_block, P->kahan);
  if (P->mb < 0) {
    delete P;
    return fail(DMW_EINVAL, "meas_block must be 0, 4, 16, 32, 64, 96 or 128 (at most 64 with kahan)");
  }
  // Many sub-steps per sample: the phase is touched once per sample, so it is
  // parked in device memory, compensated, and the sub-step loop runs without
  // phase registers (four wavefronts per SIMD instead of two).
  if (run->meas_block == 0 && run->n_sub >= kParkMinSub && wave->n_meas > 16) {
    P->park = true;
    P->kahan = true;
    P->mb = dmw::kParkBlock;
  }
  P->n_chunks = (wave->n_meas + P->mb - 1) / P->mb;
  P->n_meas_pad = P->n_chunks * P->mb;
  P->grid_x = (unsigned)((N + dmw::kBlock - 1) / dmw::kBlock);
  P->n_waves = (long long)P->grid_x * (dmw::kBlock / dmw::kWave);

#define TRY(x)        \
  do {                \
    int r_ = (x);     \
    if (r_) {         \
      delete P;       \
      return r_;      \
    }                 \
  } while (0)

  dmw::WalkArgs& a = P->args;
  a.size = geom->size;
  a.inv_size = geom->size > 0.0f ? dmw::recip_rn(geom->size) : 0.0f;
  a.size_sq_cut = geom->size > 0.0f ? dmw::sqrt_cut(geom->size) : 0.0f;
  a.size_sq_cut_gt = geom->size > 0.0f ? dmw::sqrt_cut(std::nextafter(geom->size, INFINITY)) : 0.0f;
  a.cell = geom->cell;
  a.inv_cell = geom->cell > 0.0f ? 1.0f / geom->cell : 0.0f;
  a.eps = geom->eps;
  a.nudge = geom->nudge;
  a.rhoD = geom->rho_over_D;
  a.kappaD = geom->kappa_over_D;
  a.ax = geom->axes[0];
  a.ay = geom->axes[1];
  a.az = geom->axes[2];
  a.aux = geom->aux;
  if (myel) {
    const dmw_myelin& m = *geom->myelin;
    P->myelin = m;
    P->geom.myelin = &P->myelin;
    for (int c = 0; c < 3; ++c) {
      a.m_step[c] = m.step_len[c];
      a.m_q2[c] = m.dt_over_T2[c];
      a.m_wf[c] = m.spin_weight[c];
    }
    for (int k = 0; k < 4; ++k) a.m_k2[k] = m.two_kappa_over_D[k];
    a.m_n0 = m.n_start[0];
    a.m_n01 = m.n_start[1];
    // Ensemble normalisation: sum of the start compartments' spin weights, with
    // the kernel's own start rule (index ranges, or classify_position of r0).
    double wsum = 0.0;
    if (run->r0) {
      const float R0 = geom->size * geom->size, R1 = geom->aux * geom->aux;
      for (long long w = 0; w < N; ++w) {
        const float x = run->r0[w * 3], y = run->r0[w * 3 + 1];
        const float rr = x * x + y * y;
        wsum += (double)m.spin_weight[rr < R0 ? 0 : (rr < R1 ? 1 : 2)];
      }
    } else {
      // start compartments by global walker index, counted over this call's range
      const long long f = run->first_walker;
      const long long c0 = std::min(N, std::max(0LL, (long long)m.n_start[0] - f));
      const long long c01 = std::min(N, std::max(0LL, (long long)m.n_start[1] - f));
      wsum = (double)c0 * m.spin_weight[0] + (double)(c01 - c0) * m.spin_weight[1] +
             (double)(N - c01) * m.spin_weight[2];
    }
    P->inv_norm = 1.0 / wsum;
  }
  a.n_obj = 0;
  // LDS per workgroup: object tables and grid (dynamic) plus, for the matrix-core
  // phase update, 8 KB of operand stage and start points and 128 x mb bytes of
  // waveform operands.  The grid is coarsened to fit; only if even one cell does
  // not fit (and no meas_block was requested) is a narrower block taken.
  auto static_lds = [&](int mb) -> size_t { return (mb >= 32 && !P->kahan) ? 8192 + 128 * (size_t)mb : 0; };
  auto fit_grid = [&](float step, int dim, size_t obj_bytes, Grid& G, double hyst = 0.0,
                      int max_list = 0) -> bool {
    for (;;) {
      const size_t budget = 65536 - static_lds(P->mb);
      G = build_grid(*geom, step, dim, budget, obj_bytes, hyst, max_list);
      if ((size_t)geom->n_obj * obj_bytes + (G.start.size() + G.items.size()) * 4 <= budget) return true;
      if (run->meas_block != 0 || P->mb <= 16 || P->park) return false;
      P->mb = P->mb > 32 ? P->mb / 2 : 16;
    }
  };
  const char* lds_msg = "object tables and neighbour grid exceed the LDS of a workgroup";
  if (pmyel) {
    const dmw_packed_myelin& pm = *geom->packed_myelin;
    const int n = geom->n_obj;
    float smax = 0.0f;
    std::vector<float> rows((size_t)n * 12, 0.0f);
    for (int k = 0; k < n; ++k) {
      float* r0 = &rows[(size_t)k * 4];
      r0[0] = geom->centers[k * 2];
      r0[1] = geom->centers[k * 2 + 1];
      r0[2] = pm.inner_radii[k];
      r0[3] = geom->radii[k];
      float* r1 = &rows[((size_t)n + k) * 4];
      for (int c = 0; c < 3; ++c) {
        r1[c] = pm.step_len[k * 3 + c];
        smax = std::max(smax, r1[c]);
      }
      float* r2 = &rows[((size_t)2 * n + k) * 4];
      for (int c = 0; c < 4; ++c) r2[c] = pm.two_kappa_over_D[k * 4 + c];
    }
    Grid G;
    if (!fit_grid(smax, 2, 48, G)) {
      delete P;
      return fail(DMW_EUNSUPPORTED, lds_msg);
    }
    TRY(upload(P->objs, rows.data(), rows.size() * 4, st));
    TRY(upload(P->cstart, G.start.data(), G.start.size() * 4, st));
    TRY(upload(P->citems, G.items.data(), std::max<size_t>(1, G.items.size()) * 4, st));
    TRY(upload(P->pm_start, pm.start, ((size_t)2 * n + 1) * 8, st));
    a.n_obj = n;
    a.grid_n = G.n;
    a.grid_h = G.h;
    a.grid_inv_h = 1.0f / G.h;
    a.n_items = (int)G.items.size();
    a.exact_minimage = 1;
    P->lds = (size_t)n * 48 + (G.start.size() + G.items.size()) * 4;
    for (int c = 0; c < 3; ++c) a.pm_q2[c] = pm.dt_over_T2[c];
    a.pm_rhoD = pm.rho_over_D;
    a.pm_wmy = pm.myelin_weight;
    a.pm_start = P->pm_start.as<long long>();
    const long long f = run->first_walker;
    const long long n_my = std::max(0LL, std::min(f + N, (long long)pm.start[2 * n]) -
                                             std::max(f, (long long)pm.start[n]));
    P->inv_norm = 1.0 / ((double)(N - n_my) + (double)n_my * (double)pm.myelin_weight);
  }
  if (packed) {
    const int dim = geom->kind == DMW_PACKED_CYLINDERS ? 2 : 3;
    // Permeable packings walk the screened path: cells expanded by half a side
    // (a walker keeps its cell's list until it leaves that box), less where the
    // list would exceed the screen's object cache; the list starts carry each
    // cell's box level.
    const bool screened = geom->wall == DMW_WALL_PERMEATE;
    const double hyst = screened ? 0.5 : 0.0;
    Grid G;
    if (!fit_grid(run->step_len, dim, 16, G, hyst, screened ? dmw::SCR_MAXC : 0)) {
      delete P;
      return fail(DMW_EUNSUPPORTED, lds_msg);
    }
    const std::vector<int> starts = G.encoded_start();
    std::vector<float> objs((size_t)geom->n_obj * 4, 0.0f);
    for (int k = 0; k < geom->n_obj; ++k) {
      for (int d = 0; d < dim; ++d) objs[k * 4 + d] = geom->centers[k * dim + d];
      objs[k * 4 + 3] = geom->radii[k];
    }
    TRY(upload(P->objs, objs.data(), objs.size() * 4, st));
    TRY(upload(P->cstart, starts.data(), starts.size() * 4, st));
    TRY(upload(P->citems, G.items.data(), std::max<size_t>(1, G.items.size()) * 4, st));
    a.n_obj = geom->n_obj;
    a.grid_n = G.n;
    a.grid_h = G.h;
    a.grid_inv_h = 1.0f / G.h;
    a.n_items = (int)G.items.size();
    a.exact_minimage = G.exact_minimage ? 1 : 0;
    P->lds = (size_t)geom->n_obj * 16 + (G.start.size() + G.items.size()) * 4;
    if (screened) {
      // The screen caches nearest images, which needs images unambiguous over a
      // cell's box (no true-division minimum images); its margin covers the float32
      // rounding of positions, images and the reference's roots (screen.h).
      double rmax = 0.0;
      for (int k = 0; k < geom->n_obj; ++k) rmax = std::max(rmax, (double)geom->radii[k]);
      const double m = dmw::screen_margin(geom->cell, rmax, run->step_len);
      const double smax = (double)run->step_len;
      // cached lanes test at most 32 listed objects by bit mask
      int max_len = 0;
      for (size_t c = 0; c + 1 < G.start.size(); ++c) max_len = std::max(max_len, G.start[c + 1] - G.start[c]);
      a.scr_cache = (G.exact_minimage || max_len > 32) ? 0 : 1;
      a.scr_w = (float)(hyst * G.h);
      a.scr_m2 = (float)(2.0 * m);
      a.scr_cin = (float)(m * m);
      a.scr_cout = (float)(m * m + 0.25 * smax * smax * (1.0 + 0x1p-20));
    }
  }
  P->n_chunks = (wave->n_meas + P->mb - 1) / P->mb;
  P->n_meas_pad = P->n_chunks * P->mb;
  double g_scale = 1.0;
  // Waveform: (n_meas, n_t, 3) -> (n_t, n_meas_pad, 3) scaled by gamma*dt_sub.
  {
    std::vector<float> gs((size_t)wave->n_t * P->n_meas_pad * 3, 0.0f);
    const double gdt = run->gamma_dt_sub;
    for (int m = 0; m < wave->n_meas; ++m)
      for (int t = 0; t < wave->n_t; ++t)
        for (int c = 0; c < 3; ++c)
          gs[((size_t)t * P->n_meas_pad + m) * 3 + c] =
              (float)(gdt * (double)wave->G[((size_t)m * wave->n_t + t) * 3 + c]);
    TRY(upload(P->Gs, gs.data(), gs.size() * 4, st));
    // float16 hi/lo operands and column sums of the matrix-core phase update
    g_scale = dmw::phase_g_scale(gs.data(), gs.size());
    std::vector<uint2> gh;
    dmw::phase_half_table(gs.data(), wave->n_t, P->n_meas_pad, g_scale, gh);
    TRY(upload(P->GsH, gh.data(), gh.size() * sizeof(uint2), st));
    std::vector<float> gsum;
    dmw::phase_column_sums(gs.data(), wave->n_t, P->n_meas_pad, wave->echo_idx, wave->n_echo, gsum);
    TRY(upload(P->Gsum, gsum.data(), gsum.size() * 4, st));
  }
  if (wave->chi) TRY(upload(P->chi, wave->chi, (size_t)wave->n_t * 4, st));
  if (wave->n_echo > 0) TRY(upload(P->echo, wave->echo_idx, (size_t)wave->n_echo * 4, st));

  a.objs = P->objs.as<float4>();
  a.cell_start = P->cstart.as<int>();
  a.cell_items = P->citems.as<int>();
  a.n_sub = run->n_sub;
  a.step_len = run->step_len;
  a.dt_sub = run->dt_sub;
  a.inv_T2 = run->inv_T2;
  a.inv_T1 = run->inv_T1;
  a.has_t2 = run->has_t2;
  a.has_t1 = run->has_t1;
  a.n_t = wave->n_t;
  a.t_begin = 0;
  a.t_end = wave->n_t;
  a.n_meas = wave->n_meas;
  a.n_meas_pad = P->n_meas_pad;
  a.Gs = P->Gs.as<float>();
  a.GsH = P->GsH.as<uint2>();
  a.Gsum = P->Gsum.as<float>();
  {
    double step_max = run->step_len;
    if (myel)
      for (int c = 0; c < 3; ++c) step_max = std::max(step_max, (double)a.m_step[c]);
    if (pmyel)
      for (int k = 0; k < 3 * geom->n_obj; ++k)
        step_max = std::max(step_max, (double)geom->packed_myelin->step_len[k]);
    const double p_scale = dmw::phase_p_scale(wave->n_t, run->n_sub, step_max, 1.0);
    a.p_scale = (float)p_scale;
    a.inv_pg_scale = (float)(1.0 / (p_scale * g_scale));
  }
  a.chi = wave->chi ? P->chi.as<float>() : nullptr;
  a.n_echo = wave->n_echo;
  a.echo_idx = wave->n_echo > 0 ? P->echo.as<int>() : nullptr;
  a.n_walkers = N;
  a.first_walker = run->first_walker;
  a.instance = 0;

  if (run->r0) {
    std::vector<float> soa((size_t)N * 3);
    for (long long w = 0; w < N; ++w)
      for (int c = 0; c < 3; ++c) soa[(size_t)c * N + w] = run->r0[w * 3 + c];
    TRY(upload(P->r0, soa.data(), soa.size() * 4, st));
  }
  a.r0 = run->r0 ? P->r0.as<float>() : nullptr;
  if (run->inj_normals) {
    std::vector<float> t((size_t)S * 3 * N);
    for (long long w = 0; w < N; ++w)
      for (long long s = 0; s < S; ++s)
        for (int c = 0; c < 3; ++c) t[((size_t)s * 3 + c) * N + w] = run->inj_normals[((size_t)w * S + s) * 3 + c];
    TRY(upload(P->inj_n, t.data(), t.size() * 4, st));
    if (run->inj_uniforms) {
      std::vector<float> u((size_t)S * N);
      for (long long w = 0; w < N; ++w)
        for (long long s = 0; s < S; ++s) u[(size_t)s * N + w] = run->inj_uniforms[(size_t)w * S + s];
      TRY(upload(P->inj_u, u.data(), u.size() * 4, st));
    }
  }
  a.inj_n = run->inj_normals ? P->inj_n.as<float>() : nullptr;
  a.inj_u = (run->inj_normals && run->inj_uniforms) ? P->inj_u.as<float>() : nullptr;

  TRY(P->partials.alloc((size_t)P->n_waves * P->n_meas_pad * 2 * sizeof(double)));
  if (P->park)   // per wavefront and pass: mb/4 groups x (phase, compensation) x 64 lanes of float4
    TRY(P->park_buf.alloc((size_t)P->n_chunks * P->n_waves * (P->mb / 4) * 2 * dmw::kWave * 16));
  a.park = P->park ? P->park_buf.as<float4>() : nullptr;
  if (wave->n_echo > 0)
    TRY(P->echo_partials.alloc((size_t)P->n_waves * wave->n_echo * P->n_meas_pad * 2 * sizeof(double)));
  TRY(P->sig.alloc((size_t)std::max(1, wave->n_echo) * P->n_meas_pad * 2 * sizeof(double)));
  a.partials = P->partials.as<double>();
  a.echo_partials = wave->n_echo > 0 ? P->echo_partials.as<double>() : nullptr;
  hipError_t e = hipSuccess;
  if (!g_arena) {   // allocation-free calls keep no timing events
    e = hipEventCreate(&P->ev0);
    if (e == hipSuccess) e = hipEventCreate(&P->ev1);
  }
  if (e != hipSuccess) {
    delete P;
    return hip_fail(e, "hipEventCreate");
  }
  e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    delete P;
    return hip_fail(e, "upload");
  }
#undef TRY
  *out = P;
  return DMW_OK;
}

static hipError_t launch_geom(int kind, int wall, int mb, bool kahan, bool park, dim3 grid, size_t lds,
                              hipStream_t st, const dmw::WalkArgs& a) {
  switch (kind) {
    case DMW_FREE: return dmw::launch_walk_free(wall, mb, kahan, park, grid, lds, st, a);
    case DMW_BOX1D: return dmw::launch_walk_box1d(wall, mb, kahan, park, grid, lds, st, a);
    case DMW_CYLINDER: return dmw::launch_walk_cyl(wall, mb, kahan, park, grid, lds, st, a);
    case DMW_SPHERE: return dmw::launch_walk_sph(wall, mb, kahan, park, grid, lds, st, a);
    case DMW_PACKED_CYLINDERS: return dmw::launch_walk_pcyl(wall, mb, kahan, park, grid, lds, st, a);
    case DMW_PACKED_SPHERES: return dmw::launch_walk_psph(wall, mb, kahan, park, grid, lds, st, a);
    case DMW_ELLIPSOID: return dmw::launch_walk_ell(wall, mb, kahan, park, grid, lds, st, a);
    case DMW_PERMEABLE_SLAB: return dmw::launch_walk_slab(wall, mb, kahan, park, grid, lds, st, a);
    case DMW_PERMEABLE_SHELL: return dmw::launch_walk_shell(wall, mb, kahan, park, grid, lds, st, a);
    case DMW_PERMEABLE_SHELL_CYL: return dmw::launch_walk_shellc(wall, mb, kahan, park, grid, lds, st, a);
    case DMW_MYELINATED_CYLINDER: return dmw::launch_walk_myel(wall, mb, kahan, park, grid, lds, st, a);
    case DMW_PACKED_MYELINATED: return dmw::launch_walk_pmyel(wall, mb, kahan, park, grid, lds, st, a);
    default: return hipErrorInvalidValue;
  }
}

// Outputs a result asks for, as a bit mask (a walk split over several calls
// keeps the set it started with).
static unsigned out_mask_of(const dmw_result* out) {
  const void* f[] = {out->signal_im, out->final_pos, out->comp_origin, out->comp_final, out->comp_full,
                     out->pos_full,  out->n_cross,   out->log_w,       out->phi};
  unsigned m = out->want_counters ? 1u : 0u;
  for (int i = 0; i < 9; ++i)
    if (f[i]) m |= 2u << i;
  return m;
}

static int phase_form_of(const dmw_plan* P) {
  return P->park ? DMW_PHASE_PARKED
                 : (P->kahan ? DMW_PHASE_COMPENSATED
                             : (P->mb >= 32 ? DMW_PHASE_MATRIX_CORES : DMW_PHASE_REGISTERS));
}

// Walks samples [t_begin, t_end) of the plan's walk.  The launch that starts at
// sample 0 clears the counters; one that stops before n_t keeps the walkers in
// P->wstate (two halves: the launch reads one and writes the other, since the
// measurement passes of one launch read the state while its primary pass
// writes) and returns kernel_ms only; the launch that reaches n_t reduces the
// ensemble and fills *out.
static int plan_walk(dmw_plan* P, uint64_t seed, int t_begin, int t_end, dmw_result* out) {
  hipStream_t st;
  int rc = ensure_stream_on(P->device, &st);
  if (rc) return rc;
  const long long N = P->run.n_walkers;
  dmw::WalkArgs a = P->args;
  a.seed_lo = (uint32_t)(seed & 0xFFFFFFFFu);
  a.seed_hi = (uint32_t)(seed >> 32);
  a.t_begin = t_begin;
  a.t_end = t_end;
  if (t_begin > 0 || t_end < P->n_t) {
    rc = P->wstate.alloc((size_t)2 * 3 * N * sizeof(float4));
    if (rc) return rc;
    float4* half = P->wstate.as<float4>();
    a.state_in = half + (size_t)P->state_half * 3 * N;
    a.state_out = half + (size_t)(1 - P->state_half) * 3 * N;
  }

#define OUTBUF(field, buf, bytes)                   \
  if (out->field) {                                 \
    rc = P->buf.alloc(bytes);                       \
    if (rc) return rc;                              \
  }
  OUTBUF(final_pos, o_pos, (size_t)N * 3 * 4);
  OUTBUF(comp_origin, o_c0, (size_t)N * 4);
  OUTBUF(comp_final, o_cf, (size_t)N * 4);
  OUTBUF(comp_full, o_cfull, (size_t)N * P->n_t * 4);
  OUTBUF(pos_full, o_pfull, (size_t)N * P->n_t * 3 * 4);
  OUTBUF(n_cross, o_ncross, (size_t)N * 4);
  OUTBUF(log_w, o_logw, (size_t)N * 4);
  OUTBUF(phi, o_phi, (size_t)N * P->n_meas_pad * 4);
#undef OUTBUF
  a.final_pos = out->final_pos ? P->o_pos.as<float>() : nullptr;
  a.comp_origin = out->comp_origin ? P->o_c0.as<int>() : nullptr;
  a.comp_final = out->comp_final ? P->o_cf.as<int>() : nullptr;
  a.comp_full = out->comp_full ? P->o_cfull.as<int>() : nullptr;
  a.pos_full = out->pos_full ? P->o_pfull.as<float>() : nullptr;
  a.n_cross = out->n_cross ? P->o_ncross.as<uint32_t>() : nullptr;
  a.log_w = out->log_w ? P->o_logw.as<float>() : nullptr;
  a.phi = out->phi ? P->o_phi.as<float>() : nullptr;
  a.counters = nullptr;
  if (out->want_counters) {
    rc = P->o_counters.alloc(16);
    if (rc) return rc;
    if (t_begin == 0) HIP_TRY(hipMemsetAsync(P->o_counters.p, 0, 16, st), "hipMemsetAsync");
    a.counters = P->o_counters.as<unsigned long long>();
  }

  const dim3 grid(P->grid_x, P->n_chunks);
  if (P->ev0) HIP_TRY(hipEventRecord(P->ev0, st), "hipEventRecord");
  HIP_TRY(launch_geom(P->geom.kind, P->geom.wall, P->mb, P->kahan, P->park, grid, P->lds, st, a), "walk kernel launch");
  if (P->ev1) HIP_TRY(hipEventRecord(P->ev1, st), "hipEventRecord");
  out->phase_form = phase_form_of(P);
  if (t_end < P->n_t) {
    HIP_TRY(hipStreamSynchronize(st), "walk");
    float ms = 0.0f;
    if (P->ev0) HIP_TRY(hipEventElapsedTime(&ms, P->ev0, P->ev1), "hipEventElapsedTime");
    out->kernel_ms = ms;
    P->state_half = 1 - P->state_half;
    return DMW_OK;
  }
  // CPMG returns the echo rows; a plain walk returns the signal after the last sample.
  const double inv_n = P->inv_norm;
  double* sig = P->sig.as<double>();
  if (P->n_echo > 0) {
    HIP_TRY(dmw::launch_reduce_partials(P->echo_partials.as<double>(), P->n_waves,
                                        P->n_echo * P->n_meas_pad * 2, inv_n, sig, st),
            "reduce launch");
  } else {
    HIP_TRY(dmw::launch_reduce_partials(P->partials.as<double>(), P->n_waves, P->n_meas_pad * 2, inv_n,
                                        sig, st),
            "reduce launch");
  }
  const int rows = std::max(1, P->n_echo);
  std::vector<double> hs((size_t)rows * P->n_meas_pad * 2);
  HIP_TRY(hipMemcpyAsync(hs.data(), sig, hs.size() * 8, hipMemcpyDeviceToHost, st), "hipMemcpyAsync D2H");
  HIP_TRY(hipStreamSynchronize(st), "walk");
  float ms = 0.0f;
  if (P->ev0) HIP_TRY(hipEventElapsedTime(&ms, P->ev0, P->ev1), "hipEventElapsedTime");
  out->kernel_ms = ms;
  for (double v : hs)
    if (std::isnan(v))
      return fail(DMW_EHIP, "walk produced a NaN signal (non-finite input, or a phase operand beyond the float16 range)");
  for (int r = 0; r < rows; ++r)
    for (int m = 0; m < P->n_meas; ++m) {
      out->signal[(size_t)r * P->n_meas + m] = hs[((size_t)r * P->n_meas_pad + m) * 2];
      if (out->signal_im) out->signal_im[(size_t)r * P->n_meas + m] = hs[((size_t)r * P->n_meas_pad + m) * 2 + 1];
    }

  auto d2h = [&](void* dst, const DevBuf& b, size_t bytes) -> int {
    HIP_TRY(hipMemcpy(dst, b.p, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    return DMW_OK;
  };
  if (out->final_pos) {
    std::vector<float> t((size_t)N * 3);
    if ((rc = d2h(t.data(), P->o_pos, t.size() * 4))) return rc;
    for (long long w = 0; w < N; ++w)
      for (int c = 0; c < 3; ++c) out->final_pos[w * 3 + c] = t[(size_t)c * N + w];
  }
  if (out->comp_origin && (rc = d2h(out->comp_origin, P->o_c0, (size_t)N * 4))) return rc;
  if (out->comp_final && (rc = d2h(out->comp_final, P->o_cf, (size_t)N * 4))) return rc;
  if (out->n_cross && (rc = d2h(out->n_cross, P->o_ncross, (size_t)N * 4))) return rc;
  if (out->log_w && (rc = d2h(out->log_w, P->o_logw, (size_t)N * 4))) return rc;
  if (out->comp_full) {
    std::vector<int> t((size_t)N * P->n_t);
    if ((rc = d2h(t.data(), P->o_cfull, t.size() * 4))) return rc;
    for (int s = 0; s < P->n_t; ++s)
      for (long long w = 0; w < N; ++w) out->comp_full[(size_t)w * P->n_t + s] = t[(size_t)s * N + w];
  }
  if (out->pos_full) {
    std::vector<float> t((size_t)N * P->n_t * 3);
    if ((rc = d2h(t.data(), P->o_pfull, t.size() * 4))) return rc;
    for (int s = 0; s < P->n_t; ++s)
      for (int c = 0; c < 3; ++c)
        for (long long w = 0; w < N; ++w)
          out->pos_full[((size_t)w * P->n_t + s) * 3 + c] = t[((size_t)s * 3 + c) * N + w];
  }
  if (out->phi) {
    std::vector<float> t((size_t)N * P->n_meas_pad);
    if ((rc = d2h(t.data(), P->o_phi, t.size() * 4))) return rc;
    for (int m = 0; m < P->n_meas; ++m)
      for (long long w = 0; w < N; ++w) out->phi[(size_t)w * P->n_meas + m] = t[(size_t)m * N + w];
  }
  if (out->want_counters) {
    unsigned long long c[2];
    if ((rc = d2h(c, P->o_counters, 16))) return rc;
    out->n_hits = c[0];
    out->n_cand = c[1];
  }
  return DMW_OK;
}

int dmw_plan_execute(dmw_plan* P, uint64_t seed, dmw_result* out) {
  if (!P || !out || !out->signal) return fail(DMW_EINVAL, "null argument");
  P->t_next = 0;   // abandons a walk in progress
  return plan_walk(P, seed, 0, P->n_t, out);
}

int dmw_plan_advance(dmw_plan* P, uint64_t seed, int32_t n_samples, dmw_result* out, int32_t* t_done) {
  if (!P || !out || !out->signal || !t_done) return fail(DMW_EINVAL, "null argument");
  if (n_samples < 1) return fail(DMW_EINVAL, "n_samples must be >= 1");
  const unsigned mask = out_mask_of(out);
  if (P->t_next == 0) {
    P->walk_seed = seed;
    P->out_mask = mask;
    P->state_half = 0;
  } else {
    if (seed != P->walk_seed)
      return fail(DMW_EINVAL, "dmw_plan_advance: the seed differs from the one the walk in progress started with");
    if (mask != P->out_mask)
      return fail(DMW_EINVAL, "dmw_plan_advance: the outputs asked for differ from those the walk in progress started with");
  }
  const int t0 = P->t_next;
  const int t1 = (int)std::min<long long>((long long)P->n_t, (long long)t0 + n_samples);
  if ((t0 > 0 || t1 < P->n_t) && !P->park)
    return fail(DMW_EUNSUPPORTED,
                "a walk split over several calls needs the parked phase form (meas_block 0, n_sub >= " +
                    std::to_string(kParkMinSub) + ", n_meas > 16); walk every sample in one call");
  const int rc = plan_walk(P, seed, t0, t1, out);
  P->t_next = (rc == DMW_OK && t1 < P->n_t) ? t1 : 0;
  *t_done = rc == DMW_OK ? t1 : 0;
  return rc;
}

int dmw_plan_rewind(dmw_plan* P) {
  if (!P) return fail(DMW_EINVAL, "null argument");
  P->t_next = 0;
  return DMW_OK;
}

int dmw_plan_destroy(dmw_plan* P) {
  delete P;
  return DMW_OK;
}

int dmw_walk(const dmw_geometry* geom, const dmw_waveform* wave, const dmw_run* run, dmw_result* out) {
  if (!run || !out) return fail(DMW_EINVAL, "null argument");
  dmw_plan* P = nullptr;
  int rc = dmw_plan_create(&P, geom, wave, run);
  if (rc) return rc;
  rc = dmw_plan_execute(P, run->seed, out);
  dmw_plan_destroy(P);
  return rc;
}

// dmw_walk with every device buffer (staged inputs, partial sums, outputs) in the
// device's scratch array: no hipMalloc / hipFree.  Set-up on the first call on a
// device: its stream and the array's device address.  Calls on one device are
// serialised here (they share the array, which keeps the last call's contents);
// every buffer is fully written before it is read, except the counters, which are
// cleared per call.
int dmw_walk_noalloc(const dmw_geometry* geom, const dmw_waveform* wave, const dmw_run* run,
                     dmw_result* out) {
  if (!run || !out) return fail(DMW_EINVAL, "null argument");
  const int dev = g_device;
  if (dev < 0 || dev >= kMaxDevices) return fail(DMW_ENODEV, "device index out of range");
  hipStream_t st;
  int rc = ensure_stream_on(dev, &st);   // makes dev current
  if (rc) return rc;
  std::call_once(g_scratch_once[dev], [dev] {
    void* ptr = nullptr;
    g_scratch_err[dev] = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    g_scratch_ptr[dev] = g_scratch_err[dev] == hipSuccess ? ptr : nullptr;
  });
  if (g_scratch_err[dev] != hipSuccess || !g_scratch_ptr[dev])
    return hip_fail(g_scratch_err[dev] != hipSuccess ? g_scratch_err[dev] : hipErrorInvalidSymbol,
                    "hipGetSymbolAddress(scratch)");
  std::lock_guard<std::mutex> lock(g_scratch_mutex[dev]);
  Arena arena{static_cast<char*>(g_scratch_ptr[dev]), kScratchBytes, 0};
  g_arena = &arena;
  dmw_plan* P = nullptr;
  rc = dmw_plan_create(&P, geom, wave, run);
  if (!rc) rc = dmw_plan_execute(P, run->seed, out);
  if (P) dmw_plan_destroy(P);
  g_arena = nullptr;
  return rc;
}


// Walk-once producer: the walk kernel with a one-measurement zero waveform and
// the float16 trajectory channels as its outputs (dmipy_sim/core.py:972-1529).
int dmw_traj_create(dmw_traj** out, const dmw_geometry* geom, const dmw_run* run, int32_t n_t,
                    int32_t channels, const float* lab_from_native) {
  if (!out || !geom || !run || n_t < 1) return fail(DMW_EINVAL, "null argument or empty trajectory");
  *out = nullptr;
  const long long N = run->n_walkers;
  if (N < 1) return fail(DMW_EINVAL, "n_walkers must be >= 1");
  if (geom->kind == DMW_MYELINATED_CYLINDER)
    return fail(DMW_EUNSUPPORTED, "MyelinatedCylinder uses a dedicated fused step kernel (no replay walk)");
  if (geom->kind == DMW_PACKED_MYELINATED)
    return fail(DMW_EUNSUPPORTED, "PackedMyelinatedCylinders trajectories are not provided by this producer");
  const bool has_radius = geom->kind == DMW_CYLINDER || geom->kind == DMW_SPHERE ||
                          geom->kind == DMW_PERMEABLE_SLAB || geom->kind == DMW_PERMEABLE_SHELL ||
                          geom->kind == DMW_PERMEABLE_SHELL_CYL;
  if ((channels & 2) && !(has_radius && geom->wall == DMW_WALL_PERMEATE))
    channels &= ~2;  // no occupancy rule (no radius, or impermeable): identically 0 (filled by the caller)
  std::vector<float> G((size_t)n_t * 3, 0.0f);
  dmw_waveform wave{1, n_t, G.data(), nullptr, 0, nullptr};
  dmw_run r = *run;
  r.kahan = 0;
  r.meas_block = 4;
  r.has_t2 = 0;
  r.has_t1 = 0;
  dmw_plan* P = nullptr;
  int rc = dmw_plan_create(&P, geom, &wave, &r);
  if (rc) return rc;
  hipStream_t st;
  rc = ensure_stream_on(P->device, &st);
  if (rc) {
    dmw_plan_destroy(P);
    return rc;
  }
  if (P->mb != 4 || P->kahan) {
    dmw_plan_destroy(P);
    return fail(DMW_EHIP, "producer walk needs the 4-measurement kernel variant");
  }
  dmw_traj* T = new dmw_traj();
  T->N = N;
  T->n_t = n_t;
  auto bail = [&](hipError_t e, const char* what) {
    delete T;
    dmw_plan_destroy(P);
    return hip_fail(e, what);
  };
  hipError_t e = hipMalloc((void**)&T->traj, (size_t)3 * n_t * N * 2);
  if (e == hipSuccess && (channels & 1)) e = hipMalloc((void**)&T->dlog, (size_t)n_t * N * 2);
  if (e == hipSuccess && (channels & 2)) e = hipMalloc((void**)&T->comp, (size_t)n_t * N * 2);
  if (e != hipSuccess) return bail(e, "hipMalloc trajectories");
  dmw::WalkArgs a = P->args;
  a.seed_lo = (uint32_t)(run->seed & 0xFFFFFFFFu);
  a.seed_hi = (uint32_t)(run->seed >> 32);
  a.traj_h = T->traj;
  a.dlog_h = T->dlog;
  a.comp_h = T->comp;
  a.traj_rotate = lab_from_native ? 1 : 0;
  for (int i = 0; i < 9; ++i) a.traj_R[i] = lab_from_native ? lab_from_native[i] : 0.0f;
  const dim3 grid(P->grid_x, P->n_chunks);
  e = hipEventRecord(P->ev0, st);
  if (e == hipSuccess) e = launch_geom(P->geom.kind, P->geom.wall, P->mb, P->kahan, P->park, grid, P->lds, st, a);
  if (e == hipSuccess) e = hipEventRecord(P->ev1, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  float ms = 0.0f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, P->ev0, P->ev1);
  if (e != hipSuccess) return bail(e, "trajectory walk");
  T->walk_ms = ms;
  dmw_plan_destroy(P);
  *out = T;
  return DMW_OK;
}

}  // extern "C"
