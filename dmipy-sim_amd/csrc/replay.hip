// replay.hip -- the replay contraction of the walk-once path on the matrix cores.
//
// Reference: dmipy_sim/trajectories.py:360-641 (replay) -- for every measurement
// m and walker w
//   phi[m,w]   = GAMMA dt_traj sum_t G[m,t,:] . r[w,t,:]      (float64 GEMM, 550-552)
//   log_w[m,w] = log_w_scalar[m] + (rho/D) sum_t chi[m,t] dlog[w,t]   (602-612)
//   S[m]       = mean_w exp(log_w[m,w]) cos(phi[m,w])          (637)
// with float16 positions saved by the producer walk (core.py:972-1529).
//
// Layout: trajectories stay in HBM as (3 n_t, N) float16 (row k = 3t + c), so
// the B operand of C[m, w] = sum_k G[m, k] T[k, w] is one coalesced 32-byte
// row segment per 16 walkers.  A workgroup of four wavefronts owns 64 walkers;
// each wavefront owns 16 walkers and MT tiles of 16 measurements, accumulating
// with v_mfma_f64_16x16x4f64 (lane l: A[l%16][l/16], B[l/16][l%16]; register r
// of lane l holds D[l/16 + 4r][l%16], checked on the GPU).  The A operands are
// pre-tiled on the host in lane order so every A load is one coalesced 512-byte
// read.  The epilogue evaluates exp(log_w) cos(phi) in float64, sums the 16
// walkers of each row across lanes, the four wavefronts in LDS, and writes one
// float64 partial per (workgroup, measurement); the fixed-order reduction of
// reduce.hip divides by N.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/dmipy_walk.h"
#include "batch.h"
#include "traj.h"
#include "walk_launch.h"

namespace dmw {

typedef double doublex4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ double h2d(uint16_t h) { return (double)(float)__builtin_bit_cast(_Float16, h); }

struct ReplayArgs {
  const uint16_t* traj;   // (3 n_t, N)
  const uint16_t* dlog;   // (n_t, N) or null
  const double* Gt;       // A tiles of G: ((ks * MTT + mt) * 64 + lane)
  const double* Ct;       // A tiles of chi (per-measurement chi), same layout over n_t
  const double* chi;      // (n_t) chi for the walker-scalar surface term
  const double* lw0;      // (M_pad) walker-independent log-weight
  double scale;           // GAMMA * dt_traj
  double rhoD;            // surface relaxivity / D
  long long N;
  int n_t, K, nk, nk2, MTT, M_pad, n_meas;
  double* partials;       // (signal workgroups, M_pad)
  double* phi_buf;        // (M_pad, N) phases
  double* lw_buf;         // (N) walker log-weight part (one chi row) or (M_pad, N) (per-measurement chi)
  double* phi_out;        // (n_meas, N) or null
  double* lw_out;         // (n_meas, N) or null
};

// SURF: 0 no surface term, 1 one chi for all measurements (walker scalar),
// 2 per-measurement chi (a second matrix-core contraction over n_t).
template <int MT, int SURF>
__global__ __launch_bounds__(256) void replay_kernel(const ReplayArgs a) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int kq = lane >> 4, jq = lane & 15;
  const long long w = (long long)blockIdx.x * 64 + wv * 16 + jq;
  const bool wok = w < a.N;
  const size_t wc = (size_t)(wok ? w : a.N - 1);
  const int mt0 = blockIdx.y * MT;

  doublex4 acc[MT];
#pragma unroll
  for (int i = 0; i < MT; ++i) acc[i] = doublex4{0.0, 0.0, 0.0, 0.0};
  // Branch-free operand loads (the padded k rows read row K-1 and are zeroed by
  // the mask), so the loads of consecutive k-steps are not serialised behind a
  // branch and a wait.
  const uint16_t* tcol = a.traj + wc;
  for (int ks = 0; ks < a.nk; ++ks) {
    const int k = 4 * ks + kq;
    const int kc = k < a.K ? k : a.K - 1;
    const double keep = k < a.K ? 1.0 : 0.0;
    const double b = keep * h2d(tcol[(size_t)kc * a.N]);
    const double* gt = a.Gt + ((size_t)ks * a.MTT + mt0) * 64 + lane;
#pragma unroll
    for (int i = 0; i < MT; ++i) acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(gt[i * 64], b, acc[i], 0, 0, 0);
  }

  // Phase rows of this pass; with a boundary channel also the log-weight part
  // that depends on the walker.
  // Row offsets (16 i + 4 r) N are wavefront-uniform; one per-lane base pointer.
  double* prow = a.phi_buf + (size_t)(16 * mt0 + kq) * a.N + wc;
#pragma unroll
  for (int i = 0; i < MT; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r)
      if (wok) prow[(size_t)(16 * i + 4 * r) * a.N] = a.scale * acc[i][r];
  if (SURF == 1 && blockIdx.y == 0) {
    double s_w = 0.0;
    for (int t = kq; t < a.n_t; t += 4) s_w += a.chi[t] * h2d(a.dlog[(size_t)t * a.N + wc]);
    s_w += __shfl_xor(s_w, 16, 64);
    s_w += __shfl_xor(s_w, 32, 64);
    if (wok && kq == 0) a.lw_buf[w] = a.rhoD * s_w;
  } else if (SURF == 2) {
    doublex4 acc2[MT];
#pragma unroll
    for (int i = 0; i < MT; ++i) acc2[i] = doublex4{0.0, 0.0, 0.0, 0.0};
    for (int ks = 0; ks < a.nk2; ++ks) {
      const int t = 4 * ks + kq;
      const int tc = t < a.n_t ? t : a.n_t - 1;
      const double b = (t < a.n_t ? 1.0 : 0.0) * h2d(a.dlog[(size_t)tc * a.N + wc]);
      const double* ct = a.Ct + ((size_t)ks * a.MTT + mt0) * 64 + lane;
#pragma unroll
      for (int i = 0; i < MT; ++i) acc2[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(ct[i * 64], b, acc2[i], 0, 0, 0);
    }
    double* lrow = a.lw_buf + (size_t)(16 * mt0 + kq) * a.N + wc;
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (wok) lrow[(size_t)(16 * i + 4 * r) * a.N] = a.rhoD * acc2[i][r];
  }
}

// Signal terms: one thread per walker, 16 measurements per workgroup row
// (grid.y); v = exp(log_w0[m] + walker part) cos(phi) in float64, summed over the
// workgroup's 256 walkers in a fixed order into partials[blockIdx.x][m].
template <int SURF>
__global__ __launch_bounds__(256) void replay_signal_kernel(const ReplayArgs a) {
  __shared__ double red[4][16];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const long long w = (long long)blockIdx.x * 256 + threadIdx.x;
  const bool wok = w < a.N;
  const int m0 = blockIdx.y * 16;
  const double lw_w = (SURF == 1 && wok) ? a.lw_buf[w] : 0.0;
  for (int j = 0; j < 16; ++j) {
    const int m = m0 + j;
    double v = 0.0;
    if (wok && m < a.n_meas) {
      const double phi = a.phi_buf[(size_t)m * a.N + w];
      double lw = a.lw0[m];
      if (SURF == 1) lw = lw + lw_w;
      if (SURF == 2) lw = lw + a.lw_buf[(size_t)m * a.N + w];
      v = exp(lw) * cos(phi);
      if (a.phi_out) a.phi_out[(size_t)m * a.N + w] = phi;
      if (a.lw_out) a.lw_out[(size_t)m * a.N + w] = lw;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) red[wv][j] = v;
  }
  __syncthreads();
  if (threadIdx.x < 16) {
    const int j = threadIdx.x;
    a.partials[(size_t)blockIdx.x * a.M_pad + m0 + j] = (red[0][j] + red[1][j]) + (red[2][j] + red[3][j]);
  }
}

// (rows, cols) -> (cols, rows) for 16-bit words, 64 x 64 tiles through LDS.
__global__ __launch_bounds__(256) void transpose16_kernel(const uint16_t* __restrict__ in, uint16_t* __restrict__ out,
                                                          long long rows, long long cols, long long row_tiles) {
  __shared__ uint16_t tile[64][66];
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const long long c0 = (long long)blockIdx.x * 64;
  for (long long by = blockIdx.y; by < row_tiles; by += gridDim.y) {
    const long long r0 = by * 64;
    for (int i = ty; i < 64; i += 4) {
      const long long r = r0 + i, c = c0 + tx;
      if (r < rows && c < cols) tile[i][tx] = in[r * cols + c];
    }
    __syncthreads();
    for (int i = ty; i < 64; i += 4) {
      const long long orow = c0 + i, ocol = r0 + tx;
      if (orow < cols && ocol < rows) out[orow * rows + ocol] = tile[tx][i];
    }
    __syncthreads();
  }
}

hipError_t launch_transpose16(const uint16_t* in, uint16_t* out, long long rows, long long cols, hipStream_t st) {
  const long long col_tiles = (cols + 63) / 64, row_tiles = (rows + 63) / 64;
  if (col_tiles > 0x7fffffffLL) return hipErrorInvalidValue;
  const unsigned gy = (unsigned)std::min<long long>(row_tiles, 65535);
  hipLaunchKernelGGL(transpose16_kernel, dim3((unsigned)col_tiles, gy), dim3(256), 0, st, in, out, rows, cols,
                     row_tiles);
  return hipGetLastError();
}

template <int MT, int SURF>
static hipError_t launch_mt(dim3 grid, hipStream_t st, const ReplayArgs& a) {
  hipLaunchKernelGGL((replay_kernel<MT, SURF>), grid, dim3(256), 0, st, a);
  return hipGetLastError();
}

static hipError_t launch_signal(int surf, dim3 grid, hipStream_t st, const ReplayArgs& a) {
  switch (surf) {
    case 0: hipLaunchKernelGGL((replay_signal_kernel<0>), grid, dim3(256), 0, st, a); break;
    case 1: hipLaunchKernelGGL((replay_signal_kernel<1>), grid, dim3(256), 0, st, a); break;
    default: hipLaunchKernelGGL((replay_signal_kernel<2>), grid, dim3(256), 0, st, a); break;
  }
  return hipGetLastError();
}

template <int MT>
static hipError_t launch_surf(int surf, dim3 grid, hipStream_t st, const ReplayArgs& a) {
  switch (surf) {
    case 0: return launch_mt<MT, 0>(grid, st, a);
    case 1: return launch_mt<MT, 1>(grid, st, a);
    default: return launch_mt<MT, 2>(grid, st, a);
  }
}

}  // namespace dmw

namespace {

struct Scratch {
  std::vector<void*> p;
  ~Scratch() {
    for (void* q : p) (void)hipFree(q);
  }
  hipError_t alloc(void** out, size_t bytes) {
    hipError_t e = hipMalloc(out, std::max<size_t>(bytes, 16));
    if (e == hipSuccess) p.push_back(*out);
    return e;
  }
};

#define RP_TRY(x, what)                                      \
  do {                                                       \
    hipError_t e_ = (x);                                     \
    if (e_ != hipSuccess) return dmw::batch_hip_fail(e_, what); \
  } while (0)

// A-operand tiles in lane order: out[((ks * MTT + mt) * 64 + l)] = src[16 mt + l%16][4 ks + l/16].
std::vector<double> tile_a(const double* src, int rows, int cols, int MTT, int nk) {
  std::vector<double> out((size_t)nk * MTT * 64, 0.0);
  for (int ks = 0; ks < nk; ++ks)
    for (int mt = 0; mt < MTT; ++mt)
      for (int l = 0; l < 64; ++l) {
        const int m = 16 * mt + (l & 15), k = 4 * ks + (l >> 4);
        if (m < rows && k < cols) out[((size_t)ks * MTT + mt) * 64 + l] = src[(size_t)m * cols + k];
      }
  return out;
}

}  // namespace

extern "C" {

int dmw_traj_from_host(dmw_traj** out, int64_t n_walkers, int32_t n_t, const uint16_t* traj,
                       const uint16_t* dlog) {
  if (!out || !traj || n_walkers < 1 || n_t < 1) return dmw::batch_fail(DMW_EINVAL, "null or empty trajectory");
  *out = nullptr;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return dmw::batch_fail(DMW_ENODEV, "no HIP device");
  dmw_traj* T = new dmw_traj();
  T->N = n_walkers;
  T->n_t = n_t;
  const size_t nt3 = (size_t)3 * n_t;
  Scratch s;
  void* tmp = nullptr;
  auto fail = [&](hipError_t e, const char* what) {
    delete T;
    return dmw::batch_hip_fail(e, what);
  };
  hipError_t e = hipMalloc((void**)&T->traj, nt3 * n_walkers * 2);
  if (e != hipSuccess) return fail(e, "hipMalloc trajectories");
  e = s.alloc(&tmp, nt3 * n_walkers * 2);
  if (e != hipSuccess) return fail(e, "hipMalloc trajectories");
  e = hipMemcpy(tmp, traj, nt3 * n_walkers * 2, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = dmw::launch_transpose16((const uint16_t*)tmp, T->traj, n_walkers, (long long)nt3, nullptr);
  if (e == hipSuccess && dlog) {
    e = hipMalloc((void**)&T->dlog, (size_t)n_t * n_walkers * 2);
    if (e == hipSuccess) e = hipMemcpy(tmp, dlog, (size_t)n_t * n_walkers * 2, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = dmw::launch_transpose16((const uint16_t*)tmp, T->dlog, n_walkers, n_t, nullptr);
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) return fail(e, "trajectory upload");
  *out = T;
  return DMW_OK;
}

int dmw_traj_download(dmw_traj* T, uint16_t* traj, uint16_t* dlog, uint16_t* comp) {
  if (!T) return dmw::batch_fail(DMW_EINVAL, "null trajectory set");
  if ((dlog && !T->dlog) || (comp && !T->comp))
    return dmw::batch_fail(DMW_EINVAL, "channel not recorded by this trajectory set");
  const size_t nt3 = (size_t)3 * T->n_t;
  Scratch s;
  void* tmp = nullptr;
  RP_TRY(s.alloc(&tmp, nt3 * T->N * 2), "hipMalloc");
  if (traj) {
    RP_TRY(dmw::launch_transpose16(T->traj, (uint16_t*)tmp, (long long)nt3, T->N, nullptr), "transpose");
    RP_TRY(hipMemcpy(traj, tmp, nt3 * T->N * 2, hipMemcpyDeviceToHost), "D2H");
  }
  if (dlog) {
    RP_TRY(dmw::launch_transpose16(T->dlog, (uint16_t*)tmp, T->n_t, T->N, nullptr), "transpose");
    RP_TRY(hipMemcpy(dlog, tmp, (size_t)T->n_t * T->N * 2, hipMemcpyDeviceToHost), "D2H");
  }
  if (comp) {
    RP_TRY(dmw::launch_transpose16(T->comp, (uint16_t*)tmp, T->n_t, T->N, nullptr), "transpose");
    RP_TRY(hipMemcpy(comp, tmp, (size_t)T->n_t * T->N * 2, hipMemcpyDeviceToHost), "D2H");
  }
  return DMW_OK;
}

int dmw_traj_info(dmw_traj* T, int64_t* n_walkers, int32_t* n_t, int32_t* channels, double* walk_ms) {
  if (!T) return dmw::batch_fail(DMW_EINVAL, "null trajectory set");
  if (n_walkers) *n_walkers = T->N;
  if (n_t) *n_t = T->n_t;
  if (channels) *channels = (T->dlog ? 1 : 0) | (T->comp ? 2 : 0);
  if (walk_ms) *walk_ms = T->walk_ms;
  return DMW_OK;
}

int dmw_traj_destroy(dmw_traj* T) {
  delete T;
  return DMW_OK;
}

int dmw_replay_run(dmw_traj* T, const dmw_replay* rp, double* signals, double* phi, double* log_w,
                   double* kernel_ms) {
  if (!T || !rp || !signals || !rp->G || !rp->log_w0) return dmw::batch_fail(DMW_EINVAL, "null argument");
  if (rp->n_meas < 1 || rp->n_t != T->n_t)
    return dmw::batch_fail(DMW_EINVAL, "the waveform must be resampled to the trajectory's n_t");
  const bool surf = rp->rho_over_D != 0.0;
  if (surf && (!T->dlog || !rp->chi || (rp->chi_rows != 1 && rp->chi_rows != rp->n_meas)))
    return dmw::batch_fail(DMW_EINVAL,
                           "surface relaxivity needs the boundary channel and chi of 1 or n_meas rows");
  const int n_meas = rp->n_meas, n_t = T->n_t;
  if (rp->meas_tiles != 0 && rp->meas_tiles != 1 && rp->meas_tiles != 2 && rp->meas_tiles != 4 &&
      rp->meas_tiles != 8)
    return dmw::batch_fail(DMW_EINVAL, "meas_tiles must be 0, 1, 2, 4 or 8");
  const int MT = rp->meas_tiles > 0 ? rp->meas_tiles : (n_meas <= 16 ? 1 : n_meas <= 32 ? 2 : n_meas <= 64 ? 4 : 8);
  const int MTT = ((n_meas + 16 * MT - 1) / (16 * MT)) * MT;
  const int M_pad = 16 * MTT;
  const int K = 3 * n_t, nk = (K + 3) / 4, nk2 = (n_t + 3) / 4;
  const int surf_mode = !surf ? 0 : (rp->chi_rows == 1 ? 1 : 2);

  std::vector<double> Gt = tile_a(rp->G, n_meas, K, MTT, nk);
  std::vector<double> lw0(M_pad, 0.0);
  for (int m = 0; m < n_meas; ++m) lw0[m] = rp->log_w0[m];
  std::vector<double> Ct;
  if (surf_mode == 2) Ct = tile_a(rp->chi, n_meas, n_t, MTT, nk2);

  const unsigned gx = (unsigned)((T->N + 63) / 64);
  const dim3 grid(gx, MTT / MT);
  Scratch s;
  void *dGt, *dCt = nullptr, *dchi = nullptr, *dlw0, *dpart, *dsig, *dphi = nullptr, *dlw = nullptr;
  hipStream_t st = nullptr;
  RP_TRY(s.alloc(&dGt, Gt.size() * 8), "hipMalloc");
  RP_TRY(hipMemcpy(dGt, Gt.data(), Gt.size() * 8, hipMemcpyHostToDevice), "H2D");
  RP_TRY(s.alloc(&dlw0, (size_t)M_pad * 8), "hipMalloc");
  RP_TRY(hipMemcpy(dlw0, lw0.data(), (size_t)M_pad * 8, hipMemcpyHostToDevice), "H2D");
  if (surf_mode == 2) {
    RP_TRY(s.alloc(&dCt, Ct.size() * 8), "hipMalloc");
    RP_TRY(hipMemcpy(dCt, Ct.data(), Ct.size() * 8, hipMemcpyHostToDevice), "H2D");
  }
  if (surf_mode == 1) {
    RP_TRY(s.alloc(&dchi, (size_t)n_t * 8), "hipMalloc");
    RP_TRY(hipMemcpy(dchi, rp->chi, (size_t)n_t * 8, hipMemcpyHostToDevice), "H2D");
  }
  const unsigned gs = (unsigned)((T->N + 255) / 256);
  void *dphib, *dlwb = nullptr;
  RP_TRY(s.alloc(&dphib, (size_t)M_pad * T->N * 8), "hipMalloc phases");
  if (surf_mode == 1) RP_TRY(s.alloc(&dlwb, (size_t)T->N * 8), "hipMalloc");
  if (surf_mode == 2) RP_TRY(s.alloc(&dlwb, (size_t)M_pad * T->N * 8), "hipMalloc");
  RP_TRY(s.alloc(&dpart, (size_t)gs * M_pad * 8), "hipMalloc");
  RP_TRY(s.alloc(&dsig, (size_t)M_pad * 8), "hipMalloc");
  if (phi) RP_TRY(s.alloc(&dphi, (size_t)n_meas * T->N * 8), "hipMalloc");
  if (log_w) RP_TRY(s.alloc(&dlw, (size_t)n_meas * T->N * 8), "hipMalloc");

  dmw::ReplayArgs a{};
  a.traj = T->traj;
  a.dlog = T->dlog;
  a.Gt = (const double*)dGt;
  a.Ct = (const double*)dCt;
  a.chi = (const double*)dchi;
  a.lw0 = (const double*)dlw0;
  a.scale = rp->scale;
  a.rhoD = surf ? rp->rho_over_D : 0.0;
  a.N = T->N;
  a.n_t = n_t;
  a.K = K;
  a.nk = nk;
  a.nk2 = nk2;
  a.MTT = MTT;
  a.M_pad = M_pad;
  a.n_meas = n_meas;
  a.partials = (double*)dpart;
  a.phi_buf = (double*)dphib;
  a.lw_buf = (double*)dlwb;
  a.phi_out = (double*)dphi;
  a.lw_out = (double*)dlw;

  hipEvent_t e0, e1;
  RP_TRY(hipEventCreate(&e0), "hipEventCreate");
  RP_TRY(hipEventCreate(&e1), "hipEventCreate");
  hipError_t e = hipEventRecord(e0, st);
  if (e == hipSuccess) {
    switch (MT) {
      case 1: e = dmw::launch_surf<1>(surf_mode, grid, st, a); break;
      case 2: e = dmw::launch_surf<2>(surf_mode, grid, st, a); break;
      case 4: e = dmw::launch_surf<4>(surf_mode, grid, st, a); break;
      default: e = dmw::launch_surf<8>(surf_mode, grid, st, a); break;
    }
  }
  if (e == hipSuccess) e = hipEventRecord(e1, st);
  if (e == hipSuccess) e = dmw::launch_signal(surf_mode, dim3(gs, M_pad / 16), st, a);
  if (e == hipSuccess)
    e = dmw::launch_reduce_partials((const double*)dpart, gs, M_pad, 1.0 / (double)T->N, (double*)dsig, st);
  std::vector<double> hs(M_pad);
  if (e == hipSuccess) e = hipMemcpy(hs.data(), dsig, (size_t)M_pad * 8, hipMemcpyDeviceToHost);
  float ms = 0.0f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (e != hipSuccess) return dmw::batch_hip_fail(e, "replay");
  for (int m = 0; m < n_meas; ++m) signals[m] = hs[m];
  if (phi) RP_TRY(hipMemcpy(phi, dphi, (size_t)n_meas * T->N * 8, hipMemcpyDeviceToHost), "D2H");
  if (log_w) RP_TRY(hipMemcpy(log_w, dlw, (size_t)n_meas * T->N * 8, hipMemcpyDeviceToHost), "D2H");
  if (kernel_ms) *kernel_ms = ms;
  return DMW_OK;
}

}  // extern "C"
