// mlp.hip -- the learned surrogate: a GELU MLP with BF16 weights and FP32
// accumulation on the matrix cores (v_mfma_f32_16x16x32_bf16), trained by its
// own forward, backward and Adam kernels.
//
// No counterpart in the reference; it maps substrate parameters plus a
// waveform embedding to the walk's signal and is trained on signals the walk
// produces (dmipy_sim_amd/surrogate.py).
//
// Layers: x (n_in, padded to a multiple of 32) -> [W_l, b_l, GELU] x depth ->
// w_out . h + b_out.  Per layer l the forward pass stores Z_l = X_l W_l + b_l
// (fp32) and Y_l = GELU(Z_l) (bf16, the next layer's input).  Backward:
//   dZ_out-layer    = dy * w_out * GELU'(Z_last)                 (bf16)
//   dY_{l-1}        = dZ_l . W_l^T, dZ_{l-1} = dY_{l-1} GELU'(Z_{l-1})
//   dW_l            = X_l^T . dZ_l (fp32),  db_l = sum_rows dZ_l
// Master weights, gradients and Adam moments are fp32; the MFMA operands are
// bf16 copies refreshed by the Adam kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/dmipy_walk.h"
#include "batch.h"

namespace dmw {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(__bf16 v) { return (float)v; }
__device__ __forceinline__ __bf16 f2bf(float v) { return (__bf16)v; }

// GELU, tanh form: 0.5 x (1 + tanh(c (x + 0.044715 x^3))), c = sqrt(2/pi), evaluated
// through the identity 0.5 (1 + tanh(u)) = 1 / (1 + exp(-2u)) = s: GELU = x s and
// GELU' = s + 2 x s (1 - s) c (1 + 0.134145 x^2).  exp via the hardware exp2
// (relative error ~1e-7, far below the bf16 rounding of the activations).
// The exponent -2 u log2(e) = x (A + B x^2) with A = -2 c log2(e) and
// B = 0.044715 A folded into one fma and two products.
__device__ __forceinline__ float gelu_sig(float x) {
  const float e = x * fmaf(-0.1029432f, x * x, -2.3022082f);
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(e));
}
__device__ __forceinline__ float gelu(float x) { return x * gelu_sig(x); }
__device__ __forceinline__ float gelu_grad(float x) {
  const float s = gelu_sig(x);
  return s + 2.0f * x * s * (1.0f - s) * 0.7978845608f * (1.0f + 0.134145f * x * x);
}

enum Epi { EPI_FWD = 0, EPI_DX = 1, EPI_DW = 2 };

constexpr int TM = 64, TN = 64, TK = 32, LDSK = 40;  // tile and padded LDS row (bf16)

// C[M x N] = A[M x K] . B[K x N] with bf16 operands and fp32 accumulation.
//   TA: A stored transposed (element (m,k) at A[k*lda + m]); else A[m*lda + k].
//   TB: B stored transposed (element (k,n) at B[n*ldb + k]); else B[k*ldb + n].
// M, N multiples of 64, K a multiple of 32 (checked on the host).  256 threads:
// 4 wavefronts in a 2x2 grid, each owning a 32x32 block = 2x2 MFMA 16x16 tiles.
//
// EPI_DW with colsum != nullptr: the blocks of the first row of tiles also sum
// the B operand over K (colsum[n] = sum_k B[k][n], the bias gradient when B is
// dZ) from the tiles already staged in LDS: thread (n = tid % 64, q = tid / 64)
// adds B[k][n] for k in the q-th quarter of each K tile.
template <int EPI, bool TA, bool TB>
__global__ __launch_bounds__(256) void gemm_bf16(int M, int N, int K, const __bf16* __restrict__ A, int lda,
                                                 const __bf16* __restrict__ B, int ldb,
                                                 const float* __restrict__ bias, float* __restrict__ Zout,
                                                 __bf16* __restrict__ Yout, const float* __restrict__ Zprev,
                                                 float* __restrict__ Cout, int ldc, float* __restrict__ colsum) {
  __shared__ __bf16 As[TM][LDSK];
  __shared__ __bf16 Bs[TN][LDSK];
  __shared__ float red[4][TN];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
  const int wm = (wv >> 1) * 32, wn = (wv & 1) * 32;
  const bool do_colsum = (EPI == EPI_DW) && colsum != nullptr && blockIdx.y == 0;
  float csum = 0.0f;
  f32x4 acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};

  for (int k0 = 0; k0 < K; k0 += TK) {
    // A tile 64 x 32: each thread moves 8 elements.
    if (!TA) {
      const int r = tid >> 2, c = (tid & 3) * 8;
      const bf16x8 v = *reinterpret_cast<const bf16x8*>(A + (size_t)(m0 + r) * lda + k0 + c);
      *reinterpret_cast<bf16x8*>(&As[r][c]) = v;
    } else {
      const int kk = tid >> 3, mm = (tid & 7) * 8;
      const bf16x8 v = *reinterpret_cast<const bf16x8*>(A + (size_t)(k0 + kk) * lda + m0 + mm);
      for (int e = 0; e < 8; ++e) As[mm + e][kk] = v[e];
    }
    // B tile 32 x 64 stored as Bs[n][k].
    if (TB) {
      const int r = tid >> 2, c = (tid & 3) * 8;
      const bf16x8 v = *reinterpret_cast<const bf16x8*>(B + (size_t)(n0 + r) * ldb + k0 + c);
      *reinterpret_cast<bf16x8*>(&Bs[r][c]) = v;
    } else {
      const int kk = tid >> 3, nn = (tid & 7) * 8;
      const bf16x8 v = *reinterpret_cast<const bf16x8*>(B + (size_t)(k0 + kk) * ldb + n0 + nn);
      for (int e = 0; e < 8; ++e) Bs[nn + e][kk] = v[e];
    }
    __syncthreads();
    const int kq = (lane >> 4) * 8, rq = lane & 15;
    bf16x8 af[2], bfr[2];
    for (int i = 0; i < 2; ++i) af[i] = *reinterpret_cast<const bf16x8*>(&As[wm + 16 * i + rq][kq]);
    for (int j = 0; j < 2; ++j) bfr[j] = *reinterpret_cast<const bf16x8*>(&Bs[wn + 16 * j + rq][kq]);
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
    if (do_colsum) {
      const bf16x8 v = *reinterpret_cast<const bf16x8*>(&Bs[lane][wv * 8]);
      for (int e = 0; e < 8; ++e) csum += bf2f(v[e]);
    }
    __syncthreads();
  }
  if (do_colsum) {
    red[wv][lane] = csum;
    __syncthreads();
    if (wv == 0) colsum[n0 + lane] = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
  }
  // D layout: lane holds column (lane % 16), rows 4 (lane / 16) + r of each tile.
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + 16 * i + 4 * (lane >> 4) + r;
        const int n = n0 + wn + 16 * j + (lane & 15);
        const float v = acc[i][j][r];
        if (EPI == EPI_FWD) {
          const float z = v + bias[n];
          Zout[(size_t)m * ldc + n] = z;
          Yout[(size_t)m * ldc + n] = f2bf(gelu(z));
        } else if (EPI == EPI_DX) {
          Yout[(size_t)m * ldc + n] = f2bf(v * gelu_grad(Zprev[(size_t)m * ldc + n]));
        } else {
          Cout[(size_t)m * ldc + n] = v;
        }
      }
}

// Inference GEMM: Y = bf16(GELU(A . W + bias)) with W given transposed (Wt[n][k],
// row-major, ldb = K) so both operand tiles load as contiguous 16-byte rows.
// 128 x 128 tiles, K in steps of 32, 256 threads = 2 x 2 wavefronts of 64 x 64
// (4 x 4 MFMA 16x16x32 tiles each), double-buffered LDS fed by a register
// prefetch of the next K step.  Per output the MFMA sequence (ascending K steps of
// 32) and the epilogue are those of gemm_bf16<EPI_FWD>, so the two paths give the
// same bits.  Tile order is XCD-aware: consecutive workgroups land on the 8 XCDs in
// turn, so each XCD is given whole rows of N tiles and reuses its A strip from L2.
// M a multiple of 128, N of 128, K of 32 (checked on the host).
constexpr int BM = 128, BN = 128, BK = 32, BLDK = 40;

__global__ __launch_bounds__(256) void gemm_fwd_big(int M, int N, int K, const __bf16* __restrict__ A, int lda,
                                                    const __bf16* __restrict__ Wt, int ldb,
                                                    const float* __restrict__ bias, __bf16* __restrict__ Y,
                                                    int ldy) {
  __shared__ __bf16 As[2][BM][BLDK];
  __shared__ __bf16 Bs[2][BN][BLDK];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int nN = N / BN, nMt = M / BM;
  int tile = blockIdx.x;
  if (nMt % 8 == 0) {
    const int xcd = blockIdx.x & 7, j = blockIdx.x >> 3;
    tile = ((j / nN) * 8 + xcd) * nN + j % nN;
  }
  const int m0 = (tile / nN) * BM, n0 = (tile % nN) * BN;
  const int wm = (wv >> 1) * 64, wn = (wv & 1) * 64;
  f32x4 acc[4][4];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  const int lr = tid >> 2, lc = (tid & 3) * 8;
  const __bf16* a0 = A + (size_t)(m0 + lr) * lda + lc;
  const __bf16* a1 = A + (size_t)(m0 + lr + 64) * lda + lc;
  const __bf16* b0 = Wt + (size_t)(n0 + lr) * ldb + lc;
  const __bf16* b1 = Wt + (size_t)(n0 + lr + 64) * ldb + lc;
  bf16x8 ra0 = *reinterpret_cast<const bf16x8*>(a0), ra1 = *reinterpret_cast<const bf16x8*>(a1);
  bf16x8 rb0 = *reinterpret_cast<const bf16x8*>(b0), rb1 = *reinterpret_cast<const bf16x8*>(b1);
  *reinterpret_cast<bf16x8*>(&As[0][lr][lc]) = ra0;
  *reinterpret_cast<bf16x8*>(&As[0][lr + 64][lc]) = ra1;
  *reinterpret_cast<bf16x8*>(&Bs[0][lr][lc]) = rb0;
  *reinterpret_cast<bf16x8*>(&Bs[0][lr + 64][lc]) = rb1;
  __syncthreads();
  const int nk = K / BK;
  const int kq = (lane >> 4) * 8, rq = lane & 15;
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    const bool more = kt + 1 < nk;
    if (more) {
      const int ko = (kt + 1) * BK;
      ra0 = *reinterpret_cast<const bf16x8*>(a0 + ko);
      ra1 = *reinterpret_cast<const bf16x8*>(a1 + ko);
      rb0 = *reinterpret_cast<const bf16x8*>(b0 + ko);
      rb1 = *reinterpret_cast<const bf16x8*>(b1 + ko);
    }
    bf16x8 af[4], bfr[4];
    for (int i = 0; i < 4; ++i) af[i] = *reinterpret_cast<const bf16x8*>(&As[cur][wm + 16 * i + rq][kq]);
    for (int j = 0; j < 4; ++j) bfr[j] = *reinterpret_cast<const bf16x8*>(&Bs[cur][wn + 16 * j + rq][kq]);
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
    if (more) {
      *reinterpret_cast<bf16x8*>(&As[cur ^ 1][lr][lc]) = ra0;
      *reinterpret_cast<bf16x8*>(&As[cur ^ 1][lr + 64][lc]) = ra1;
      *reinterpret_cast<bf16x8*>(&Bs[cur ^ 1][lr][lc]) = rb0;
      *reinterpret_cast<bf16x8*>(&Bs[cur ^ 1][lr + 64][lc]) = rb1;
    }
    __syncthreads();
  }
  for (int j = 0; j < 4; ++j) {
    const int n = n0 + wn + 16 * j + (lane & 15);
    const float bn = bias[n];
    for (int i = 0; i < 4; ++i)
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + 16 * i + 4 * (lane >> 4) + r;
        Y[(size_t)m * ldy + n] = f2bf(gelu(acc[i][j][r] + bn));
      }
  }
}

// Layer-fused inference for width 512: one workgroup (8 wavefronts) per strip of
// ROWS = 64 or 128 rows keeps the strip's activations in LDS (ROWS x 520 bf16:
// 65 or 130 KB, so two or one workgroups per CU) through every hidden layer and
// the output layer, so activations never touch
// HBM; only the input strip is read and one float per row written.  Per layer,
// wavefront w computes features [64w, 64w + 64) of all ROWS rows as the
// transposed product Y^T = W^T X^T: A operands are rows of the transposed
// weights W^T (streamed from L2, each wavefront's next K step prefetched into
// registers), B operands are activation rows read from LDS, and each lane's
// accumulator holds four consecutive features of one row, written back as one
// 8-byte LDS store after a barrier (every wavefront has then read the layer's
// input).  Per output the K steps of 32 run in ascending order and the epilogue
// bf16(GELU(acc + b)) and the output-layer dot product (per lane strided, then a
// butterfly) are those of gemm_fwd_big / out_fwd_kernel.
constexpr int FH = 512, FLD = FH + 8, FMAXL = 16;
// Strip rows of the fused inference kernel: 64 (two workgroups per CU) or 128 (one).
#ifndef DMW_FUSED_ROWS
#define DMW_FUSED_ROWS 128
#endif
static_assert(DMW_FUSED_ROWS == 0 || DMW_FUSED_ROWS == 64 || DMW_FUSED_ROWS == 128, "fused rows: 0, 64 or 128");

struct FusedArgs {
  int depth, k0;
  const __bf16* X0;              // (B, k0) input rows
  const __bf16* Wt;              // transposed bf16 weights: layer l at Wt + w_off[l], (FH, K_l) row-major
  const float* P;                // fp32 parameters (biases, output layer)
  long long w_off[FMAXL], b_off[FMAXL];
  long long wout_off, bout_off;
  float* y;
};

template <int ROWS>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(ROWS == 64 ? 4 : 2)))
void mlp_infer_fused(const FusedArgs a) {
  constexpr int RT = ROWS / 16;        // row tiles of 16
  __shared__ __attribute__((aligned(16))) __bf16 X[ROWS][FLD];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const long long r0 = (long long)blockIdx.x * ROWS;
  {
    const int cpr = a.k0 / 8;    // 16-byte chunks per input row
    for (int c = tid; c < ROWS * cpr; c += 512) {
      const int r = c / cpr, q = (c - r * cpr) * 8;
      *reinterpret_cast<bf16x8*>(&X[r][q]) = *reinterpret_cast<const bf16x8*>(a.X0 + (r0 + r) * a.k0 + q);
    }
  }
  __syncthreads();
  const int rq = lane & 15, kq = (lane >> 4) * 8;
  const int f0 = wv * 64;
  for (int l = 0; l < a.depth; ++l) {
    const int K = l == 0 ? a.k0 : FH;
    const __bf16* W = a.Wt + a.w_off[l];
    f32x4 acc[4][RT];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < RT; ++j) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const __bf16* wp[4];
    bf16x8 wf[4], wn[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      wp[i] = W + (size_t)(f0 + 16 * i + rq) * K + kq;
      wf[i] = *reinterpret_cast<const bf16x8*>(wp[i]);
    }
    for (int k = 0; k < K; k += 32) {
      const bool more = k + 32 < K;
      if (more) {
#pragma unroll
        for (int i = 0; i < 4; ++i) wn[i] = *reinterpret_cast<const bf16x8*>(wp[i] + k + 32);
      }
      bf16x8 xf[RT];
#pragma unroll
      for (int j = 0; j < RT; ++j) xf[j] = *reinterpret_cast<const bf16x8*>(&X[16 * j + rq][k + kq]);
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < RT; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf[i], xf[j], acc[i][j], 0, 0, 0);
      if (more) {
#pragma unroll
        for (int i = 0; i < 4; ++i) wf[i] = wn[i];
      }
    }
    __syncthreads();   // every wavefront has read this layer's input
    const float* bias = a.P + a.b_off[l];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int f = f0 + 16 * i + 4 * (lane >> 4);
      const float b0 = bias[f], b1 = bias[f + 1], b2 = bias[f + 2], b3 = bias[f + 3];
#pragma unroll
      for (int j = 0; j < RT; ++j) {
        typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
        const bf16x4 v = {f2bf(gelu(acc[i][j][0] + b0)), f2bf(gelu(acc[i][j][1] + b1)),
                          f2bf(gelu(acc[i][j][2] + b2)), f2bf(gelu(acc[i][j][3] + b3))};
        *reinterpret_cast<bf16x4*>(&X[16 * j + rq][f]) = v;
      }
    }
    __syncthreads();
  }
  // output layer: wavefront w takes rows 16w .. 16w + 15
  float w8[FH / 64];
#pragma unroll
  for (int q = 0; q < FH / 64; ++q) w8[q] = a.P[a.wout_off + lane + 64 * q];
  const float bo = a.P[a.bout_off];
  for (int rr = 0; rr < ROWS / 8; ++rr) {
    const int row = wv * (ROWS / 8) + rr;
    float s = 0.0f;
#pragma unroll
    for (int q = 0; q < FH / 64; ++q) s += bf2f(X[row][lane + 64 * q]) * w8[q];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) a.y[r0 + row] = s + bo;
  }
}

// y[b] = Y[b] . w + b0; with targets also the squared error and dy = 2 (y - t) / B.
__global__ void out_fwd_kernel(int B, int H, const __bf16* __restrict__ Y, const float* __restrict__ w,
                               const float* __restrict__ b0, const float* __restrict__ t,
                               float* __restrict__ y, float* __restrict__ dy, float* __restrict__ sq) {
  const int b = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (b >= B) return;
  float s = 0.0f;
  for (int i = lane; i < H; i += 64) s += bf2f(Y[(size_t)b * H + i]) * w[i];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) {
    const float v = s + b0[0];
    y[b] = v;
    if (t) {
      const float e = v - t[b];
      dy[b] = 2.0f * e / (float)B;
      sq[b] = e * e;
    }
  }
}

// Output layer backward: dZ_last[b][i] = dy[b] w[i] GELU'(Z_last[b][i]).
__global__ void out_bwd_dz_kernel(int B, int H, const float* __restrict__ dy, const float* __restrict__ w,
                                  const float* __restrict__ Z, __bf16* __restrict__ dZ) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)B * H) return;
  const int i = (int)(idx % H);
  const int b = (int)(idx / H);
  dZ[idx] = f2bf(dy[b] * w[i] * gelu_grad(Z[idx]));
}

// Output-layer weight gradient dw[i] = sum_b Y[b][i] dy[b] in two deterministic
// stages.  Stage 1: block (column tile x, row chunk y of 64 rows), thread
// (i = tid % 64, q = tid / 64) sums rows q, q+4, ... of the chunk; the four
// partial sums are combined in LDS into part[y][i].  Stage 2 adds the chunks in
// order and also forms db0 = sum_b dy[b] (thread H).
// Blocks of the first column tile also write the chunk's sum of dy to
// part[n_chunks * H + chunk] for the bias.
__global__ __launch_bounds__(256) void out_bwd_w_part_kernel(int B, int H, const __bf16* __restrict__ Y,
                                                             const float* __restrict__ dy,
                                                             float* __restrict__ part) {
  __shared__ float red[4][64];
  const int i = blockIdx.x * 64 + (threadIdx.x & 63), q = threadIdx.x >> 6;
  const int r0 = blockIdx.y * 64;
  float s = 0.0f;
  for (int r = r0 + q; r < r0 + 64 && r < B; r += 4) s += bf2f(Y[(size_t)r * H + i]) * dy[r];
  red[q][threadIdx.x & 63] = s;
  __syncthreads();
  if (q == 0) {
    const int c = threadIdx.x & 63;
    part[(size_t)blockIdx.y * H + i] = (red[0][c] + red[1][c]) + (red[2][c] + red[3][c]);
  }
  if (blockIdx.x == 0 && threadIdx.x < 64) {
    const int r = r0 + threadIdx.x;
    float v = r < B ? dy[r] : 0.0f;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (threadIdx.x == 0) part[(size_t)gridDim.y * H + blockIdx.y] = v;
  }
}

__global__ void out_bwd_w_sum_kernel(int B, int H, int n_chunks, const float* __restrict__ part,
                                     const float* __restrict__ dy, float* __restrict__ dw,
                                     float* __restrict__ db0) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > H) return;
  float s = 0.0f;
  if (i < H) {
    for (int c = 0; c < n_chunks; ++c) s += part[(size_t)c * H + i];
    dw[i] = s;
  } else {
    for (int c = 0; c < n_chunks; ++c) s += part[(size_t)n_chunks * H + c];
    db0[0] = s;
  }
}

// Adam with bias correction; refreshes the bf16 weight copy of matrix entries.
__global__ void adam_kernel(long long n, float* __restrict__ p, const float* __restrict__ g,
                            float* __restrict__ m1, float* __restrict__ m2, float lr, float b1, float b2,
                            float eps, float c1, float c2, const int* __restrict__ bf_index,
                            __bf16* __restrict__ pbf, const int* __restrict__ bfT_index,
                            __bf16* __restrict__ pbfT) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.0f - b1) * gi;
  const float v = b2 * m2[i] + (1.0f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = v;
  const float upd = lr * (a / c1) / (sqrtf(v / c2) + eps);
  const float np = p[i] - upd;
  p[i] = np;
  const int j = bf_index[i];
  if (j >= 0) {
    pbf[j] = f2bf(np);
    pbfT[bfT_index[i]] = f2bf(np);
  }
}

__global__ void to_bf16_kernel(long long n, const float* __restrict__ p, const int* __restrict__ bf_index,
                               __bf16* __restrict__ pbf, const int* __restrict__ bfT_index,
                               __bf16* __restrict__ pbfT) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int j = bf_index[i];
  if (j >= 0) {
    pbf[j] = f2bf(p[i]);
    pbfT[bfT_index[i]] = f2bf(p[i]);
  }
}

// Gather a training batch: rows idx[b] of the resident dataset.
__global__ void gather_kernel(int B, int K, const int* __restrict__ idx, const __bf16* __restrict__ X,
                              const float* __restrict__ T, __bf16* __restrict__ Xb, float* __restrict__ Tb) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)B * K) return;
  const int b = (int)(e / K), k = (int)(e % K);
  const int r = idx[b];
  Xb[e] = X[(size_t)r * K + k];
  if (k == 0) Tb[b] = T[r];
}

__global__ void f32_to_bf16_kernel(long long n, const float* __restrict__ a, __bf16* __restrict__ b) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = f2bf(a[i]);
}

}  // namespace dmw

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
struct dmw_mlp {
  int n_in = 0, k0 = 0, H = 0, depth = 0, max_batch = 0;
  int fused_rows = DMW_FUSED_ROWS;  // rows per strip of the layer-fused inference kernel (0: per-layer kernels)
  long long n_params = 0;          // internal (padded) parameter count
  std::vector<long long> w_off, b_off;  // offsets of W_l, b_l in the flat fp32 vector
  long long wout_off = 0, bout_off = 0;
  std::vector<long long> wbf_off;  // offsets of W_l in the bf16 copy
  long long n_bf = 0;
  hipStream_t st = nullptr;
  float *p = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;
  int* bf_index = nullptr;
  __bf16* pbf = nullptr;
  int* bfT_index = nullptr;  // W_l element (k, n) -> its place in the transposed copy (n, k)
  __bf16* pbfT = nullptr;    // bf16 W_l^T, the inference GEMM's B operand
  std::vector<__bf16*> Y;   // Y[0] = input batch (B x k0), Y[l+1] = GELU(Z_l) (B x H)
  std::vector<float*> Z;    // Z[l] (B x H)
  __bf16 *dZa = nullptr, *dZb = nullptr;
  float *out = nullptr, *dy = nullptr, *sq = nullptr, *tb = nullptr;
  float* part = nullptr;           // (max_batch / 64) x (H + 1) output-layer partial sums
  // resident dataset
  __bf16* dX = nullptr;
  float* dT = nullptr;
  long long n_data = 0;
  int* dIdx = nullptr;
  long long idx_cap = 0;
  int adam_t = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  std::vector<void*> allocs;
  ~dmw_mlp() {
    for (void* a : allocs) (void)hipFree(a);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
  }
  template <class T>
  hipError_t alloc(T** ptr, size_t count) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, count * sizeof(T));
    if (e == hipSuccess) {
      allocs.push_back(q);
      *ptr = reinterpret_cast<T*>(q);
    }
    return e;
  }
};

namespace {

#define MLP_TRY(x, what)                                   \
  do {                                                     \
    hipError_t e_ = (x);                                   \
    if (e_ != hipSuccess) return dmw::batch_hip_fail(e_, what); \
  } while (0)

template <int EPI, bool TA, bool TB>
hipError_t gemm(hipStream_t st, int M, int N, int K, const __bf16* A, int lda, const __bf16* B, int ldb,
                const float* bias, float* Z, __bf16* Y, const float* Zprev, float* C, int ldc,
                float* colsum = nullptr) {
  hipLaunchKernelGGL((dmw::gemm_bf16<EPI, TA, TB>), dim3(N / dmw::TN, M / dmw::TM), dim3(256), 0, st, M, N,
                     K, A, lda, B, ldb, bias, Z, Y, Zprev, C, ldc, colsum);
  return hipGetLastError();
}

unsigned blocks(long long n, int t) { return (unsigned)((n + t - 1) / t); }

// Forward pass of a batch already in m->Y[0]; logits in m->out; with targets
// also dy and squared errors.  Without targets (inference) the pre-activations
// are not kept, and batches of 128-row tiles take the large-tile GEMM.
int forward(dmw_mlp* m, int B, const float* targets_dev) {
  hipStream_t st = m->st;
  const int fr = m->fused_rows;
  if (targets_dev == nullptr && fr > 0 && B % fr == 0 && m->H == dmw::FH && m->depth <= dmw::FMAXL &&
      m->k0 <= dmw::FH) {
    dmw::FusedArgs fa{};
    fa.depth = m->depth;
    fa.k0 = m->k0;
    fa.X0 = m->Y[0];
    fa.Wt = m->pbfT;
    fa.P = m->p;
    for (int l = 0; l < m->depth; ++l) {
      fa.w_off[l] = m->wbf_off[l];
      fa.b_off[l] = m->b_off[l];
    }
    fa.wout_off = m->wout_off;
    fa.bout_off = m->bout_off;
    fa.y = m->out;
    if (fr == 64)
      hipLaunchKernelGGL(dmw::mlp_infer_fused<64>, dim3(B / 64), dim3(512), 0, st, fa);
    else
      hipLaunchKernelGGL(dmw::mlp_infer_fused<128>, dim3(B / 128), dim3(512), 0, st, fa);
    MLP_TRY(hipGetLastError(), "mlp fused inference");
    return DMW_OK;
  }
  const bool big = targets_dev == nullptr && B % dmw::BM == 0 && m->H % dmw::BN == 0;
  for (int l = 0; l < m->depth; ++l) {
    const int K = l == 0 ? m->k0 : m->H;
    if (big) {
      const unsigned n_tiles = (unsigned)((B / dmw::BM) * (m->H / dmw::BN));
      hipLaunchKernelGGL(dmw::gemm_fwd_big, dim3(n_tiles), dim3(256), 0, st, B, m->H, K, m->Y[l], K,
                         m->pbfT + m->wbf_off[l], K, m->p + m->b_off[l], m->Y[l + 1], m->H);
      MLP_TRY(hipGetLastError(), "mlp forward gemm");
      continue;
    }
    MLP_TRY((gemm<dmw::EPI_FWD, false, false>(st, B, m->H, K, m->Y[l], K, m->pbf + m->wbf_off[l], m->H,
                                              m->p + m->b_off[l], m->Z[l], m->Y[l + 1], nullptr, nullptr,
                                              m->H)),
            "mlp forward gemm");
  }
  hipLaunchKernelGGL(dmw::out_fwd_kernel, dim3(blocks(B, 4)), dim3(256), 0, st, B, m->H, m->Y[m->depth],
                     m->p + m->wout_off, m->p + m->bout_off, targets_dev, m->out, m->dy, m->sq);
  MLP_TRY(hipGetLastError(), "mlp output");
  return DMW_OK;
}

int backward(dmw_mlp* m, int B) {
  hipStream_t st = m->st;
  const int H = m->H, D = m->depth;
  const int n_chunks = (B + 63) / 64;
  hipLaunchKernelGGL(dmw::out_bwd_w_part_kernel, dim3(H / 64, n_chunks), dim3(256), 0, st, B, H, m->Y[D], m->dy,
                     m->part);
  hipLaunchKernelGGL(dmw::out_bwd_w_sum_kernel, dim3(blocks(H + 1, 256)), dim3(256), 0, st, B, H, n_chunks,
                     m->part, m->dy, m->g + m->wout_off, m->g + m->bout_off);
  hipLaunchKernelGGL(dmw::out_bwd_dz_kernel, dim3(blocks((long long)B * H, 256)), dim3(256), 0, st, B, H,
                     m->dy, m->p + m->wout_off, m->Z[D - 1], m->dZa);
  MLP_TRY(hipGetLastError(), "mlp output backward");
  __bf16* dz = m->dZa;
  __bf16* dz_next = m->dZb;
  for (int l = D - 1; l >= 0; --l) {
    const int K = l == 0 ? m->k0 : H;
    // dW_l = Y_l^T dZ_l  (K x H); db_l = colsum dZ_l from the same LDS tiles
    MLP_TRY((gemm<dmw::EPI_DW, true, false>(st, K, H, B, m->Y[l], K, dz, H, nullptr, nullptr, nullptr, nullptr,
                                            m->g + m->w_off[l], H, m->g + m->b_off[l])),
            "mlp dW gemm");
    if (l > 0) {
      // dZ_{l-1} = (dZ_l W_l^T) * GELU'(Z_{l-1})
      MLP_TRY((gemm<dmw::EPI_DX, false, true>(st, B, H, H, dz, H, m->pbf + m->wbf_off[l], H, nullptr, nullptr,
                                              dz_next, m->Z[l - 1], nullptr, H)),
              "mlp dX gemm");
      std::swap(dz, dz_next);
    }
  }
  return DMW_OK;
}

int upload_batch(dmw_mlp* m, const float* x, int B) {
  // host fp32 (B, n_in) -> device bf16 (B, k0) with zero padding
  std::vector<float> pad((size_t)B * m->k0, 0.0f);
  for (int b = 0; b < B; ++b) std::memcpy(&pad[(size_t)b * m->k0], x + (size_t)b * m->n_in, m->n_in * 4);
  float* tmp = reinterpret_cast<float*>(m->Z[0]);  // scratch large enough (B x H >= B x k0)
  MLP_TRY(hipStreamSynchronize(m->st), "mlp sync");
  MLP_TRY(hipMemcpy(tmp, pad.data(), pad.size() * 4, hipMemcpyHostToDevice), "H2D");
  hipLaunchKernelGGL(dmw::f32_to_bf16_kernel, dim3(blocks((long long)pad.size(), 256)), dim3(256), 0, m->st,
                     (long long)pad.size(), tmp, m->Y[0]);
  MLP_TRY(hipGetLastError(), "mlp input convert");
  return DMW_OK;
}

int check_batch(dmw_mlp* m, int B) {
  if (B < 64 || B % 64 != 0 || B > m->max_batch)
    return dmw::batch_fail(DMW_EINVAL, "batch must be a multiple of 64 and at most max_batch");
  return DMW_OK;
}

}  // namespace

extern "C" {

int dmw_mlp_create(dmw_mlp** out, int32_t n_in, int32_t width, int32_t depth, int32_t max_batch) {
  if (!out || n_in < 1 || width < 64 || width % 64 != 0 || depth < 1 || max_batch < 64 || max_batch % 64 != 0)
    return dmw::batch_fail(DMW_EINVAL, "mlp: n_in >= 1, width a multiple of 64, depth >= 1, max_batch a multiple of 64");
  *out = nullptr;
  dmw_mlp* m = new dmw_mlp();
  m->n_in = n_in;
  m->k0 = ((n_in + 63) / 64) * 64;
  m->H = width;
  m->depth = depth;
  m->max_batch = max_batch;
  long long off = 0, boff = 0;
  for (int l = 0; l < depth; ++l) {
    const int K = l == 0 ? m->k0 : width;
    m->w_off.push_back(off);
    m->wbf_off.push_back(boff);
    off += (long long)K * width;
    boff += (long long)K * width;
    m->b_off.push_back(off);
    off += width;
  }
  m->wout_off = off;
  off += width;
  m->bout_off = off;
  off += 1;
  m->n_params = off;
  m->n_bf = boff;
  auto fail = [&](hipError_t e, const char* what) {
    delete m;
    return dmw::batch_hip_fail(e, what);
  };
#define MC(x, w)                       \
  do {                                 \
    hipError_t e_ = (x);               \
    if (e_ != hipSuccess) return fail(e_, w); \
  } while (0)
  MC(hipStreamCreateWithFlags(&m->st, hipStreamNonBlocking), "hipStreamCreate");
  MC(m->alloc(&m->p, m->n_params), "hipMalloc");
  MC(m->alloc(&m->g, m->n_params), "hipMalloc");
  MC(m->alloc(&m->m1, m->n_params), "hipMalloc");
  MC(m->alloc(&m->m2, m->n_params), "hipMalloc");
  MC(m->alloc(&m->bf_index, m->n_params), "hipMalloc");
  MC(m->alloc(&m->pbf, m->n_bf), "hipMalloc");
  MC(m->alloc(&m->bfT_index, m->n_params), "hipMalloc");
  MC(m->alloc(&m->pbfT, m->n_bf), "hipMalloc");
  m->Y.resize(depth + 1);
  m->Z.resize(depth);
  MC(m->alloc(&m->Y[0], (size_t)max_batch * m->k0), "hipMalloc");
  for (int l = 0; l < depth; ++l) {
    MC(m->alloc(&m->Y[l + 1], (size_t)max_batch * width), "hipMalloc");
    MC(m->alloc(&m->Z[l], (size_t)max_batch * std::max(width, m->k0)), "hipMalloc");
  }
  MC(m->alloc(&m->dZa, (size_t)max_batch * width), "hipMalloc");
  MC(m->alloc(&m->dZb, (size_t)max_batch * width), "hipMalloc");
  MC(m->alloc(&m->out, max_batch), "hipMalloc");
  MC(m->alloc(&m->dy, max_batch), "hipMalloc");
  MC(m->alloc(&m->sq, max_batch), "hipMalloc");
  MC(m->alloc(&m->tb, max_batch), "hipMalloc");
  MC(m->alloc(&m->part, (size_t)(max_batch / 64) * (width + 1)), "hipMalloc");
  MC(hipEventCreate(&m->e0), "hipEventCreate");
  MC(hipEventCreate(&m->e1), "hipEventCreate");
  // bf16 index map: weight matrices map to their bf16 copy, biases to -1
  std::vector<int> idx(m->n_params, -1);
  for (int l = 0; l < depth; ++l) {
    const int K = l == 0 ? m->k0 : width;
    for (long long i = 0; i < (long long)K * width; ++i) idx[m->w_off[l] + i] = (int)(m->wbf_off[l] + i);
  }
  MC(hipMemcpy(m->bf_index, idx.data(), idx.size() * 4, hipMemcpyHostToDevice), "H2D");
  std::vector<int> idxT(m->n_params, -1);
  for (int l = 0; l < depth; ++l) {
    const int K = l == 0 ? m->k0 : width;
    for (int k = 0; k < K; ++k)
      for (int n = 0; n < width; ++n)
        idxT[m->w_off[l] + (long long)k * width + n] = (int)(m->wbf_off[l] + (long long)n * K + k);
  }
  MC(hipMemcpy(m->bfT_index, idxT.data(), idxT.size() * 4, hipMemcpyHostToDevice), "H2D");
  std::vector<float> zeros(m->n_params, 0.0f);
  MC(hipMemcpy(m->p, zeros.data(), zeros.size() * 4, hipMemcpyHostToDevice), "H2D");
  MC(hipMemcpy(m->m1, zeros.data(), zeros.size() * 4, hipMemcpyHostToDevice), "H2D");
  MC(hipMemcpy(m->m2, zeros.data(), zeros.size() * 4, hipMemcpyHostToDevice), "H2D");
#undef MC
  *out = m;
  return DMW_OK;
}

int dmw_mlp_destroy(dmw_mlp* m) {
  delete m;
  return DMW_OK;
}

int64_t dmw_mlp_num_params(dmw_mlp* m) {
  // public (unpadded) layout: W_0 is (n_in x H)
  return m ? m->n_params - (long long)(m->k0 - m->n_in) * m->H : 0;
}

// Public flat layout: for each layer W_l (K_l x H, row-major; K_0 = n_in), b_l (H);
// then w_out (H), b_out (1).
static void map_public(dmw_mlp* m, std::vector<long long>& pos) {
  pos.clear();
  for (int l = 0; l < m->depth; ++l) {
    const int K = l == 0 ? m->n_in : m->H;
    for (long long i = 0; i < (long long)K * m->H; ++i) pos.push_back(m->w_off[l] + i);
    for (int j = 0; j < m->H; ++j) pos.push_back(m->b_off[l] + j);
  }
  for (int j = 0; j < m->H; ++j) pos.push_back(m->wout_off + j);
  pos.push_back(m->bout_off);
}

int dmw_mlp_set_params(dmw_mlp* m, const float* params, int64_t n) {
  if (!m || !params || n != dmw_mlp_num_params(m)) return dmw::batch_fail(DMW_EINVAL, "mlp: parameter count mismatch");
  std::vector<long long> pos;
  map_public(m, pos);
  std::vector<float> full(m->n_params, 0.0f);
  for (size_t i = 0; i < pos.size(); ++i) full[pos[i]] = params[i];
  MLP_TRY(hipMemcpy(m->p, full.data(), full.size() * 4, hipMemcpyHostToDevice), "H2D");
  std::vector<float> zeros(m->n_params, 0.0f);
  MLP_TRY(hipMemcpy(m->m1, zeros.data(), zeros.size() * 4, hipMemcpyHostToDevice), "H2D");
  MLP_TRY(hipMemcpy(m->m2, zeros.data(), zeros.size() * 4, hipMemcpyHostToDevice), "H2D");
  m->adam_t = 0;
  hipLaunchKernelGGL(dmw::to_bf16_kernel, dim3(blocks(m->n_params, 256)), dim3(256), 0, m->st, m->n_params, m->p,
                     m->bf_index, m->pbf, m->bfT_index, m->pbfT);
  MLP_TRY(hipGetLastError(), "to_bf16");
  MLP_TRY(hipStreamSynchronize(m->st), "mlp set params");
  return DMW_OK;
}

static int copy_public(dmw_mlp* m, const float* dev, float* out) {
  std::vector<long long> pos;
  map_public(m, pos);
  std::vector<float> full(m->n_params);
  MLP_TRY(hipStreamSynchronize(m->st), "mlp sync");
  MLP_TRY(hipMemcpy(full.data(), dev, full.size() * 4, hipMemcpyDeviceToHost), "D2H");
  for (size_t i = 0; i < pos.size(); ++i) out[i] = full[pos[i]];
  return DMW_OK;
}

int dmw_mlp_get_params(dmw_mlp* m, float* params, int64_t n) {
  if (!m || !params || n != dmw_mlp_num_params(m)) return dmw::batch_fail(DMW_EINVAL, "mlp: parameter count mismatch");
  return copy_public(m, m->p, params);
}

int dmw_mlp_forward(dmw_mlp* m, const float* x, int32_t batch, float* y) {
  if (!m || !x || !y) return dmw::batch_fail(DMW_EINVAL, "null argument");
  int rc = check_batch(m, batch);
  if (rc) return rc;
  if ((rc = upload_batch(m, x, batch))) return rc;
  if ((rc = forward(m, batch, nullptr))) return rc;
  MLP_TRY(hipMemcpyAsync(y, m->out, (size_t)batch * 4, hipMemcpyDeviceToHost, m->st), "D2H");
  MLP_TRY(hipStreamSynchronize(m->st), "mlp forward");
  return DMW_OK;
}

int dmw_mlp_gradients(dmw_mlp* m, const float* x, const float* t, int32_t batch, float* grads, float* loss) {
  if (!m || !x || !t || !grads) return dmw::batch_fail(DMW_EINVAL, "null argument");
  int rc = check_batch(m, batch);
  if (rc) return rc;
  if ((rc = upload_batch(m, x, batch))) return rc;
  MLP_TRY(hipMemcpyAsync(m->tb, t, (size_t)batch * 4, hipMemcpyHostToDevice, m->st), "H2D");
  if ((rc = forward(m, batch, m->tb))) return rc;
  if ((rc = backward(m, batch))) return rc;
  if (loss) {
    std::vector<float> sq(batch);
    MLP_TRY(hipMemcpyAsync(sq.data(), m->sq, (size_t)batch * 4, hipMemcpyDeviceToHost, m->st), "D2H");
    MLP_TRY(hipStreamSynchronize(m->st), "mlp gradients");
    double s = 0;
    for (float v : sq) s += v;
    *loss = (float)(s / batch);
  }
  return copy_public(m, m->g, grads);
}

int dmw_mlp_set_data(dmw_mlp* m, const float* x, const float* t, int64_t n) {
  if (!m || !x || !t || n < 1) return dmw::batch_fail(DMW_EINVAL, "null argument");
  if (m->dX) {
    (void)hipFree(m->dX);
    (void)hipFree(m->dT);
    m->allocs.erase(std::remove(m->allocs.begin(), m->allocs.end(), (void*)m->dX), m->allocs.end());
    m->allocs.erase(std::remove(m->allocs.begin(), m->allocs.end(), (void*)m->dT), m->allocs.end());
    m->dX = nullptr;
    m->dT = nullptr;
  }
  MLP_TRY(m->alloc(&m->dX, (size_t)n * m->k0), "hipMalloc dataset");
  MLP_TRY(m->alloc(&m->dT, (size_t)n), "hipMalloc dataset");
  const long long chunk = 65536;
  std::vector<float> pad((size_t)chunk * m->k0);
  float* tmp = nullptr;
  MLP_TRY(hipMalloc(&tmp, pad.size() * 4), "hipMalloc");
  for (long long r0 = 0; r0 < n; r0 += chunk) {
    const long long rows = std::min(chunk, n - r0);
    std::fill(pad.begin(), pad.end(), 0.0f);
    for (long long r = 0; r < rows; ++r)
      std::memcpy(&pad[(size_t)r * m->k0], x + (size_t)(r0 + r) * m->n_in, m->n_in * 4);
    hipError_t e = hipMemcpy(tmp, pad.data(), (size_t)rows * m->k0 * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(dmw::f32_to_bf16_kernel, dim3(blocks(rows * m->k0, 256)), dim3(256), 0, m->st,
                         rows * m->k0, tmp, m->dX + (size_t)r0 * m->k0);
      e = hipStreamSynchronize(m->st);
    }
    if (e != hipSuccess) {
      (void)hipFree(tmp);
      return dmw::batch_hip_fail(e, "dataset upload");
    }
  }
  (void)hipFree(tmp);
  MLP_TRY(hipMemcpy(m->dT, t, (size_t)n * 4, hipMemcpyHostToDevice), "H2D");
  m->n_data = n;
  return DMW_OK;
}

int dmw_mlp_train(dmw_mlp* m, int32_t batch, int32_t n_steps, const int32_t* indices, const float* lr,
                  float beta1, float beta2, float eps, float* losses, double* train_ms) {
  if (!m || !indices || !lr || !m->dX) return dmw::batch_fail(DMW_EINVAL, "mlp: dataset not set or null argument");
  int rc = check_batch(m, batch);
  if (rc) return rc;
  const long long need = (long long)batch * n_steps;
  for (long long i = 0; i < need; ++i)
    if (indices[i] < 0 || indices[i] >= m->n_data) return dmw::batch_fail(DMW_EINVAL, "mlp: index out of range");
  if (need > m->idx_cap) {
    if (m->dIdx) {
      (void)hipFree(m->dIdx);
      m->allocs.erase(std::remove(m->allocs.begin(), m->allocs.end(), (void*)m->dIdx), m->allocs.end());
      m->dIdx = nullptr;
    }
    MLP_TRY(m->alloc(&m->dIdx, (size_t)need), "hipMalloc indices");
    m->idx_cap = need;
  }
  MLP_TRY(hipMemcpyAsync(m->dIdx, indices, (size_t)need * 4, hipMemcpyHostToDevice, m->st), "H2D");
  std::vector<float> sq_host(batch);
  MLP_TRY(hipEventRecord(m->e0, m->st), "event");
  for (int s = 0; s < n_steps; ++s) {
    hipLaunchKernelGGL(dmw::gather_kernel, dim3(blocks((long long)batch * m->k0, 256)), dim3(256), 0, m->st, batch,
                       m->k0, m->dIdx + (size_t)s * batch, m->dX, m->dT, m->Y[0], m->tb);
    MLP_TRY(hipGetLastError(), "gather");
    if ((rc = forward(m, batch, m->tb))) return rc;
    if ((rc = backward(m, batch))) return rc;
    m->adam_t += 1;
    const float c1 = 1.0f - std::pow(beta1, (float)m->adam_t), c2 = 1.0f - std::pow(beta2, (float)m->adam_t);
    hipLaunchKernelGGL(dmw::adam_kernel, dim3(blocks(m->n_params, 256)), dim3(256), 0, m->st, m->n_params, m->p,
                       m->g, m->m1, m->m2, lr[s], beta1, beta2, eps, c1, c2, m->bf_index, m->pbf, m->bfT_index, m->pbfT);
    MLP_TRY(hipGetLastError(), "adam");
    if (losses) {
      MLP_TRY(hipMemcpyAsync(sq_host.data(), m->sq, (size_t)batch * 4, hipMemcpyDeviceToHost, m->st), "D2H");
      MLP_TRY(hipStreamSynchronize(m->st), "train step");
      double acc = 0;
      for (float v : sq_host) acc += v;
      losses[s] = (float)(acc / batch);
    }
  }
  MLP_TRY(hipEventRecord(m->e1, m->st), "event");
  MLP_TRY(hipStreamSynchronize(m->st), "train");
  if (train_ms) {
    float ms = 0;
    (void)hipEventElapsedTime(&ms, m->e0, m->e1);
    *train_ms = ms;
  }
  return DMW_OK;
}

int dmw_mlp_predict(dmw_mlp* m, int64_t n, float* y, double* kernel_ms) {
  if (!m || !y || !m->dX || n < 1 || n > m->n_data) return dmw::batch_fail(DMW_EINVAL, "mlp: dataset not set");
  const int B = m->max_batch;
  std::vector<int> idx(B);
  if ((long long)B > m->idx_cap) {
    if (m->dIdx) {
      (void)hipFree(m->dIdx);
      m->allocs.erase(std::remove(m->allocs.begin(), m->allocs.end(), (void*)m->dIdx), m->allocs.end());
      m->dIdx = nullptr;
    }
    MLP_TRY(m->alloc(&m->dIdx, (size_t)B), "hipMalloc indices");
    m->idx_cap = B;
  }
  double total = 0;
  for (long long r0 = 0; r0 < n; r0 += B) {
    const long long rows = std::min<long long>(B, n - r0);
    const int Bp = (int)(((rows + 63) / 64) * 64);
    for (int b = 0; b < Bp; ++b) idx[b] = (int)std::min<long long>(r0 + b, n - 1);
    MLP_TRY(hipMemcpyAsync(m->dIdx, idx.data(), (size_t)Bp * 4, hipMemcpyHostToDevice, m->st), "H2D");
    MLP_TRY(hipEventRecord(m->e0, m->st), "event");
    hipLaunchKernelGGL(dmw::gather_kernel, dim3(blocks((long long)Bp * m->k0, 256)), dim3(256), 0, m->st, Bp,
                       m->k0, m->dIdx, m->dX, m->dT, m->Y[0], m->tb);
    MLP_TRY(hipGetLastError(), "gather");
    int rc = forward(m, Bp, nullptr);
    if (rc) return rc;
    MLP_TRY(hipEventRecord(m->e1, m->st), "event");
    MLP_TRY(hipMemcpyAsync(y + r0, m->out, (size_t)rows * 4, hipMemcpyDeviceToHost, m->st), "D2H");
    MLP_TRY(hipStreamSynchronize(m->st), "predict");
    float ms = 0;
    (void)hipEventElapsedTime(&ms, m->e0, m->e1);
    total += ms;
  }
  if (kernel_ms) *kernel_ms = total;
  return DMW_OK;
}

}  // extern "C"
