// Learned initializer: BF16 MFMA GEMMs with FP32 accumulation, the elementwise pieces
// of an MLP's training step (bias gradient, Adam on FP32 master weights with a BF16
// copy), diffraction-pattern features and object-patch gather / stitch.  gfx950 only.
//
// GEMM: C[M, N] = op(A)[M, K] . op(B)[K, N], row-major storage with leading dimensions;
// op(A) is A (M x K, "N") or the transpose of a K x M array ("T"), op(B) is B (K x N,
// "N") or the transpose of an N x K array ("T").  A workgroup of 256 threads (4 waves,
// 2 x 2) computes a 128 x 128 tile; each wave a 64 x 64 block as 2 x 2
// v_mfma_f32_32x32x16_bf16 tiles.  K advances 32 per LDS stage; both operand tiles sit in
// LDS k-contiguous ([m][k] and [n][k]) so every lane reads its 8 consecutive k values as
// one 16-byte load.  M and N must be multiples of 128, K of 32.
#include "phx_common.h"
#include "../../include/phaser_mi355x.h"

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int BM = 128, BN = 128, BK = 32;
constexpr int LDK = BK + 8;  // LDS row stride (bf16): 80 bytes, staggers rows across banks

__device__ __forceinline__ __bf16 to_bf16(float v) { return (__bf16)v; }  // round to nearest even

template <bool TA, bool TB>
__global__ void __launch_bounds__(256) gemm_bf16_k(const __bf16* __restrict__ A, int64_t lda,
                                                   const __bf16* __restrict__ B, int64_t ldb, int K,
                                                   float* __restrict__ C, __bf16* __restrict__ Cb, int64_t ldc,
                                                   const float* __restrict__ bias, int relu,
                                                   const __bf16* __restrict__ mask, int64_t ldm, int accumulate) {
  __shared__ __align__(16) __bf16 sA[BM * LDK];
  __shared__ __align__(16) __bf16 sB[BN * LDK];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
  const int64_t m0 = (int64_t)blockIdx.y * BM, n0 = (int64_t)blockIdx.x * BN;
  f32x16 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;

  for (int k0 = 0; k0 < K; k0 += BK) {
    __syncthreads();
    // ---- stage A tile (BM x BK) into sA[m][k]
    if (!TA) {
      // A row-major M x K: 8 consecutive k per 16-byte load; 512 loads over 256 threads
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const int id = tid + 256 * t, r = id >> 2, kc = (id & 3) * 8;
        *(bf16x8*)&sA[r * LDK + kc] = *(const bf16x8*)&A[(m0 + r) * lda + k0 + kc];
      }
    } else {
      // A stored K x M: 8 consecutive m per load, written transposed
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const int id = tid + 256 * t, kr = id >> 4, mc = (id & 15) * 8;
        const bf16x8 v = *(const bf16x8*)&A[(int64_t)(k0 + kr) * lda + m0 + mc];
#pragma unroll
        for (int e = 0; e < 8; ++e) sA[(mc + e) * LDK + kr] = v[e];
      }
    }
    // ---- stage B tile (BK x BN) into sB[n][k]
    if (TB) {
      // B stored N x K: k-contiguous
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const int id = tid + 256 * t, r = id >> 2, kc = (id & 3) * 8;
        *(bf16x8*)&sB[r * LDK + kc] = *(const bf16x8*)&B[(n0 + r) * ldb + k0 + kc];
      }
    } else {
      // B row-major K x N: 8 consecutive n per load, written transposed
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const int id = tid + 256 * t, kr = id >> 4, nc = (id & 15) * 8;
        const bf16x8 v = *(const bf16x8*)&B[(int64_t)(k0 + kr) * ldb + n0 + nc];
#pragma unroll
        for (int e = 0; e < 8; ++e) sB[(nc + e) * LDK + kr] = v[e];
      }
    }
    __syncthreads();
    // ---- two K=16 MFMA steps; lane l holds row (l % 32), k = 8 (l / 32) .. + 7
#pragma unroll
    for (int kk = 0; kk < BK; kk += 16) {
      const int kl = kk + 8 * (lane >> 5);
      bf16x8 a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) a[i] = *(const bf16x8*)&sA[(wm + 32 * i + (lane & 31)) * LDK + kl];
#pragma unroll
      for (int j = 0; j < 2; ++j) b[j] = *(const bf16x8*)&sB[(wn + 32 * j + (lane & 31)) * LDK + kl];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
  }
  // ---- epilogue: D element r of lane l is (row 8 (r/4) + 4 (l/32) + r%4, column l % 32)
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int64_t row = m0 + wm + 32 * i + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const int64_t col = n0 + wn + 32 * j + (lane & 31);
        float v = acc[i][j][r];
        if (accumulate) v += C[row * ldc + col];
        if (bias) v += bias[col];
        if (relu) v = fmaxf(v, 0.f);
        if (mask && !((float)mask[row * ldm + col] > 0.f)) v = 0.f;
        if (C) C[row * ldc + col] = v;
        if (Cb) Cb[row * ldc + col] = to_bf16(v);
      }
}

// NT GEMM, the form every training and inference GEMM of the network takes: C[M, N] =
// A[M, K] . B[N, K]^T with both operands k-contiguous (the network keeps a transposed
// BF16 copy of its weights and of the activations and output gradients the weight
// gradients need).  256 threads (4 waves, 2 x 2) per 128 x BN tile, each wave a 64 x BN/2
// block of 32 x 32 v_mfma_f32_32x32x16_bf16 tiles.  K advances 64 per stage.
//
// Pipeline.  A stage's MFMAs take a few hundred cycles per SIMD, far less than the
// latency of its operand loads, and a CU holds one workgroup for these grids, so the
// loads run D stages ahead: D register slots hold stages s+1 .. s+D while stage s is
// multiplied out of one of two LDS buffers, and stage s+1 goes from its slot to the
// other buffer after the MFMAs; one barrier per stage separates the buffers' reads and
// writes.  (The stage loop is unrolled by D so every slot index is static.)  Row stride
// 72 BF16 (144 bytes, 36 dwords): the 16-byte stores of a stage (8 lanes per 128-byte
// row) and the 16-byte MFMA operand reads (8 rows per half-wave group, 4 dwords each)
// both land on 32 distinct banks.  BN = 64 doubles the workgroups of a 1024 x 2048
// output (256, one per CU) at the cost of a smaller wave block.  The epilogue of
// gemm_bf16_k, plus an optional transposed BF16 copy Ct[n * ldct + m].
constexpr int NT_BM = 128, NT_BK = 64, NT_LDK = NT_BK + 8;

template <int BN, int D>
__global__ void __launch_bounds__(256) gemm_nt_k(const __bf16* __restrict__ A, int64_t lda,
                                                 const __bf16* __restrict__ B, int64_t ldb, int K,
                                                 float* __restrict__ C, __bf16* __restrict__ Cb, int64_t ldc,
                                                 __bf16* __restrict__ Ct, int64_t ldct, const float* __restrict__ bias,
                                                 int relu, const __bf16* __restrict__ mask, int64_t ldm,
                                                 int accumulate) {
  constexpr int WN = BN / 2, NJ = WN / 32;
  constexpr int A_LD = NT_BM * NT_BK / 8 / 256, B_LD = BN * NT_BK / 8 / 256;  // 16-byte loads per thread
  __shared__ __align__(16) __bf16 sA[2][NT_BM * NT_LDK];
  __shared__ __align__(16) __bf16 sB[2][BN * NT_LDK];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 64, wn = (wave & 1) * WN;
  const int64_t m0 = (int64_t)blockIdx.y * NT_BM, n0 = (int64_t)blockIdx.x * BN;
  f32x16 acc[2][NJ];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < NJ; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;

  // thread -> (row, 8-wide k chunk) of a stage: 8 threads per 64-wide row
  bf16x8 ra[D][A_LD], rb[D][B_LD];
  auto gload = [&](int slot, int k0) __attribute__((always_inline)) {
#pragma unroll
    for (int t = 0; t < A_LD; ++t) {
      const int id = tid + 256 * t, r = id >> 3, kc = (id & 7) * 8;
      ra[slot][t] = *(const bf16x8*)&A[(m0 + r) * lda + k0 + kc];
    }
#pragma unroll
    for (int t = 0; t < B_LD; ++t) {
      const int id = tid + 256 * t, r = id >> 3, kc = (id & 7) * 8;
      rb[slot][t] = *(const bf16x8*)&B[(n0 + r) * ldb + k0 + kc];
    }
  };
  auto sstore = [&](int slot, int buf) __attribute__((always_inline)) {
#pragma unroll
    for (int t = 0; t < A_LD; ++t) {
      const int id = tid + 256 * t, r = id >> 3, kc = (id & 7) * 8;
      *(bf16x8*)&sA[buf][r * NT_LDK + kc] = ra[slot][t];
    }
#pragma unroll
    for (int t = 0; t < B_LD; ++t) {
      const int id = tid + 256 * t, r = id >> 3, kc = (id & 7) * 8;
      *(bf16x8*)&sB[buf][r * NT_LDK + kc] = rb[slot][t];
    }
  };
  auto compute = [&](int buf) __attribute__((always_inline)) {
    const __bf16* a_s = sA[buf];
    const __bf16* b_s = sB[buf];
#pragma unroll
    for (int kk = 0; kk < NT_BK; kk += 16) {
      // lane l holds row (l % 32), k = kk + 8 (l / 32) .. + 7
      const int kl = kk + 8 * (lane >> 5);
      bf16x8 a[2], b[NJ];
#pragma unroll
      for (int i = 0; i < 2; ++i) a[i] = *(const bf16x8*)&a_s[(wm + 32 * i + (lane & 31)) * NT_LDK + kl];
#pragma unroll
      for (int j = 0; j < NJ; ++j) b[j] = *(const bf16x8*)&b_s[(wn + 32 * j + (lane & 31)) * NT_LDK + kl];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
  };

  // The loads and stores run unconditionally (stages past the end reload the last one and
  // are never multiplied): a branch around them would make the compiler wait for every
  // load in flight, newest included, before each store.
  const int nk = K / NT_BK;  // a multiple of D (the host picks D)
#pragma unroll
  for (int u = 0; u < D; ++u) gload(u, min(u, nk - 1) * NT_BK);
  sstore(0, 0);
  __syncthreads();
  for (int k0 = 0; k0 < nk; k0 += D) {
#pragma unroll
    for (int u = 0; u < D; ++u) {
      const int s = k0 + u;  // the stage multiplied now; slot u held it and is free again
      gload(u, min(s + D, nk - 1) * NT_BK);
      // buffer s & 1: u & 1 for even D (k0 a multiple of D)
      compute(D % 2 == 0 ? (u & 1) : (s & 1));
      sstore((u + 1) % D, D % 2 == 0 ? ((u + 1) & 1) : ((s + 1) & 1));
      __syncthreads();
    }
  }
  // epilogue: D element r of lane l is (row 8 (r/4) + 4 (l/32) + r%4, column l % 32).  Each
  // option is a uniform branch around a whole 16-element loop, so a tile's loads (C to
  // accumulate, the mask) are issued together rather than one element at a time; a lane's
  // bias is one value per column block.  Ct gets the lane's 4 consecutive rows of a
  // column as one 8-byte store.
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const int64_t col = n0 + wn + 32 * j + (lane & 31);
    const float bv = bias ? bias[col] : 0.f;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int64_t row0 = m0 + wm + 32 * i + 4 * (lane >> 5);
      auto row = [&](int r) { return row0 + 8 * (r >> 2) + (r & 3); };
      float v[16];
#pragma unroll
      for (int r = 0; r < 16; ++r) v[r] = acc[i][j][r];
      if (accumulate) {
        float c[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) c[r] = C[row(r) * ldc + col];
#pragma unroll
        for (int r = 0; r < 16; ++r) v[r] += c[r];
      }
#pragma unroll
      for (int r = 0; r < 16; ++r) v[r] += bv;
      if (relu) {
#pragma unroll
        for (int r = 0; r < 16; ++r) v[r] = fmaxf(v[r], 0.f);
      }
      if (mask) {
        __bf16 mk[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) mk[r] = mask[row(r) * ldm + col];
#pragma unroll
        for (int r = 0; r < 16; ++r) v[r] = ((float)mk[r] > 0.f) ? v[r] : 0.f;
      }
      if (C) {
#pragma unroll
        for (int r = 0; r < 16; ++r) C[row(r) * ldc + col] = v[r];
      }
      if (Cb) {
#pragma unroll
        for (int r = 0; r < 16; ++r) Cb[row(r) * ldc + col] = to_bf16(v[r]);
      }
      if (Ct) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          bf16x4 t;
#pragma unroll
          for (int e = 0; e < 4; ++e) t[e] = to_bf16(v[4 * q + e]);
          *(bf16x4*)&Ct[col * ldct + row(4 * q)] = t;
        }
      }
    }
  }
}

// out[c][r] = in[r][c] for a rows x cols BF16 array (leading dimensions ldi, ldo), 64 x 64
// tiles through LDS (row stride 66 elements: the column reads of a tile spread over the
// banks).  Full tiles of 16-byte-aligned rows move 8 elements per access; edge tiles go
// element by element.
__global__ void __launch_bounds__(256) transpose_bf16_k(const __bf16* __restrict__ in, int64_t ldi, int64_t rows,
                                                        int64_t cols, __bf16* __restrict__ out, int64_t ldo) {
  __shared__ __bf16 t[64 * 66];
  const int64_t r0 = (int64_t)blockIdx.y * 64, c0 = (int64_t)blockIdx.x * 64;
  const bool full = r0 + 64 <= rows && c0 + 64 <= cols && ldi % 8 == 0 && ldo % 8 == 0 &&
                    reinterpret_cast<uintptr_t>(in) % 16 == 0 && reinterpret_cast<uintptr_t>(out) % 16 == 0;
  if (full) {
    for (int v = threadIdx.x; v < 64 * 8; v += 256) {  // row y, elements 8 x8 .. + 7
      const int y = v >> 3, x8 = (v & 7) * 8;
      const bf16x8 a = *(const bf16x8*)&in[(r0 + y) * ldi + c0 + x8];
#pragma unroll
      for (int e = 0; e < 8; ++e) t[y * 66 + x8 + e] = a[e];
    }
    __syncthreads();
    for (int v = threadIdx.x; v < 64 * 8; v += 256) {  // out row c0 + y, columns r0 + x8 .. + 7
      const int y = v >> 3, x8 = (v & 7) * 8;
      bf16x8 a;
#pragma unroll
      for (int e = 0; e < 8; ++e) a[e] = t[(x8 + e) * 66 + y];
      *(bf16x8*)&out[(c0 + y) * ldo + r0 + x8] = a;
    }
    return;
  }
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int y = e >> 6, x = e & 63;
    if (r0 + y < rows && c0 + x < cols) t[y * 66 + x] = in[(r0 + y) * ldi + c0 + x];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int y = e >> 6, x = e & 63;  // out row c0 + y, column r0 + x
    if (c0 + y < cols && r0 + x < rows) out[(c0 + y) * ldo + r0 + x] = t[x * 66 + y];
  }
}

// out[c] = sum_r in[r, c] (FP32): a workgroup of 1024 threads owns 64 columns; thread
// (strip t, column x) sums rows t, t + 16, t + 32, .. in order (the 64 lanes of a wave
// read 64 consecutive columns of a row), then the 16 strip sums are added in strip order.
__global__ void __launch_bounds__(1024) colsum_k(const float* __restrict__ in, int64_t rows, int64_t cols, int64_t ld,
                                                 float* __restrict__ out) {
  __shared__ float part[16][64];
  const int x = threadIdx.x & 63, t = threadIdx.x >> 6;
  const int64_t c = (int64_t)blockIdx.x * 64 + x;
  float s = 0.f;
  if (c < cols) {
#pragma unroll 8
    for (int64_t r = t; r < rows; r += 16) s += in[r * ld + c];
  }
  part[t][x] = s;
  __syncthreads();
  if (t == 0 && c < cols) {
    float v = part[0][x];
    for (int k = 1; k < 16; ++k) v += part[k][x];
    out[c] = v;
  }
}

// Adam (Kingma & Ba) on FP32 master weights, BF16 copy refreshed: bias corrections c1 =
// 1/(1 - b1^t), c2 = 1/(1 - b2^t) are computed on the host in FP64.
__global__ void adam_k(float* __restrict__ w, __bf16* __restrict__ wb, const float* __restrict__ g,
                       float* __restrict__ mo, float* __restrict__ ve, int64_t n, float lr, float b1, float b2,
                       float eps, float c1, float c2, float weight_decay) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i] + weight_decay * w[i];
  const float m = b1 * mo[i] + (1.f - b1) * gi;
  const float v = b2 * ve[i] + (1.f - b2) * gi * gi;
  mo[i] = m;
  ve[i] = v;
  const float nw = w[i] - lr * (m * c1) / (sqrtf(v * c2) + eps);
  w[i] = nw;
  if (wb) wb[i] = to_bf16(nw);
}

__global__ void f32_to_bf16_k(const float* __restrict__ in, __bf16* __restrict__ out, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = to_bf16(in[i]);
}

// d = scale (pred - target) (FP32 and BF16), loss partial sum of squares in FP64; a
// grid-stride loop over at most MSE_BLOCKS workgroups, so the loss takes few atomics
constexpr int MSE_BLOCKS = 512;
__global__ void mse_grad_k(const float* __restrict__ pred, const float* __restrict__ target, int64_t n, float scale,
                           float* __restrict__ d, __bf16* __restrict__ db, double* __restrict__ loss) {
  __shared__ double red[256 / 64];
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float e = pred[i] - target[i];
    const float gi = scale * e;
    if (d) d[i] = gi;
    if (db) db[i] = to_bf16(gi);
    acc += (double)e * e;
  }
  const double s = phx::block_sum_f64(acc, red);
  if (threadIdx.x == 0 && loss) phx::atomic_add_f64(loss, s);
}

// Features: 2 x 2 binned corner-origin pattern, log(1 + I / scale), as BF16 rows.
__global__ void pattern_features_k(const float* __restrict__ pats, const int32_t* __restrict__ index, int64_t count,
                                   int n, int bin, float inv_scale, __bf16* __restrict__ out) {
  const int h = n / bin, hh = h * h;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count * hh) return;
  const int64_t r = t / hh;
  const int e = (int)(t % hh), y = e / h, x = e % h;
  const int64_t p = index ? index[r] : r;
  const float* P = pats + p * (int64_t)n * n + (int64_t)(bin * y) * n + bin * x;
  float s = 0.f;  // the bin x bin block summed row by row
  for (int i = 0; i < bin; ++i)
    for (int j = 0; j < bin; ++j) s += P[i * n + j];
  out[r * hh + e] = to_bf16(log1pf(fmaxf(s, 0.f) * inv_scale));
}

// Patch targets: for position r, the ps x ps window starting at (y0, x0)[r] of the
// projected object prod_s O_s: phase (minus `phase0`) and amplitude.
__global__ void object_patches_k(const float2* __restrict__ obj, int S, int H, int W, const int32_t* __restrict__ starts,
                                 int64_t count, int ps, float phase0, float* __restrict__ out) {
  const int pp = ps * ps;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count * pp) return;
  const int64_t r = t / pp;
  const int e = (int)(t % pp);
  const int y = starts[2 * r] + e / ps, x = starts[2 * r + 1] + e % ps;
  float2 prod = make_float2(1.f, 0.f);
  if ((unsigned)y < (unsigned)H && (unsigned)x < (unsigned)W)
    for (int s = 0; s < S; ++s) prod = phx::cmul(prod, obj[((int64_t)s * H + y) * W + x]);
  out[r * 2 * pp + e] = atan2f(prod.y, prod.x) - phase0;
  out[r * 2 * pp + pp + e] = sqrtf(phx::abs2f(prod));
}

// Stitch predicted patches (phase, amplitude rows) into per-pixel sums over the
// positions covering each pixel (atomics: the average is order-insensitive up to
// rounding).
__global__ void stitch_patches_k(const float* __restrict__ pred, int64_t ldp, const int32_t* __restrict__ starts,
                                 int64_t count, int ps, int H, int W, float* __restrict__ acc_phase,
                                 float* __restrict__ acc_amp, float* __restrict__ acc_n) {
  const int pp = ps * ps;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count * pp) return;
  const int64_t r = t / pp;
  const int e = (int)(t % pp);
  const int y = starts[2 * r] + e / ps, x = starts[2 * r + 1] + e % ps;
  if ((unsigned)y >= (unsigned)H || (unsigned)x >= (unsigned)W) return;
  const int64_t o = (int64_t)y * W + x;
  phx::atomic_add_f32(&acc_phase[o], pred[r * ldp + e]);
  phx::atomic_add_f32(&acc_amp[o], pred[r * ldp + pp + e]);
  phx::atomic_add_f32(&acc_n[o], 1.f);
}

// O_s = a^(1/S) exp(i phi / S) from the averaged projected phase/amplitude; pixels no
// patch covers stay 1 (flat start)
__global__ void stitch_finish_k(const float* __restrict__ acc_phase, const float* __restrict__ acc_amp,
                                const float* __restrict__ acc_n, int S, int64_t HW, float2* __restrict__ obj) {
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= HW) return;
  const float cnt = acc_n[o];
  float2 v = make_float2(1.f, 0.f);
  if (cnt > 0.f) {
    const float ph = acc_phase[o] / cnt / (float)S;
    const float am = powf(fmaxf(acc_amp[o] / cnt, 1e-6f), 1.f / (float)S);
    v = make_float2(am * cosf(ph), am * sinf(ph));
  }
  for (int s = 0; s < S; ++s) obj[(int64_t)s * HW + o] = v;
}

inline unsigned blocks_for(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

// 128 x 128 tiles when they give a workgroup per CU or more, else 128 x 64
int launch_gemm_nt(const void* A, int64_t lda, const void* B, int64_t ldb, int64_t M, int64_t N, int64_t K, float* C,
                   void* Cb, int64_t ldc, void* Ct, int64_t ldct, const float* bias, int relu, const void* mask,
                   int64_t ldm, int accumulate, hipStream_t stream) {
  if (M / NT_BM > 65535) return PHX_EUNSUPPORTED;
  const bool wide = (M / NT_BM) * (N / 128) >= 256;
  const bool deep = (K / NT_BK) % 4 == 0;  // 4 stages in flight when the stage count allows
  const __bf16* a = (const __bf16*)A;
  const __bf16* b = (const __bf16*)B;
  const __bf16* mk = (const __bf16*)mask;
#define G(BN, D)                                                                                                      \
  hipLaunchKernelGGL((gemm_nt_k<BN, D>), dim3((unsigned)(N / BN), (unsigned)(M / NT_BM)), dim3(256), 0, stream, a, lda, \
                     b, ldb, (int)K, C, (__bf16*)Cb, ldc, (__bf16*)Ct, ldct, bias, relu, mk, ldm, accumulate)
  if (wide) {
    if (deep) G(128, 4);
    else G(128, 1);
  } else {
    if (deep) G(64, 4);
    else G(64, 1);
  }
#undef G
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

}  // namespace

extern "C" {

int phx_gemm_bf16_nt(const void* A, int64_t lda, const void* B, int64_t ldb, int64_t M, int64_t N, int64_t K, float* C,
                     void* Cb, int64_t ldc, void* Ct, int64_t ldct, const float* bias, int relu, const void* mask,
                     int64_t ldm, int accumulate, hipStream_t stream) {
  if (!A || !B || (!C && !Cb && !Ct) || M <= 0 || N <= 0 || K <= 0) return PHX_EINVAL;
  if (M % NT_BM || N % 128 || K % NT_BK) return PHX_EUNSUPPORTED;
  if ((accumulate && !C) || (mask && ldm < N) || ((C || Cb) && ldc < N) || (Ct && ldct < M)) return PHX_EINVAL;
  if (Ct && (ldct % 4 || reinterpret_cast<uintptr_t>(Ct) % 8)) return PHX_EINVAL;  // 8-byte transposed stores
  if (lda < K || ldb < K || (lda % 8) || (ldb % 8)) return PHX_EINVAL;
  return launch_gemm_nt(A, lda, B, ldb, M, N, K, C, Cb, ldc, Ct, ldct, bias, relu, mask, ldm, accumulate, stream);
}

int phx_transpose_bf16(const void* in, int64_t ldi, int64_t rows, int64_t cols, void* out, int64_t ldo,
                       hipStream_t stream) {
  if (!in || !out || rows <= 0 || cols <= 0 || ldi < cols || ldo < rows || in == out) return PHX_EINVAL;
  if ((rows + 63) / 64 > 65535) return PHX_EUNSUPPORTED;
  hipLaunchKernelGGL(transpose_bf16_k, dim3((unsigned)((cols + 63) / 64), (unsigned)((rows + 63) / 64)), dim3(256), 0,
                     stream, (const __bf16*)in, ldi, rows, cols, (__bf16*)out, ldo);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_mse_grad_t(const float* pred, const float* target, int64_t rows, int64_t cols, float scale, float* d, void* db,
                   void* dbt, double* loss, hipStream_t stream) {
  if (!pred || !target || rows <= 0 || cols <= 0 || (dbt && !db)) return PHX_EINVAL;
  if (dbt && (rows + 63) / 64 > 65535) return PHX_EUNSUPPORTED;
  const int64_t n = rows * cols;
  hipLaunchKernelGGL(mse_grad_k, dim3(std::min<unsigned>(blocks_for(n, 256), MSE_BLOCKS)), dim3(256), 0, stream, pred,
                     target, n, scale, d, (__bf16*)db, loss);
  PHX_CHECK_LAUNCH();
  if (dbt) {
    // the transpose through LDS tiles (coalesced both ways)
    hipLaunchKernelGGL(transpose_bf16_k, dim3((unsigned)((cols + 63) / 64), (unsigned)((rows + 63) / 64)), dim3(256),
                       0, stream, (const __bf16*)db, cols, rows, cols, (__bf16*)dbt, rows);
    PHX_CHECK_LAUNCH();
  }
  return PHX_OK;
}

int phx_gemm_bf16(const void* A, int64_t lda, int trans_a, const void* B, int64_t ldb, int trans_b, int64_t M,
                  int64_t N, int64_t K, float* C, void* Cb, int64_t ldc, const float* bias, int relu, const void* mask,
                  int64_t ldm, int accumulate, hipStream_t stream) {
  if (!A || !B || (!C && !Cb) || M <= 0 || N <= 0 || K <= 0) return PHX_EINVAL;
  if (M % BM || N % BN || K % BK) return PHX_EUNSUPPORTED;
  if ((accumulate && !C) || (mask && ldm < N) || ldc < N) return PHX_EINVAL;
  if ((!trans_a && lda < K) || (trans_a && lda < M) || (!trans_b && ldb < N) || (trans_b && ldb < K)) return PHX_EINVAL;
  if ((lda % 8) || (ldb % 8)) return PHX_EINVAL;  // 16-byte operand loads
  if (M / BM > 65535) return PHX_EUNSUPPORTED;
  if (!trans_a && trans_b && K % NT_BK == 0)
    return launch_gemm_nt(A, lda, B, ldb, M, N, K, C, Cb, ldc, nullptr, 0, bias, relu, mask, ldm, accumulate, stream);
  const dim3 grid((unsigned)(N / BN), (unsigned)(M / BM));
  const __bf16* a = (const __bf16*)A;
  const __bf16* b = (const __bf16*)B;
  __bf16* cb = (__bf16*)Cb;
  const __bf16* mk = (const __bf16*)mask;
#define G(TA, TB) \
  hipLaunchKernelGGL((gemm_bf16_k<TA, TB>), grid, dim3(256), 0, stream, a, lda, b, ldb, (int)K, C, cb, ldc, bias, relu, mk, ldm, accumulate)
  if (!trans_a && !trans_b) G(false, false);
  else if (!trans_a && trans_b) G(false, true);
  else if (trans_a && !trans_b) G(true, false);
  else G(true, true);
#undef G
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_colsum_f32(const float* in, int64_t rows, int64_t cols, int64_t ld, float* out, hipStream_t stream) {
  if (!in || !out || rows < 0 || cols <= 0 || ld < cols) return PHX_EINVAL;
  hipLaunchKernelGGL(colsum_k, dim3(blocks_for(cols, 64)), dim3(1024), 0, stream, in, rows, cols, ld, out);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_adam_f32_bf16(float* w, void* wb, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2,
                      float eps, float c1, float c2, float weight_decay, hipStream_t stream) {
  if (!w || !g || !m || !v || n <= 0) return PHX_EINVAL;
  hipLaunchKernelGGL(adam_k, dim3(blocks_for(n, 256)), dim3(256), 0, stream, w, (__bf16*)wb, g, m, v, n, lr, b1, b2,
                     eps, c1, c2, weight_decay);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_f32_to_bf16(const float* in, void* out, int64_t n, hipStream_t stream) {
  if (!in || !out || n <= 0) return PHX_EINVAL;
  hipLaunchKernelGGL(f32_to_bf16_k, dim3(blocks_for(n, 256)), dim3(256), 0, stream, in, (__bf16*)out, n);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_mse_grad(const float* pred, const float* target, int64_t n, float scale, float* d, void* db, double* loss,
                 hipStream_t stream) {
  if (!pred || !target || n <= 0) return PHX_EINVAL;
  hipLaunchKernelGGL(mse_grad_k, dim3(std::min<unsigned>(blocks_for(n, 256), MSE_BLOCKS)), dim3(256), 0, stream, pred,
                     target, n, scale, d, (__bf16*)db, loss);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_pattern_features(const float* patterns, const int32_t* index, int64_t count, int64_t n, int64_t bin,
                         float inv_scale, void* out, hipStream_t stream) {
  if (!patterns || !out || count <= 0 || bin < 1 || n < bin || n % bin) return PHX_EINVAL;
  const int64_t total = count * (n / bin) * (n / bin);
  hipLaunchKernelGGL(pattern_features_k, dim3(blocks_for(total, 256)), dim3(256), 0, stream, patterns, index, count,
                     (int)n, (int)bin, inv_scale, (__bf16*)out);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_object_patches(const void* obj, int64_t S, int64_t H, int64_t W, const int32_t* starts, int64_t count,
                       int64_t ps, float phase0, float* out, hipStream_t stream) {
  if (!obj || !starts || !out || count <= 0 || ps <= 0 || S <= 0) return PHX_EINVAL;
  const int64_t total = count * ps * ps;
  hipLaunchKernelGGL(object_patches_k, dim3(blocks_for(total, 256)), dim3(256), 0, stream, (const float2*)obj, (int)S,
                     (int)H, (int)W, starts, count, (int)ps, phase0, out);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_stitch_patches(const float* pred, int64_t ldp, const int32_t* starts, int64_t count, int64_t ps, int64_t S,
                       int64_t H, int64_t W, float* acc_phase, float* acc_amp, float* acc_n, void* obj,
                       hipStream_t stream) {
  if (!pred || !starts || !acc_phase || !acc_amp || !acc_n || !obj || count <= 0 || ps <= 0 || ldp < 2 * ps * ps)
    return PHX_EINVAL;
  const int64_t total = count * ps * ps;
  hipLaunchKernelGGL(stitch_patches_k, dim3(blocks_for(total, 256)), dim3(256), 0, stream, pred, ldp, starts, count,
                     (int)ps, (int)H, (int)W, acc_phase, acc_amp, acc_n);
  PHX_CHECK_LAUNCH();
  hipLaunchKernelGGL(stitch_finish_k, dim3(blocks_for(H * W, 256)), dim3(256), 0, stream, acc_phase, acc_amp, acc_n,
                     (int)S, H * W, (float2*)obj);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

}  // extern "C"
