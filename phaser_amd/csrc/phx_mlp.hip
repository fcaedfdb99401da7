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

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
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

// out[c] = sum_r in[r, c] (FP32), one thread per column, rows in order
__global__ void colsum_k(const float* __restrict__ in, int64_t rows, int64_t cols, int64_t ld, float* __restrict__ out) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  float s = 0.f;
  for (int64_t r = 0; r < rows; ++r) s += in[r * ld + c];
  out[c] = s;
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

// d = scale (pred - target) (FP32 and BF16), loss partial sum of squares in FP64
__global__ void mse_grad_k(const float* __restrict__ pred, const float* __restrict__ target, int64_t n, float scale,
                           float* __restrict__ d, __bf16* __restrict__ db, double* __restrict__ loss) {
  __shared__ double red[256 / 64];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float e = 0.f;
  if (i < n) {
    e = pred[i] - target[i];
    const float gi = scale * e;
    if (d) d[i] = gi;
    if (db) db[i] = to_bf16(gi);
  }
  const double s = phx::block_sum_f64((double)e * e, red);
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

}  // namespace

extern "C" {

int phx_gemm_bf16(const void* A, int64_t lda, int trans_a, const void* B, int64_t ldb, int trans_b, int64_t M,
                  int64_t N, int64_t K, float* C, void* Cb, int64_t ldc, const float* bias, int relu, const void* mask,
                  int64_t ldm, int accumulate, hipStream_t stream) {
  if (!A || !B || (!C && !Cb) || M <= 0 || N <= 0 || K <= 0) return PHX_EINVAL;
  if (M % BM || N % BN || K % BK) return PHX_EUNSUPPORTED;
  if ((accumulate && !C) || (mask && ldm < N) || ldc < N) return PHX_EINVAL;
  if ((!trans_a && lda < K) || (trans_a && lda < M) || (!trans_b && ldb < N) || (trans_b && ldb < K)) return PHX_EINVAL;
  if ((lda % 8) || (ldb % 8)) return PHX_EINVAL;  // 16-byte operand loads
  if (M / BM > 65535) return PHX_EUNSUPPORTED;
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
  hipLaunchKernelGGL(colsum_k, dim3(blocks_for(cols, 256)), dim3(256), 0, stream, in, rows, cols, ld, out);
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
  hipLaunchKernelGGL(mse_grad_k, dim3(blocks_for(n, 256)), dim3(256), 0, stream, pred, target, n, scale, d,
                     (__bf16*)db, loss);
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
