// LDS-resident batched 2-D complex FFT for n x n planes (n = 16, 32, 64, 128) on gfx950.
//
// One workgroup owns one plane.  The plane lives in LDS (n x (n+1) float2, padded row
// stride), so a 128^2 complex64 plane takes 129 KiB of the 160 KiB LDS of a CU.
//
// Each 1-D length-n transform is a four-step decomposition n = R1 * R2 (R1, R2 <= 16)
// done in two register passes with an LDS exchange:
//   forward stage 1 : for each n2 < R2, DFT_R1 over x[R2*n1 + n2]; twiddle W_n^(n2*k1);
//                     write back in place at R2*k1 + n2.
//   forward stage 2 : for each k1 < R1, DFT_R2 over the R2 contiguous entries R2*k1 + j;
//                     write back in place.  Position p = R2*k1 + k2 now holds X[k1 + R1*k2]
//                     ("scrambled" order).
//   inverse stage A : exact reverse of forward stage 2 (inverse DFT_R2, in place).
//   inverse stage B : conj twiddle, inverse DFT_R1 over the stride-R2 entries; output is in
//                     natural order, in place.
// So a forward transform takes natural order to scrambled order and an inverse transform
// takes scrambled order back to natural order, without any permutation pass.  Pointwise
// work between them (propagators, noise-model updates, shift filters) is done in the
// scrambled order with frequency-domain constants stored in that same order.
//
// Transforms are unnormalised: forward uses exp(-2 pi i nk/n), inverse exp(+2 pi i nk/n).
// The (i)fftshift of the centred real-space convention cancels in every product the
// engines form (fft2 -> pointwise -> ifft2), so it is not applied here; callers scale by
// 1/n^2 per forward+inverse pair (ortho normalisation).
//
// The first stage of the row pass can read straight from global memory (fused load)
// and the last stage of the inverse row pass can write straight to global memory
// (fused store).  Forward stage 2 and inverse stage A of the column pass touch the same
// R2 entries, so a pointwise multiply between a forward and an inverse transform is done
// in registers with no LDS round trip (cols_f2_mul_iA).
#pragma once
#include "phx_common.h"

namespace phx {

// exp(-2 pi i k / 16), k in [0, 16)
__device__ __forceinline__ float2 w16(int k) {
  constexpr float C8 = 0.92387953251128674f, S8 = 0.38268343236508978f, R2f = 0.70710678118654752f;
  constexpr float c[16] = {1.f, C8, R2f, S8, 0.f, -S8, -R2f, -C8, -1.f, -C8, -R2f, -S8, 0.f, S8, R2f, C8};
  constexpr float s[16] = {0.f, -S8, -R2f, -C8, -1.f, -C8, -R2f, -S8, 0.f, S8, R2f, C8, 1.f, C8, R2f, S8};
  return make_float2(c[k], s[k]);
}

// v * exp(-2 pi i k / 16) with the trivial factors special-cased (k is a compile-time
// constant after unrolling).
__device__ __forceinline__ float2 twmul16(float2 v, int k) {
  if (k == 0) return v;
  if (k == 4) return make_float2(v.y, -v.x);
  if (k == 8) return make_float2(-v.x, -v.y);
  if (k == 12) return make_float2(-v.y, v.x);
  return cmul(v, w16(k));
}

template <int R> struct Log2;
template <> struct Log2<2> { static constexpr int v = 1; };
template <> struct Log2<4> { static constexpr int v = 2; };
template <> struct Log2<8> { static constexpr int v = 3; };
template <> struct Log2<16> { static constexpr int v = 4; };

// In-register length-R DFT (R = 2, 4, 8, 16), natural order in and out.
template <int R, bool INV>
__device__ __forceinline__ void dft(float2 (&x)[R]) {
  constexpr int LG = Log2<R>::v;
#pragma unroll
  for (int i = 0; i < R; ++i) {
    int j = 0;
#pragma unroll
    for (int b = 0; b < LG; ++b) j |= ((i >> b) & 1) << (LG - 1 - b);
    if (j > i) {
      float2 t = x[i];
      x[i] = x[j];
      x[j] = t;
    }
  }
#pragma unroll
  for (int len = 2; len <= R; len <<= 1) {
    const int half = len >> 1;
#pragma unroll
    for (int i = 0; i < R; i += len) {
#pragma unroll
      for (int j = 0; j < half; ++j) {
        int k = j * (16 / len);
        if (INV) k = (16 - k) & 15;
        const float2 v = twmul16(x[i + j + half], k);
        const float2 u = x[i + j];
        x[i + j] = cadd(u, v);
        x[i + j + half] = csub(u, v);
      }
    }
  }
}

template <int N> struct FFTShape;
template <> struct FFTShape<16> { static constexpr int R1 = 4, R2 = 4; };
template <> struct FFTShape<32> { static constexpr int R1 = 8, R2 = 4; };
template <> struct FFTShape<64> { static constexpr int R1 = 8, R2 = 8; };
template <> struct FFTShape<128> { static constexpr int R1 = 16, R2 = 8; };

// Frequency index held at scrambled position p (one axis).
template <int N>
__host__ __device__ __forceinline__ int scr_to_freq(int p) {
  return p / FFTShape<N>::R2 + FFTShape<N>::R1 * (p % FFTShape<N>::R2);
}
// Scrambled position holding frequency index k (one axis).
template <int N>
__host__ __device__ __forceinline__ int freq_to_scr(int k) {
  return FFTShape<N>::R2 * (k % FFTShape<N>::R1) + k / FFTShape<N>::R1;
}

// LDS plane of N x N complex values with T threads per workgroup.
template <int N, int T>
struct Plane {
  static constexpr int R1 = FFTShape<N>::R1, R2 = FFTShape<N>::R2;
  static constexpr int LD = N + 1;
  static constexpr int LDS_FLOAT2 = N * LD;
  float2* s;         // N x LD
  const float2* tw;  // W_N^e, e in [0, N), in LDS

  __device__ __forceinline__ float2& at(int y, int x) { return s[y * LD + x]; }

  // ---------------- row transforms (along x) ----------------
  // Forward stage 1 with a fused source: src(y, x) -> natural-order input value.
  template <class Src>
  __device__ __forceinline__ void rows_f1(Src src) {
#pragma unroll 1
    for (int id = threadIdx.x; id < N * R2; id += T) {
      const int r = id / R2, n2 = id % R2;
      float2 v[R1];
#pragma unroll
      for (int j = 0; j < R1; ++j) v[j] = src(r, R2 * j + n2);
      dft<R1, false>(v);
      float2* row = s + r * LD;
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) row[R2 * k1 + n2] = (k1 == 0) ? v[0] : cmul(v[k1], tw[n2 * k1]);
    }
  }
  __device__ __forceinline__ void rows_f2() {
#pragma unroll 1
    for (int id = threadIdx.x; id < N * R1; id += T) {
      const int r = id / R1, k1 = id % R1;
      float2* p = s + r * LD + R2 * k1;
      float2 v[R2];
#pragma unroll
      for (int j = 0; j < R2; ++j) v[j] = p[j];
      dft<R2, false>(v);
#pragma unroll
      for (int j = 0; j < R2; ++j) p[j] = v[j];
    }
  }
  __device__ __forceinline__ void rows_iA() {
#pragma unroll 1
    for (int id = threadIdx.x; id < N * R1; id += T) {
      const int r = id / R1, k1 = id % R1;
      float2* p = s + r * LD + R2 * k1;
      float2 v[R2];
#pragma unroll
      for (int j = 0; j < R2; ++j) v[j] = p[j];
      dft<R2, true>(v);
#pragma unroll
      for (int j = 0; j < R2; ++j) p[j] = v[j];
    }
  }
  // Inverse stage B with a fused sink: dst(y, x, value) receives natural-order output.
  template <class Dst>
  __device__ __forceinline__ void rows_iB(Dst dst) {
#pragma unroll 1
    for (int id = threadIdx.x; id < N * R2; id += T) {
      const int r = id / R2, n2 = id % R2;
      const float2* row = s + r * LD;
      float2 v[R1];
#pragma unroll
      for (int j = 0; j < R1; ++j) {
        const float2 a = row[R2 * j + n2];
        v[j] = (j == 0) ? a : cmulc(a, tw[n2 * j]);
      }
      dft<R1, true>(v);
#pragma unroll
      for (int n1 = 0; n1 < R1; ++n1) dst(r, R2 * n1 + n2, v[n1]);
    }
  }
  // Inverse stage B written back into LDS.
  __device__ __forceinline__ void rows_iB_lds() {
    rows_iB([&](int y, int x, float2 v) { s[y * LD + x] = v; });
  }

  // ---------------- column transforms (along y) ----------------
  __device__ __forceinline__ void cols_f1() {
#pragma unroll 1
    for (int id = threadIdx.x; id < N * R2; id += T) {
      const int c = id % N, n2 = id / N;
      float2 v[R1];
#pragma unroll
      for (int j = 0; j < R1; ++j) v[j] = s[(R2 * j + n2) * LD + c];
      dft<R1, false>(v);
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) s[(R2 * k1 + n2) * LD + c] = (k1 == 0) ? v[0] : cmul(v[k1], tw[n2 * k1]);
    }
  }
  // Forward stage 2; fn(py, px, v) may modify v (scrambled positions) before it is stored.
  template <class Fn>
  __device__ __forceinline__ void cols_f2(Fn fn) {
#pragma unroll 1
    for (int id = threadIdx.x; id < N * R1; id += T) {
      const int c = id % N, k1 = id / N;
      float2 v[R2];
#pragma unroll
      for (int j = 0; j < R2; ++j) v[j] = s[(R2 * k1 + j) * LD + c];
      dft<R2, false>(v);
#pragma unroll
      for (int j = 0; j < R2; ++j) {
        fn(R2 * k1 + j, c, v[j]);
        s[(R2 * k1 + j) * LD + c] = v[j];
      }
    }
  }
  // Forward stage 2, pointwise mul(py, px, v) -> v', inverse stage A, all in registers.
  template <class Mul>
  __device__ __forceinline__ void cols_f2_mul_iA(Mul mul) {
#pragma unroll 1
    for (int id = threadIdx.x; id < N * R1; id += T) {
      const int c = id % N, k1 = id / N;
      float2 v[R2];
#pragma unroll
      for (int j = 0; j < R2; ++j) v[j] = s[(R2 * k1 + j) * LD + c];
      dft<R2, false>(v);
#pragma unroll
      for (int j = 0; j < R2; ++j) v[j] = mul(R2 * k1 + j, c, v[j]);
      dft<R2, true>(v);
#pragma unroll
      for (int j = 0; j < R2; ++j) s[(R2 * k1 + j) * LD + c] = v[j];
    }
  }
  // Inverse stage A with a fused source: src(py, px) -> scrambled-order input value.
  template <class Src>
  __device__ __forceinline__ void cols_iA(Src src) {
#pragma unroll 1
    for (int id = threadIdx.x; id < N * R1; id += T) {
      const int c = id % N, k1 = id / N;
      float2 v[R2];
#pragma unroll
      for (int j = 0; j < R2; ++j) v[j] = src(R2 * k1 + j, c);
      dft<R2, true>(v);
#pragma unroll
      for (int j = 0; j < R2; ++j) s[(R2 * k1 + j) * LD + c] = v[j];
    }
  }
  __device__ __forceinline__ void cols_iB() {
#pragma unroll 1
    for (int id = threadIdx.x; id < N * R2; id += T) {
      const int c = id % N, n2 = id / N;
      float2 v[R1];
#pragma unroll
      for (int j = 0; j < R1; ++j) {
        const float2 a = s[(R2 * j + n2) * LD + c];
        v[j] = (j == 0) ? a : cmulc(a, tw[n2 * j]);
      }
      dft<R1, true>(v);
#pragma unroll
      for (int n1 = 0; n1 < R1; ++n1) s[(R2 * n1 + n2) * LD + c] = v[n1];
    }
  }

  // Full forward transform from a fused source; result scrambled in LDS; fn as in cols_f2.
  template <class Src, class Fn>
  __device__ __forceinline__ void forward(Src src, Fn fn) {
    rows_f1(src);
    __syncthreads();
    rows_f2();
    __syncthreads();
    cols_f1();
    __syncthreads();
    cols_f2(fn);
    __syncthreads();
  }
  // Full inverse transform of the scrambled LDS plane into a fused sink (src as in cols_iA).
  template <class Src, class Dst>
  __device__ __forceinline__ void inverse(Src src, Dst dst) {
    cols_iA(src);
    __syncthreads();
    cols_iB();
    __syncthreads();
    rows_iA();
    __syncthreads();
    rows_iB(dst);
  }
  // forward -> pointwise -> inverse
  template <class Src, class Mul, class Dst>
  __device__ __forceinline__ void convolve(Src src, Mul mul, Dst dst) {
    rows_f1(src);
    __syncthreads();
    rows_f2();
    __syncthreads();
    cols_f1();
    __syncthreads();
    cols_f2_mul_iA(mul);
    __syncthreads();
    cols_iB();
    __syncthreads();
    rows_iA();
    __syncthreads();
    rows_iB(dst);
  }
};

// Load the twiddle table (n entries, host-computed in FP64) into LDS.
template <int N, int T>
__device__ __forceinline__ void load_twiddles(float2* tw_lds, const float2* tw_glob) {
  for (int i = threadIdx.x; i < N; i += T) tw_lds[i] = tw_glob[i];
}

}  // namespace phx
