// Object cutout gather / scatter-add / set, fills, and the standalone convention-exact
// batched fft2/ifft2.
//
// The cutout entry points replace phaser's cupy kernels get/add/set_cutouts
// (phaser/utils/_cuda_kernels.py:19-107, kernel template :163-211): same index
// semantics (start_idxs (n_cutouts, 2) int64 rows of (y, x); cutout c of object slice o
// reads obj[o, start_y + i, start_x + j]; out-of-bounds reads give `default`, out-of-
// bounds adds/sets are skipped), same layout (cutouts (n_cutouts, n_objects, cy, cx)).
#include "phx_common.h"
#include "phx_fft.h"
#include "../../include/phaser_mi355x.h"

using namespace phx;

namespace {

template <class Tv>
struct Ops;
template <>
struct Ops<float> {
  static __device__ __forceinline__ void add(float* p, float v) { atomic_add_f32(p, v); }
};
template <>
struct Ops<float2> {
  static __device__ __forceinline__ void add(float2* p, float2 v) { atomic_add_c64(p, v); }
};

enum class CutOp { Get, Add, Set };

// grid: (ceil(cx/64), ceil(cy/4), n_cutouts*n_objects); block (64, 4)
template <class Tv, CutOp OP>
__global__ void __launch_bounds__(256) cutout_k(Tv* __restrict__ obj, Tv* __restrict__ cut,
                                                const int64_t* __restrict__ start, int64_t n_cut,
                                                int64_t n_obj, int64_t oh, int64_t ow, int64_t ch,
                                                int64_t cw, Tv dflt) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i = (int64_t)blockIdx.y * blockDim.y + threadIdx.y;
  const int64_t cn = blockIdx.z;
  const int64_t o = cn % n_obj, c = cn / n_obj;
  if (c >= n_cut || i >= ch || j >= cw) return;
  const int64_t cidx = (cn * ch + i) * cw + j;
  const int64_t yi = i + start[2 * c], xj = j + start[2 * c + 1];
  if (yi < 0 || yi >= oh || xj < 0 || xj >= ow) {
    if (OP == CutOp::Get) cut[cidx] = dflt;
    return;
  }
  const int64_t oidx = (o * oh + yi) * ow + xj;
  if (OP == CutOp::Get) cut[cidx] = obj[oidx];
  else if (OP == CutOp::Add) Ops<Tv>::add(&obj[oidx], cut[cidx]);
  else obj[oidx] = cut[cidx];
}

template <class Tv, CutOp OP>
int launch_cutout(void* obj, void* cut, const int64_t* start, int64_t n_cut, int64_t n_obj, int64_t oh,
                  int64_t ow, int64_t ch, int64_t cw, Tv dflt, hipStream_t stream) {
  if (!obj || !cut || !start || n_cut < 0 || n_obj < 1 || oh < 1 || ow < 1 || ch < 1 || cw < 1) return PHX_EINVAL;
  if (n_cut == 0) return PHX_OK;
  if (n_cut * n_obj > 0x7fffffffLL || ch > 0x7fffffffLL * 4) return PHX_EINVAL;
  dim3 block(64, 4, 1);
  dim3 grid((unsigned)((cw + 63) / 64), (unsigned)((ch + 3) / 4), (unsigned)(n_cut * n_obj));
  hipLaunchKernelGGL((cutout_k<Tv, OP>), grid, block, 0, stream, (Tv*)obj, (Tv*)cut, start, n_cut, n_obj, oh,
                     ow, ch, cw, dflt);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

__global__ void fill_u32_k(uint32_t* p, uint32_t v, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// ----- convention-exact fft2 / ifft2 (phaser/utils/num.py:434-468) for testing -----
template <int N, int T>
__global__ void __launch_bounds__(T) fft2_conv_k(const float2* __restrict__ in, float2* __restrict__ out,
                                                 const float2* __restrict__ twg, int inverse) {
  static_assert(T == Plane<N>::T, "plane kernels run one workgroup of Plane<N>::T threads per plane");
  constexpr int E = Plane<N>::E;
  __shared__ float2 lds[Plane<N>::LDS_FLOAT2 + N];
  const float2* a = in + (size_t)blockIdx.x * N * N;
  float2* b = out + (size_t)blockIdx.x * N * N;
  Plane<N> pl(lds, lds + Plane<N>::LDS_FLOAT2);
  load_twiddles<N>(lds + Plane<N>::LDS_FLOAT2, twg);
  __syncthreads();
  const float sc = 1.0f / (float)N;
  float2 x[E];
  if (!inverse) {
    // fft2(a) = (-1)^(ky+kx) * FFT(a) / N  for even N
    pl.load_rows([&](int y, int xx) { return a[y * N + xx]; }, x);
    pl.forward_regs(x);
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const int ky = scr_to_freq<N>(pl.QL * k + pl.q), kx = scr_to_freq<N>(pl.rc);
      const float sg = ((ky + kx) & 1) ? -sc : sc;
      b[ky * N + kx] = cscale(x[k], sg);
    }
  } else {
    // ifft2(A) = IFFT((-1)^(ky+kx) * A) / N
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const int ky = scr_to_freq<N>(pl.QL * k + pl.q), kx = scr_to_freq<N>(pl.rc);
      const float sg = ((ky + kx) & 1) ? -sc : sc;
      x[k] = cscale(a[ky * N + kx], sg);
    }
    pl.inverse_regs(x);
    pl.store_rows([&](int y, int xx, float2 v) { b[y * N + xx] = v; }, x);
  }
}

// out[b][py][px] = in[b][freq(py)][freq(px)] (scramble) or the inverse map.
template <int N, class Tv>
__global__ void scramble_k(const Tv* __restrict__ in, Tv* __restrict__ out, int unscramble) {
  const int b = blockIdx.y;
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= N * N) return;
  const int py = e / N, px = e % N;
  const int fe = scr_to_freq<N>(py) * N + scr_to_freq<N>(px);
  const size_t base = (size_t)b * N * N;
  if (unscramble) out[base + fe] = in[base + e];
  else out[base + e] = in[base + fe];
}

template <int N>
int launch_scramble(const void* in, void* out, int64_t batch, int is_complex, int unscramble, hipStream_t stream) {
  dim3 grid((unsigned)((N * N + 255) / 256), (unsigned)batch);
  if (is_complex)
    hipLaunchKernelGGL((scramble_k<N, float2>), grid, dim3(256), 0, stream, (const float2*)in, (float2*)out, unscramble);
  else
    hipLaunchKernelGGL((scramble_k<N, float>), grid, dim3(256), 0, stream, (const float*)in, (float*)out, unscramble);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

// flag = min(flag, value) if any of the n floats is NaN or inf (one atomic per wave).
__global__ void check_finite_k(const float* p, int64_t n, int32_t* flag, int32_t value) {
  bool bad = false;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    bad |= !isfinite(p[i]);
  if (__ballot(bad) != 0ull && (threadIdx.x & 63) == 0) atomicMin(flag, value);
}

__global__ void scale_c64_k(float2* p, int64_t n, float s) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = cscale(p[i], s);
}

template <int N, int T>
int launch_fft2(const void* in, void* out, int64_t batch, int inverse, const void* tw, hipStream_t stream) {
  hipLaunchKernelGGL((fft2_conv_k<N, T>), dim3((unsigned)batch), dim3(T), 0, stream, (const float2*)in,
                     (float2*)out, (const float2*)tw, inverse);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

// ----- n = 256: split plane, one launch per axis -----
// A 256 x 256 complex64 plane (512 KiB) fits neither a CU's LDS nor one workgroup's
// registers (64 points per lane at 1024 threads), so the transform is split: a row
// pass (in -> out) and a column pass (out -> out, in place), each a launch of
// workgroups that own 16 (rows) or 32 (columns) whole lines of one plane.  A line is held by
// 16 threads; thread t of a line loads x[16 a + t], a < 16, and the length-256
// transform is the four-step 16 x 16 split
//   X[k1 + 16 k2] = sum_t W_16^(t k2) W_256^(t k1) sum_a x[16 a + t] W_16^(a k1):
// an in-register DFT_16 over a, the twiddle W_256^(t k1) (table in LDS), an exchange
// through LDS after which thread k1 holds the 16 values of index k1, and a second
// in-register DFT_16 that leaves X[k1 + 16 k2] in register k2.  The row pass reads and
// writes whole rows (a wave covers 4 rows, 512 contiguous bytes per load); the column
// pass reads and writes 32-column strips (a wave covers 2 rows of 256 bytes).  The
// column pass reads its strip whole before writing it, and no other workgroup touches
// it, so in place is safe.  Traffic: 2 x 512 KiB read + 2 x 512 KiB written per plane.
constexpr int F256_LD = 17;  // padded row of a line's 16 x 16 exchange block

// LINES lines per workgroup of 16 LINES threads: 16 rows in the row pass, 32 columns in
// the column pass (a wave then reads 2 rows x 256 contiguous bytes per load).
template <bool COL, int LINES, bool INV>
__global__ void __launch_bounds__(16 * LINES) fft256_pass_k(const float2* in, float2* out,
                                                             const float2* __restrict__ twg) {
  constexpr int TILES = 256 / LINES;
  // stride between the lines' blocks: in the column pass the LINES lines of a load are
  // adjacent lanes, so an odd stride spreads them over the banks; in the row pass a
  // wave's 4 lines at stride 272 (16 mod 32 float2 slots) meet at most 2 to a bank
  constexpr int BLK = 16 * F256_LD + (COL ? 1 : 0);
  __shared__ float2 ex[LINES * BLK];
  __shared__ float2 tw[256];
  const int tid = threadIdx.x;
  const int tile = blockIdx.x % TILES;
  const size_t plane = (size_t)(blockIdx.x / TILES) * 256 * 256;
  // row pass: line l = tid / 16 is row LINES tile + l; column pass: line l = tid % LINES
  // is column LINES tile + l (LINES adjacent threads read adjacent columns of one row)
  const int l = COL ? (tid % LINES) : (tid / 16);
  const int t = COL ? (tid / LINES) : (tid % 16);
  const int line = LINES * tile + l;
  // element e of this line sits at base + e * es
  const size_t base = plane + (COL ? (size_t)line : (size_t)line * 256);
  const int es = COL ? 256 : 1;
  for (int k = tid; k < 256; k += 16 * LINES) tw[k] = twg[k];
  const float sc = 1.0f / 256.0f;
  float2 x[16];
#pragma unroll
  for (int a = 0; a < 16; ++a) {
    float2 v = in[base + (size_t)(16 * a + t) * es];
    if (!COL && INV) {
      // ifft2(A) = IFFT((-1)^(ky+kx) A) / n: the row pass sees (ky, kx) = (line, e)
      v = cscale(v, ((line + 16 * a + t) & 1) ? -sc : sc);
    }
    x[a] = v;
  }
  __syncthreads();  // twiddle table written
  dft<16, INV>(x);
#pragma unroll
  for (int k1 = 1; k1 < 16; ++k1) {
    const float2 w = tw[(t * k1) & 255];
    x[k1] = INV ? cmulc(x[k1], w) : cmul(x[k1], w);
  }
  float2* blk = ex + l * BLK;
#pragma unroll
  for (int k1 = 0; k1 < 16; ++k1) blk[k1 * F256_LD + t] = x[k1];
  __syncthreads();
  // thread t now plays k1 = t
#pragma unroll
  for (int j = 0; j < 16; ++j) x[j] = blk[t * F256_LD + j];
  dft<16, INV>(x);
#pragma unroll
  for (int k2 = 0; k2 < 16; ++k2) {
    const int e = t + 16 * k2;
    float2 v = x[k2];
    if (COL && !INV) {
      // fft2(a) = (-1)^(ky+kx) FFT(a) / n: the column pass sees (ky, kx) = (e, line)
      v = cscale(v, ((line + e) & 1) ? -sc : sc);
    }
    out[base + (size_t)e * es] = v;
  }
}

#ifndef PHX_F256_COL_LINES
#define PHX_F256_COL_LINES 32
#endif
constexpr int F256_ROW_LINES = 16, F256_COL_LINES = PHX_F256_COL_LINES;

template <bool INV>
void launch_fft256_passes(const void* in, void* out, int64_t batch, const void* tw, hipStream_t stream) {
  hipLaunchKernelGGL((fft256_pass_k<false, F256_ROW_LINES, INV>), dim3((unsigned)(batch * (256 / F256_ROW_LINES))),
                     dim3(16 * F256_ROW_LINES), 0, stream, (const float2*)in, (float2*)out, (const float2*)tw);
  hipLaunchKernelGGL((fft256_pass_k<true, F256_COL_LINES, INV>), dim3((unsigned)(batch * (256 / F256_COL_LINES))),
                     dim3(16 * F256_COL_LINES), 0, stream, (const float2*)out, (float2*)out, (const float2*)tw);
}

int launch_fft256(const void* in, void* out, int64_t batch, int inverse, const void* tw, hipStream_t stream) {
  // 256 / LINES workgroups per plane and pass; the grid's x extent bounds the batch
  if (batch > (1LL << 20)) return PHX_EINVAL;
  if (inverse) launch_fft256_passes<true>(in, out, batch, tw, stream);
  else launch_fft256_passes<false>(in, out, batch, tw, stream);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

}  // namespace

extern "C" {

int phx_get_cutouts_f32(const void* obj, void* cutouts, const int64_t* start_idxs, int64_t n_cutouts,
                        int64_t n_objects, int64_t obj_h, int64_t obj_w, int64_t cut_h, int64_t cut_w,
                        float default_value, hipStream_t stream) {
  return launch_cutout<float, CutOp::Get>((void*)obj, cutouts, start_idxs, n_cutouts, n_objects, obj_h, obj_w,
                                          cut_h, cut_w, default_value, stream);
}

int phx_get_cutouts_c64(const void* obj, void* cutouts, const int64_t* start_idxs, int64_t n_cutouts,
                        int64_t n_objects, int64_t obj_h, int64_t obj_w, int64_t cut_h, int64_t cut_w,
                        float default_re, float default_im, hipStream_t stream) {
  return launch_cutout<float2, CutOp::Get>((void*)obj, cutouts, start_idxs, n_cutouts, n_objects, obj_h, obj_w,
                                           cut_h, cut_w, make_float2(default_re, default_im), stream);
}

int phx_add_cutouts_f32(void* obj, const void* cutouts, const int64_t* start_idxs, int64_t n_cutouts,
                        int64_t n_objects, int64_t obj_h, int64_t obj_w, int64_t cut_h, int64_t cut_w,
                        hipStream_t stream) {
  return launch_cutout<float, CutOp::Add>(obj, (void*)cutouts, start_idxs, n_cutouts, n_objects, obj_h, obj_w,
                                          cut_h, cut_w, 0.f, stream);
}

int phx_add_cutouts_c64(void* obj, const void* cutouts, const int64_t* start_idxs, int64_t n_cutouts,
                        int64_t n_objects, int64_t obj_h, int64_t obj_w, int64_t cut_h, int64_t cut_w,
                        hipStream_t stream) {
  return launch_cutout<float2, CutOp::Add>(obj, (void*)cutouts, start_idxs, n_cutouts, n_objects, obj_h, obj_w,
                                           cut_h, cut_w, make_float2(0.f, 0.f), stream);
}

int phx_set_cutouts_f32(void* obj, const void* cutouts, const int64_t* start_idxs, int64_t n_cutouts,
                        int64_t n_objects, int64_t obj_h, int64_t obj_w, int64_t cut_h, int64_t cut_w,
                        hipStream_t stream) {
  return launch_cutout<float, CutOp::Set>(obj, (void*)cutouts, start_idxs, n_cutouts, n_objects, obj_h, obj_w,
                                          cut_h, cut_w, 0.f, stream);
}

int phx_set_cutouts_c64(void* obj, const void* cutouts, const int64_t* start_idxs, int64_t n_cutouts,
                        int64_t n_objects, int64_t obj_h, int64_t obj_w, int64_t cut_h, int64_t cut_w,
                        hipStream_t stream) {
  return launch_cutout<float2, CutOp::Set>(obj, (void*)cutouts, start_idxs, n_cutouts, n_objects, obj_h, obj_w,
                                           cut_h, cut_w, make_float2(0.f, 0.f), stream);
}

int phx_fill_u32(void* ptr, uint32_t value, int64_t count, hipStream_t stream) {
  if (!ptr || count < 0) return PHX_EINVAL;
  if (count == 0) return PHX_OK;
  int64_t blocks = (count + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(fill_u32_k, dim3((unsigned)blocks), dim3(256), 0, stream, (uint32_t*)ptr, value, count);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_fft2_c64(const void* in, void* out, int64_t batch, int64_t n, int inverse, const void* twiddle,
                 hipStream_t stream) {
  if (!in || !out || !twiddle || batch < 0 || batch > 0x7fffffffLL) return PHX_EINVAL;
  if (in == out) return PHX_EINVAL;
  if (batch == 0) return PHX_OK;
  switch (n) {
    case 16: return launch_fft2<16, 64>(in, out, batch, inverse, twiddle, stream);
    case 32: return launch_fft2<32, 128>(in, out, batch, inverse, twiddle, stream);
    case 64: return launch_fft2<64, 256>(in, out, batch, inverse, twiddle, stream);
    case 128: return launch_fft2<128, 1024>(in, out, batch, inverse, twiddle, stream);
    case 192: return launch_fft2<192, 768>(in, out, batch, inverse, twiddle, stream);
    case 256: return launch_fft256(in, out, batch, inverse, twiddle, stream);
    default: return PHX_EUNSUPPORTED;
  }
}

int phx_scramble(const void* in, void* out, int64_t batch, int64_t n, int is_complex, int unscramble,
                 hipStream_t stream) {
  if (!in || !out || in == out || batch < 0 || batch > 65535) return PHX_EINVAL;
  if (batch == 0) return PHX_OK;
  switch (n) {
    case 16: return launch_scramble<16>(in, out, batch, is_complex, unscramble, stream);
    case 32: return launch_scramble<32>(in, out, batch, is_complex, unscramble, stream);
    case 64: return launch_scramble<64>(in, out, batch, is_complex, unscramble, stream);
    case 128: return launch_scramble<128>(in, out, batch, is_complex, unscramble, stream);
    case 192: return launch_scramble<192>(in, out, batch, is_complex, unscramble, stream);
    case 256: return launch_scramble<256>(in, out, batch, is_complex, unscramble, stream);
    default: return PHX_EUNSUPPORTED;
  }
}

int phx_check_finite_f32(const void* ptr, int64_t count, int32_t* flag, int32_t value, hipStream_t stream) {
  if (!ptr || !flag || count < 0) return PHX_EINVAL;
  if (count == 0) return PHX_OK;
  int64_t blocks = (count + 255) / 256;
  if (blocks > 2048) blocks = 2048;
  hipLaunchKernelGGL(check_finite_k, dim3((unsigned)blocks), dim3(256), 0, stream, (const float*)ptr, count, flag,
                     value);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_scale_c64(void* ptr, int64_t count, float scale, hipStream_t stream) {
  if (!ptr || count < 0) return PHX_EINVAL;
  if (count == 0) return PHX_OK;
  int64_t blocks = (count + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(scale_c64_k, dim3((unsigned)blocks), dim3(256), 0, stream, (float2*)ptr, count, scale);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

}  // extern "C"
