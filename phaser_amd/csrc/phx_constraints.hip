// Constraints and cost regularizers of phaser's engines (phaser/engines/common/
// regularizers.py) on gfx950: whole-object and whole-probe passes that run between group
// steps (group constraints, solvers.py:113, gradient/run.py:344-347), after an iteration
// (iter constraints, conventional/run.py:93, gradient/run.py:276-279), or inside the
// gradient engine's loss (cost regularizers, gradient/run.py:399-403).
//
// Every pass is HBM-bound elementwise or line work: one read and one write of the object
// per pass (the object filters: five passes, see below).
#include "phx_common.h"
#include "../../include/phaser_mi355x.h"

using namespace phx;

namespace {

// ------------------------------------------------------------------------------------
// ClampObjectAmplitude (regularizers.py:20-42): o * min(|o|, a) / |o|.  |o| = 0 gives
// 0/0 = NaN, as numpy does; such a NaN is recorded in `flag` (min with flag_value) when a
// flag is given.
__global__ void clamp_amplitude_k(float2* o, int64_t n, float a, int32_t* flag, int32_t flag_value) {
  bool bad = false;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float2 v = o[i];
    const float amp = hypotf(v.x, v.y);
    const float s = fminf(amp, a) / amp;
    const float2 r = make_float2(v.x * s, v.y * s);
    o[i] = r;
    bad |= !(isfinite(r.x) && isfinite(r.y));
  }
  if (flag && __ballot(bad) != 0ull && (threadIdx.x & 63) == 0) atomicMin(flag, flag_value);
}

// ------------------------------------------------------------------------------------
// RegularizeLayers (regularizers.py:86-132): per object pixel, the complex log of the S
// slices is convolved along the slice axis with a normalised Gaussian of 2r + 1 taps
// (scipy.ndimage.convolve1d, mode 'nearest': the edge slice repeats), exponentiated, and
// blended: o <- w exp(conv(log o)) + (1 - w) o.  The taps are summed in FP64 in tap
// order, as ndimage's line buffers do; log and exp in FP32.
constexpr int LAYERS_MAX_S = 128;
__global__ void layers_k(float2* obj, int S, int64_t HW, const float* taps, int r, float w, float w1) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= HW) return;
  float lre[LAYERS_MAX_S], lim[LAYERS_MAX_S];
  for (int s = 0; s < S; ++s) {
    const float2 v = obj[(int64_t)s * HW + e];
    lre[s] = logf(hypotf(v.x, v.y));
    lim[s] = atan2f(v.y, v.x);
  }
  for (int s = 0; s < S; ++s) {
    double are = 0.0, aim = 0.0;
    for (int j = 0; j <= 2 * r; ++j) {
      const int t = min(max(s + j - r, 0), S - 1);
      are += (double)taps[j] * (double)lre[t];
      aim += (double)taps[j] * (double)lim[t];
    }
    const float cre = (float)are, cim = (float)aim;
    const float m = expf(cre);
    float sn, cs;
    sincosf(cim, &sn, &cs);
    const int64_t k = (int64_t)s * HW + e;
    const float2 o = obj[k];
    const float2 nw = make_float2(w * (m * cs), w * (m * sn));
    obj[k] = make_float2(nw.x + w1 * o.x, nw.y + w1 * o.y);
  }
}

// ------------------------------------------------------------------------------------
// Cost regularizers of the gradient engine (regularizers.py:195-426) for the object:
// their loss (FP64, added to *loss) and their gradient in torch's convention
// (2 dL/dconj(O), added to grad).  c[] are the costs times the group's cost scale
// B / n_positions:
//   c[0] obj_l1       sum |O - 1|                 grad (O - 1)/|O - 1| (0 where O = 1)
//   c[1] obj_l2       sum |O - 1|^2               grad 2 (O - 1)
//   c[2] obj_tikh     sum |dx O|^2 + |dy O|^2     grad 2 (d_left - d_right) per axis
//   c[3] layers_tv    sum |O_{s+1} - O_s|         grad d_below/|d_below| - d_above/|d_above|
//   c[4] layers_tikh  sum |O_{s+1} - O_s|^2       grad 2 (d_below - d_above)
// A difference is counted in the loss once, by its first pixel.  want_grad = 0 gives the
// loss only.
struct CostCoefs {
  float c[5];
};

__device__ __forceinline__ float2 unit_or_zero(float2 d) {
  const float a = hypotf(d.x, d.y);
  return a > 0.f ? make_float2(d.x / a, d.y / a) : make_float2(0.f, 0.f);
}

__global__ void __launch_bounds__(256) cost_regularizers_k(const float2* obj, float2* grad, int S, int H, int W,
                                                           CostCoefs cc, double* loss, int want_grad) {
  __shared__ double red[4];
  const int64_t HW = (int64_t)H * W, tot = HW * S;
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < tot; i += (int64_t)gridDim.x * blockDim.x) {
    const int s = (int)(i / HW);
    const int64_t e = i - (int64_t)s * HW;
    const int y = (int)(e / W), x = (int)(e - (int64_t)y * W);
    const float2 o = obj[i];
    float gx = 0.f, gy = 0.f;
    const float2 d1 = make_float2(o.x - 1.f, o.y);
    if (cc.c[0] != 0.f) {
      acc += (double)cc.c[0] * (double)hypotf(d1.x, d1.y);
      const float2 u = unit_or_zero(d1);
      gx += cc.c[0] * u.x;
      gy += cc.c[0] * u.y;
    }
    if (cc.c[1] != 0.f) {
      acc += (double)cc.c[1] * (double)(d1.x * d1.x + d1.y * d1.y);
      gx += 2.f * cc.c[1] * d1.x;
      gy += 2.f * cc.c[1] * d1.y;
    }
    if (cc.c[2] != 0.f) {
      const float c = cc.c[2];
      if (x + 1 < W) {
        const float2 n = obj[i + 1];
        const float2 d = make_float2(n.x - o.x, n.y - o.y);
        acc += (double)c * (double)(d.x * d.x + d.y * d.y);
        gx -= 2.f * c * d.x;
        gy -= 2.f * c * d.y;
      }
      if (x > 0) {
        const float2 p = obj[i - 1];
        gx += 2.f * c * (o.x - p.x);
        gy += 2.f * c * (o.y - p.y);
      }
      if (y + 1 < H) {
        const float2 n = obj[i + W];
        const float2 d = make_float2(n.x - o.x, n.y - o.y);
        acc += (double)c * (double)(d.x * d.x + d.y * d.y);
        gx -= 2.f * c * d.x;
        gy -= 2.f * c * d.y;
      }
      if (y > 0) {
        const float2 p = obj[i - W];
        gx += 2.f * c * (o.x - p.x);
        gy += 2.f * c * (o.y - p.y);
      }
    }
    if (cc.c[3] != 0.f || cc.c[4] != 0.f) {
      if (s + 1 < S) {
        const float2 n = obj[i + HW];
        const float2 d = make_float2(n.x - o.x, n.y - o.y);
        if (cc.c[3] != 0.f) {
          acc += (double)cc.c[3] * (double)hypotf(d.x, d.y);
          const float2 u = unit_or_zero(d);
          gx -= cc.c[3] * u.x;
          gy -= cc.c[3] * u.y;
        }
        if (cc.c[4] != 0.f) {
          acc += (double)cc.c[4] * (double)(d.x * d.x + d.y * d.y);
          gx -= 2.f * cc.c[4] * d.x;
          gy -= 2.f * cc.c[4] * d.y;
        }
      }
      if (s > 0) {
        const float2 p = obj[i - HW];
        const float2 d = make_float2(o.x - p.x, o.y - p.y);
        if (cc.c[3] != 0.f) {
          const float2 u = unit_or_zero(d);
          gx += cc.c[3] * u.x;
          gy += cc.c[3] * u.y;
        }
        if (cc.c[4] != 0.f) {
          gx += 2.f * cc.c[4] * d.x;
          gy += 2.f * cc.c[4] * d.y;
        }
      }
    }
    if (want_grad) {
      const float2 g = grad[i];
      grad[i] = make_float2(g.x + gx, g.y + gy);
    }
  }
  acc = block_sum_f64(acc, red);
  if (threadIdx.x == 0 && loss) atomic_add_f64(loss, acc);
}

// ------------------------------------------------------------------------------------
// Object filters (ObjLowPass, regularizers.py:135-161; ObjGaussian, :164-192):
// O_s <- ifft2(F fft2(O_s)) for every slice, F an H x W real filter on the corner-origin
// frequency grid.  The (i)fftshift of phaser's convention (num.py:434-468) cancels in a
// product with a filter, so this is the plain DFT filter IDFT(F DFT(O)).  Object sides
// are arbitrary (518, 1030, 1540, ...), so every 1-D DFT of length N is Bluestein's:
//   X[m] = w[m] sum_k (x[k] w[k]) conj(w[m - k]),  w[k] = exp(-i pi k^2 / N),
// a circular convolution of length L = 2^p >= 2N - 1, done as FFT_L, product with the
// host-computed spectrum bhat = FFT_L(conj w) / L, inverse FFT_L.  One workgroup owns one
// line in LDS (L <= 4096: 32 KiB, sides up to 2048): the FFT is an in-place radix-2 decimation in time
// (bit-reversed load), the inverse a decimation in frequency (bit-reversed result).  The
// inverse DFT of a line is conj(DFT(conj x)) / N.
//
// Passes per filter application: DFT of every row (in place), transpose to (S, W, H),
// DFT -> filter -> inverse DFT of every column line (in place), transpose back, inverse
// DFT of every row (in place): 5 reads and 5 writes of the object.
constexpr int BS_MAX_L = 4096;
constexpr int BS_T = 256;

struct BluesteinTables {
  const float2* chirp;  // (N) w[k], FP64-rounded
  const float2* bhat;   // (L) FFT_L(conj w) / L
  const float2* tw;     // (L/2) exp(-2 pi i j / L)
  int n, l, logl;
};

__device__ __forceinline__ unsigned bitrev(unsigned v, int bits) { return __brev(v) >> (32 - bits); }

// lds <- FFT_L(lds) for data stored bit-reversed; natural-order result
__device__ __forceinline__ void lds_fft_dit(float2* lds, const float2* tw, int L, int logL) {
  for (int st = 1; st <= logL; ++st) {
    const int half = 1 << (st - 1), span = half << 1, tstep = L >> st;
    for (int j = threadIdx.x; j < (L >> 1); j += blockDim.x) {
      const int grp = j >> (st - 1), pos = j & (half - 1);
      const int i0 = grp * span + pos, i1 = i0 + half;
      const float2 w = tw[pos * tstep];
      const float2 u = lds[i0], t = cmul(lds[i1], w);
      lds[i0] = cadd(u, t);
      lds[i1] = csub(u, t);
    }
    __syncthreads();
  }
}

// lds <- unnormalised inverse FFT_L(lds) for natural-order data; bit-reversed result
__device__ __forceinline__ void lds_ifft_dif(float2* lds, const float2* tw, int L, int logL) {
  for (int st = logL; st >= 1; --st) {
    const int half = 1 << (st - 1), span = half << 1, tstep = L >> st;
    for (int j = threadIdx.x; j < (L >> 1); j += blockDim.x) {
      const int grp = j >> (st - 1), pos = j & (half - 1);
      const int i0 = grp * span + pos, i1 = i0 + half;
      const float2 w = tw[pos * tstep];
      const float2 u = lds[i0], v = lds[i1];
      lds[i0] = cadd(u, v);
      lds[i1] = cmulc(csub(u, v), w);  // (u - v) conj(w)
    }
    __syncthreads();
  }
}

// DFT of the N values v(k) (conjugated in and out with CONJ: the inverse DFT times N),
// result handed to out(m, X).  Every thread of the block calls it.
template <bool CONJ, class In, class Out>
__device__ __forceinline__ void bluestein_dft(float2* lds, const BluesteinTables& t, In in, Out out) {
  const int N = t.n, L = t.l, logL = t.logl;
  for (int k = threadIdx.x; k < L; k += blockDim.x) {
    float2 a = make_float2(0.f, 0.f);
    if (k < N) {
      float2 v = in(k);
      if (CONJ) v.y = -v.y;
      a = cmul(v, t.chirp[k]);
    }
    lds[bitrev((unsigned)k, logL)] = a;
  }
  __syncthreads();
  lds_fft_dit(lds, t.tw, L, logL);
  for (int k = threadIdx.x; k < L; k += blockDim.x) lds[k] = cmul(lds[k], t.bhat[k]);
  __syncthreads();
  lds_ifft_dif(lds, t.tw, L, logL);
  for (int m = threadIdx.x; m < N; m += blockDim.x) {
    float2 X = cmul(lds[bitrev((unsigned)m, logL)], t.chirp[m]);
    if (CONJ) X.y = -X.y;
    out(m, X);
  }
  __syncthreads();  // lds is reused by the caller's next transform
}

// MODE 0: forward DFT of each line; 1: inverse DFT (1/N); 2: forward, times filt[line %
// lines_per_slice][.], inverse (1/N).  Lines are contiguous of length t.n.
template <int MODE>
__global__ void __launch_bounds__(BS_T) bluestein_lines_k(float2* data, BluesteinTables t, const float* filt,
                                                          int64_t lines_per_slice) {
  extern __shared__ float2 lds[];
  const int64_t line = blockIdx.x;
  float2* x = data + line * t.n;
  const float invn = 1.0f / (float)t.n;
  if (MODE == 0) {
    bluestein_dft<false>(lds, t, [&](int k) { return x[k]; }, [&](int m, float2 X) { x[m] = X; });
  } else if (MODE == 1) {
    bluestein_dft<true>(lds, t, [&](int k) { return x[k]; }, [&](int m, float2 X) { x[m] = cscale(X, invn); });
  } else {
    // the forward DFT leaves c[m] = lds[bitrev m] with X[m] = w[m] c[m]; the inverse DFT's
    // input conj(f[m] X[m]) w[m] goes to the same slot bitrev m (zero for m >= N), so the
    // filtered spectrum never leaves LDS
    const float* f = filt + (line % lines_per_slice) * t.n;
    const int N = t.n, L = t.l, logL = t.logl;
    for (int k = threadIdx.x; k < L; k += blockDim.x) {
      float2 v = k < N ? x[k] : make_float2(0.f, 0.f);
      if (k < N) v = cmul(v, t.chirp[k]);
      lds[bitrev((unsigned)k, logL)] = v;
    }
    __syncthreads();
    lds_fft_dit(lds, t.tw, L, logL);
    for (int k = threadIdx.x; k < L; k += blockDim.x) lds[k] = cmul(lds[k], t.bhat[k]);
    __syncthreads();
    lds_ifft_dif(lds, t.tw, L, logL);
    for (int k = threadIdx.x; k < L; k += blockDim.x) {
      const int slot = bitrev((unsigned)k, logL);
      float2 a = make_float2(0.f, 0.f);
      if (k < N) {
        float2 Y = cscale(cmul(lds[slot], t.chirp[k]), f[k]);
        Y.y = -Y.y;
        a = cmul(Y, t.chirp[k]);
      }
      lds[slot] = a;
    }
    __syncthreads();
    lds_fft_dit(lds, t.tw, L, logL);
    for (int k = threadIdx.x; k < L; k += blockDim.x) lds[k] = cmul(lds[k], t.bhat[k]);
    __syncthreads();
    lds_ifft_dif(lds, t.tw, L, logL);
    for (int m = threadIdx.x; m < N; m += blockDim.x) {
      float2 X = cmul(lds[bitrev((unsigned)m, logL)], t.chirp[m]);
      x[m] = make_float2(X.x * invn, -X.y * invn);
    }
  }
}

// out[s][x][y] = in[s][y][x] through 32 x 32 LDS tiles (padded: no bank conflicts)
__global__ void __launch_bounds__(256) transpose_k(const float2* in, float2* out, int H, int W) {
  __shared__ float2 tile[32][33];
  const int s = blockIdx.z;
  const int64_t off = (int64_t)s * H * W;
  const int x0 = blockIdx.x * 32, y0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int r = ty; r < 32; r += 8) {
    const int y = y0 + r, x = x0 + tx;
    if (y < H && x < W) tile[r][tx] = in[off + (int64_t)y * W + x];
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int x = x0 + r, y = y0 + tx;
    if (y < H && x < W) out[off + (int64_t)x * H + y] = tile[tx][r];
  }
}

bool tables_ok(const BluesteinTables& t) {
  if (!t.chirp || !t.bhat || !t.tw || t.n < 1 || t.l > BS_MAX_L || t.l < 2 || (1 << t.logl) != t.l) return false;
  return t.l >= 2 * t.n - 1;
}

template <int MODE>
int launch_lines(float2* data, int64_t lines, const BluesteinTables& t, const float* filt, int64_t per_slice,
                 hipStream_t stream) {
  for (int64_t l0 = 0; l0 < lines; l0 += 0x7fffffffLL) {
    const int64_t cnt = lines - l0 < 0x7fffffffLL ? lines - l0 : 0x7fffffffLL;
    const size_t lds = (size_t)t.l * sizeof(float2);
    hipLaunchKernelGGL((bluestein_lines_k<MODE>), dim3((unsigned)cnt), dim3(BS_T), lds, stream, data + l0 * t.n, t,
                       filt, per_slice);
    PHX_CHECK_LAUNCH();
  }
  return PHX_OK;
}

int64_t grid_for(int64_t n, int64_t cap) {
  int64_t b = (n + 255) / 256;
  return b < 1 ? 1 : (b > cap ? cap : b);
}

}  // namespace

extern "C" {

int phx_clamp_amplitude_c64(void* obj, int64_t count, float amplitude, int32_t* flag, int32_t flag_value,
                            hipStream_t stream) {
  if (!obj || count < 0) return PHX_EINVAL;
  if (count == 0) return PHX_OK;
  hipLaunchKernelGGL(clamp_amplitude_k, dim3((unsigned)grid_for(count, 8192)), dim3(256), 0, stream, (float2*)obj,
                     count, amplitude, flag, flag_value);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_regularize_layers_c64(void* obj, int64_t n_slices, int64_t obj_h, int64_t obj_w, const float* taps,
                              int64_t radius, float weight, float one_minus_weight, hipStream_t stream) {
  if (!obj || !taps || n_slices < 1 || n_slices > LAYERS_MAX_S || obj_h < 1 || obj_w < 1 || radius < 0 ||
      radius > 4096)
    return PHX_EINVAL;
  const int64_t HW = obj_h * obj_w;
  hipLaunchKernelGGL(layers_k, dim3((unsigned)((HW + 255) / 256)), dim3(256), 0, stream, (float2*)obj,
                     (int)n_slices, HW, taps, (int)radius, weight, one_minus_weight);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_cost_regularizers_c64(const void* obj, void* grad, int64_t n_slices, int64_t obj_h, int64_t obj_w,
                              const float* coefs, double* loss, int want_grad, hipStream_t stream) {
  if (!obj || !coefs || n_slices < 1 || obj_h < 1 || obj_w < 1 || (want_grad && !grad)) return PHX_EINVAL;
  CostCoefs cc;
  for (int k = 0; k < 5; ++k) cc.c[k] = coefs[k];
  const int64_t tot = n_slices * obj_h * obj_w;
  hipLaunchKernelGGL(cost_regularizers_k, dim3((unsigned)grid_for(tot, 4096)), dim3(256), 0, stream,
                     (const float2*)obj, (float2*)grad, (int)n_slices, (int)obj_h, (int)obj_w, cc, loss, want_grad);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_object_filter_c64(void* obj, void* work, int64_t n_slices, int64_t obj_h, int64_t obj_w,
                          const float* filt_t, const void* chirp_h, const void* bhat_h, const void* tw_h,
                          int64_t l_h, const void* chirp_w, const void* bhat_w, const void* tw_w, int64_t l_w,
                          hipStream_t stream) {
  if (!obj || !work || obj == work || !filt_t || n_slices < 1 || obj_h < 1 || obj_w < 1 || n_slices > 65535)
    return PHX_EINVAL;
  auto tables = [](const void* c, const void* b, const void* tw, int64_t n, int64_t l) {
    BluesteinTables t;
    t.chirp = (const float2*)c;
    t.bhat = (const float2*)b;
    t.tw = (const float2*)tw;
    t.n = (int)n;
    t.l = (int)l;
    int lg = 0;
    while ((1LL << lg) < l && lg < 30) ++lg;
    t.logl = lg;
    return t;
  };
  const BluesteinTables th = tables(chirp_h, bhat_h, tw_h, obj_h, l_h);
  const BluesteinTables tw = tables(chirp_w, bhat_w, tw_w, obj_w, l_w);
  if (l_h > BS_MAX_L || l_w > BS_MAX_L) return PHX_EUNSUPPORTED;
  if (!tables_ok(th) || !tables_ok(tw)) return PHX_EINVAL;
  float2* o = (float2*)obj;
  float2* wk = (float2*)work;
  const int S = (int)n_slices, H = (int)obj_h, W = (int)obj_w;
  int r = launch_lines<0>(o, (int64_t)S * H, tw, nullptr, 1, stream);  // DFT of every row
  if (r) return r;
  const dim3 tg((unsigned)((W + 31) / 32), (unsigned)((H + 31) / 32), (unsigned)S);
  hipLaunchKernelGGL(transpose_k, tg, dim3(256), 0, stream, (const float2*)o, wk, H, W);
  PHX_CHECK_LAUNCH();
  // column lines of length H; line (s, kx) takes filter row kx of the transposed filter
  r = launch_lines<2>(wk, (int64_t)S * W, th, filt_t, W, stream);
  if (r) return r;
  const dim3 tg2((unsigned)((H + 31) / 32), (unsigned)((W + 31) / 32), (unsigned)S);
  hipLaunchKernelGGL(transpose_k, tg2, dim3(256), 0, stream, (const float2*)wk, o, W, H);
  PHX_CHECK_LAUNCH();
  return launch_lines<1>(o, (int64_t)S * H, tw, nullptr, 1, stream);  // inverse DFT of every row
}

}  // extern "C"
