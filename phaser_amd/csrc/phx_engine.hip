// Fused plane kernels for one group step of phaser's conventional engines (LSQ-ML and
// ePIE) on gfx950.  One workgroup owns one (position, mode) plane of n x n complex64 in
// LDS; the object cutout gather, the subpixel-shift filter, the slice transmission, the
// Fresnel propagator, the modulus constraint and the update reductions are applied
// while the plane passes through the FFT stages (phx_fft.h), so the stored waves are
// read and written once per slice step.
//
// Reference: phaser/engines/conventional/solvers.py:131-290 (LSQ-ML), :386-504 (ePIE),
// phaser/engines/common/simulation.py:241-306, noise_models.py:47-63.
#include "phx_common.h"
#include "phx_fft.h"
#include "../../include/phaser_mi355x.h"

#include <algorithm>
#include <climits>
#include <mutex>

using namespace phx;

namespace {

constexpr float TWO_PI = 6.283185307179586f;

struct RedLayout {
  int num_obj, den_obj, numP, denP, norm, pos_num, pos_den, loss, err, total;
};
__host__ __device__ inline RedLayout red_layout(int S, int bmax, int M) {
  RedLayout r;
  int o = 0;
  r.num_obj = o; o += S * bmax;
  r.den_obj = o; o += S;
  r.numP = o;    o += bmax * M;
  r.denP = o;    o += 1;
  r.norm = o;    o += bmax;
  r.pos_num = o; o += 2 * bmax;
  r.pos_den = o; o += 2;
  r.loss = o;    o += 1;
  r.err = o;     o += 8 * bmax;  // 256 x 256 planes: per-slot error partials of the 8 column tiles
  r.total = o;
  return r;
}

// params[] slots; P_VAR = 1 + gaussian_variance of the amplitude loss, P_LOSS_SCALE the
// upstream gradient of the loss (autograd's grad_output)
enum Param { P_BETA_OBJ = 0, P_BETA_PROBE, P_REG_OBJ, P_REG_PROBE, P_GAMMA, P_OFFSET, P_EPS, P_VAR, P_LOSS_SCALE };

// far-field modes of farfield_update_k
enum FarMode { FM_MODULUS = 0, FM_GRAD_AMPLITUDE = 1, FM_GRAD_POISSON = 2 };

struct Cut {
  int y0, x0, H, W;
  __device__ __forceinline__ bool inb(int y, int x) const {
    // unsigned compares joined without short-circuit: no branches in unrolled loops
    return ((unsigned)(y0 + y) < (unsigned)H) & ((unsigned)(x0 + x) < (unsigned)W);
  }
  __device__ __forceinline__ int idx(int y, int x) const { return (y0 + y) * W + (x0 + x); }
  // the n x n cutout lies inside the object (uniform per plane): no clamping needed
  __device__ __forceinline__ bool interior(int n) const {
    return y0 >= 0 && x0 >= 0 && y0 + n <= H && x0 + n <= W;
  }
};

__device__ __forceinline__ Cut cut_of(const phx_group_desc& d, int p) {
  return Cut{d.start[2 * p], d.start[2 * p + 1], d.obj_h, d.obj_w};
}

// Object value under the cutout; out of bounds gives 1 (cupy default, object.py:509).
// Branch-free: the load goes to the clamped (always valid) address and the result is
// selected, so an unrolled loop over a row stays one basic block the scheduler can pace.
__device__ __forceinline__ float2 obj_at(const float2* o, const Cut& c, int y, int x) {
  const int gy = min(max(c.y0 + y, 0), c.H - 1), gx = min(max(c.x0 + x, 0), c.W - 1);
  const float2 v = o[gy * c.W + gx];
  return c.inb(y, x) ? v : make_float2(1.f, 0.f);
}

// exp(-2 pi i (sx * kx + sy * ky)) in float32 (optics.py:151-170).  The phase is taken
// in half-turns (sincospi: exact integer range reduction, no large-argument path); it
// differs from exp of the float32-rounded product -2 pi t by that rounding only (~1e-7).
__device__ __forceinline__ float2 shift_filter(float sy, float sx, float ky, float kx) {
  const float t = sx * kx + sy * ky;
  float sn, cs;
  sincospif(-2.0f * t, &sn, &cs);
  return make_float2(cs, sn);
}

// Propagator j of plane slot b: the shared props[j], or with per-position tilts the slot's
// tilted copy in tprops (phx_tilt_propagators fills it for the group).  d.tilt is a
// kernel argument, so the choice is uniform and costs no vector registers.
__device__ __forceinline__ const float2* prop_of(const phx_group_desc& d, int j, int b) {
  const size_t NN = (size_t)d.n * d.n;
  return d.tilt != nullptr ? (const float2*)d.tprops + ((size_t)j * d.bmax + b) * NN
                           : (const float2*)d.props + (size_t)j * NN;
}
// The same for a whole slice loop: plane slot b's first propagator and the distance
// between consecutive ones, taken once before the loop.
struct PropSeq {
  const float2* base;
  size_t stride;
  __device__ __forceinline__ const float2* operator[](int j) const { return base + (size_t)j * stride; }
};
__device__ __forceinline__ PropSeq prop_seq(const phx_group_desc& d, int b) {
  const size_t NN = (size_t)d.n * d.n;
  return d.tilt != nullptr ? PropSeq{(const float2*)d.tprops + (size_t)b * NN, (size_t)d.bmax * NN}
                           : PropSeq{(const float2*)d.props, NN};
}

// Model intensity sum_m |F_m|^2 / n^2 at spectrum point e of slot b.  The forward kernels
// leave each mode's term with a plain store at intensity[(b n^2 + e) M + m] (the M terms of
// a point side by side), and the terms are added here in mode order, as numpy's sum over
// the mode axis adds them: the result does not depend on the order in which the modes'
// workgroups ran.
__device__ __forceinline__ float* intensity_term(const phx_group_desc& d, int b, int e, int m) {
  return d.intensity + ((size_t)b * d.n * d.n + e) * d.n_modes + m;
}
__device__ __forceinline__ float model_intensity(const phx_group_desc& d, int b, int e) {
  const float* I = intensity_term(d, b, e, 0);
  float s = I[0];
  for (int k = 1; k < d.n_modes; ++k) s += I[k];
  return s;
}

// model_intensity for up to four modes, as a per-point functor whose loads go out with
// the caller's other loads of the point (no separate round trip): with four modes one
// 16-byte load, otherwise up to four loads within the point's 16 bytes (a mode past M
// re-reads mode 0 and is not added).  The adds keep mode order.
struct ModeSum4 {
  const float* base;
  int M;
  __device__ __forceinline__ ModeSum4(const phx_group_desc& d, int b) {
    base = intensity_term(d, b, 0, 0);
    M = d.n_modes;
  }
  __device__ __forceinline__ float operator()(int e) const {
    if (M == 4) {
      const float4 v = *reinterpret_cast<const float4*>(base + 4 * (size_t)e);
      return ((v.x + v.y) + v.z) + v.w;
    }
    const float* I = base + (size_t)e * M;
    const float a0 = I[0], a1 = I[M > 1 ? 1 : 0], a2 = I[M > 2 ? 2 : 0], a3 = I[M > 3 ? 3 : 0];
    float s = a0;
    s = M > 1 ? s + a1 : s;
    s = M > 2 ? s + a2 : s;
    s = M > 3 ? s + a3 : s;
    return s;
  }
};

// model_intensity at every column-phase point of this lane (register k: py = QL k + q,
// px = rc), for more than four modes: each mode's loads for all the points together.
template <int N>
__device__ __forceinline__ void model_intensity_cols(const phx_group_desc& d, int b, int q, int rc,
                                                     float (&Iv)[Plane<N>::E]) {
  constexpr int E = Plane<N>::E, QL = Plane<N>::QL;
  const int M = d.n_modes;
  const float* I = intensity_term(d, b, q * N + rc, 0);
#pragma unroll
  for (int k = 0; k < E; ++k) Iv[k] = I[(size_t)k * QL * N * M];
  for (int mm = 1; mm < M; ++mm) {
#pragma unroll
    for (int k = 0; k < E; ++k) Iv[k] += I[(size_t)k * QL * N * M + mm];
  }
}

#define PHX_PLANE_SETUP(N, T)                                                    \
  static_assert(T == Plane<N>::T, "plane kernels run one workgroup of Plane<N>::T threads per plane");               \
  __shared__ float2 lds_[Plane<N>::LDS_FLOAT2 + N];                              \
  [[maybe_unused]] __shared__ double red_lds_[T / 64];                           \
  Plane<N> pl(lds_, lds_ + Plane<N>::LDS_FLOAT2);                                \
  load_twiddles<N>(lds_ + Plane<N>::LDS_FLOAT2, (const float2*)d.twiddle);       \
  __syncthreads();

// The same in two halves, for kernels that run one of several plane bodies: the kernel
// declares the LDS (one declaration shared by the bodies), the body binds the plane.
#define PHX_PLANE_LDS(N, T)                                                      \
  static_assert(T == Plane<N>::T, "plane kernels run one workgroup of Plane<N>::T threads per plane");               \
  __shared__ float2 lds_[Plane<N>::LDS_FLOAT2 + N];                              \
  [[maybe_unused]] __shared__ double red_lds_[T / 64];
#define PHX_PLANE_BIND(N)                                                        \
  Plane<N> pl(lds_, lds_ + Plane<N>::LDS_FLOAT2);                                \
  load_twiddles<N>(lds_ + Plane<N>::LDS_FLOAT2, (const float2*)d.twiddle);       \
  __syncthreads();

__device__ __forceinline__ bool slot_active(const phx_group_desc& d, int b) { return b < *d.group_size; }

// Tilted propagators of the group's slots (simulation.py:193-222): tprops[j][b][e] =
// props[j][e] exp(2 pi i dz_j (ty_p ky + tx_p kx)) at scrambled point e = py n + px, with the
// tangents of slot b's position p (FP64 on the host, float32 here).  The phase is taken
// in half-turns (sincospif: exact range reduction).  Grid (n^2 / 256, S - 1, bmax).
__global__ void __launch_bounds__(256) tilt_props_k(phx_group_desc d) {
  const int b = blockIdx.z, j = blockIdx.y;
  if (!slot_active(d, b)) return;
  const int NN = d.n * d.n;
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= NN) return;
  const int p = d.group_pos[b];
  const float ty = d.tilt[2 * p], tx = d.tilt[2 * p + 1];
  const float dz2 = 2.0f * d.slice_dz[j];
  const float h = dz2 * fmaf(ty, d.ky_s[e / d.n], tx * d.kx_s[e % d.n]);
  float sn, cs;
  sincospif(h, &sn, &cs);
  const float2 v = ((const float2*)d.props)[(size_t)j * NN + e];
  ((float2*)d.tprops)[((size_t)j * d.bmax + b) * NN + e] = cmul(v, make_float2(cs, sn));
}

// Plane (b, m) of this workgroup in a sweep launch (grid n_modes x bmax).  Workgroups go
// to the 8 XCDs round-robin by linear id L, so with bmax a multiple of 8 the planes are
// placed at L = 8 (M floor(b/8) + m) + b mod 8: every mode of one position runs on the
// same XCD and the position's object cutouts are fetched into one L2, not M of them.
__device__ __forceinline__ void sweep_plane(const phx_group_desc& d, int& b, int& m) {
  const int M = d.n_modes;
  if ((d.bmax & 7) == 0) {
    const int L = blockIdx.y * M + blockIdx.x;
    const int slot = L >> 3;
    m = slot % M;
    b = (slot / M) * 8 + (L & 7);
  } else {
    m = blockIdx.x;
    b = blockIdx.y;
  }
}

// Offset of the stored wave of slice i, slot b, mode m.  With psi_slots < S only psi_0
// is kept permanently and slices 1.. alternate between two further planes, so slice i
// is recomputed from psi_0 when the adjoint needs it.
__device__ __forceinline__ size_t psi_off(const phx_group_desc& d, int i, int b, int m) {
  const bool full = d.psi_slots <= 0 || d.psi_slots >= d.n_slices;
  const int slot = full ? i : (i == 0 ? 0 : 1 + ((i - 1) & 1));
  return (((size_t)slot * d.bmax + b) * d.n_modes + m) * (size_t)d.n * d.n;
}

// ------------------------------------------------------------------------------------
template <int N, int T>
__global__ void __launch_bounds__(T) probe_spectra_k(phx_group_desc d) {
  PHX_PLANE_SETUP(N, T);
  constexpr int NN = N * N;
  const int m = blockIdx.x;
  const float2* P = (const float2*)d.probe + (size_t)m * NN;
  float2* Pk = (float2*)d.probe_k + (size_t)m * NN;
  pl.forward([&](int y, int x) { return P[y * N + x]; },
             [&](int py, int px, float2& v) { Pk[py * N + px] = v; });
}

// gom = sum_b |prod_s O_s|^2 over the group's cutouts (solvers.py:208), in a fixed order:
// a workgroup owns 64 pixels; its wave w sums the positions b = w, w + GOM_WAVES, .. in
// increasing b, and the GOM_WAVES partials are added in wave order.  No atomics, so gom
// (and the probe step that divides by it) does not depend on scheduling.  Grid n*n/64.
constexpr int GOM_WAVES = 8;
__global__ void __launch_bounds__(64 * GOM_WAVES) group_obj_mag_k(phx_group_desc d) {
  __shared__ float part[GOM_WAVES][64];
  const int n = d.n, NN = n * n;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int e = blockIdx.x * 64 + lane;
  const int gs = *d.group_size;
  const float2* obj = (const float2*)d.object;
  const size_t HW = (size_t)d.obj_h * d.obj_w;
  constexpr int K = 4;  // positions per pass, their loads in flight together
  float acc = 0.f;
  if (e < NN) {
    const int y = e / n, x = e % n;
    for (int b0 = w; b0 < gs; b0 += K * GOM_WAVES) {
      Cut c[K];
      float2 prod[K];
#pragma unroll
      for (int u = 0; u < K; ++u) {
        const int b = b0 + u * GOM_WAVES;
        c[u] = cut_of(d, d.group_pos[b < gs ? b : w]);
        prod[u] = obj_at(obj, c[u], y, x);
      }
      for (int s = 1; s < d.n_slices; ++s) {
#pragma unroll
        for (int u = 0; u < K; ++u) prod[u] = cmul(prod[u], obj_at(obj + s * HW, c[u], y, x));
      }
      // positions b0, b0 + GOM_WAVES, .. in increasing order
#pragma unroll
      for (int u = 0; u < K; ++u)
        if (b0 + u * GOM_WAVES < gs) acc += abs2f(prod[u]);
    }
  }
  part[w][lane] = acc;
  __syncthreads();
  if (w == 0 && e < NN) {
    float t = part[0][lane];
#pragma unroll
    for (int k = 1; k < GOM_WAVES; ++k) t += part[k][lane];
    d.gom[e] = t;
  }
}

// psi0 = ifft2(fft2(P) * filt_b), illumination, norm, optional slice-0 step.
template <int N, int T, bool ILLUM>
__global__ void __launch_bounds__(T) shift_probes_k(phx_group_desc d, int propagate) {
  const int m = blockIdx.x, b = blockIdx.y;
  if (!slot_active(d, b)) return;
  PHX_PLANE_SETUP(N, T);
  constexpr int NN = N * N;
  constexpr float inv_nn = 1.0f / (float)NN;
  const int M = d.n_modes, bmax = d.bmax;
  const int p = d.group_pos[b];
  const Cut c = cut_of(d, p);
  const float sy = d.shift[2 * p], sx = d.shift[2 * p + 1];
  const float2* Pk = (const float2*)d.probe_k + (size_t)m * NN;
  float2* psi0 = (float2*)d.psi + psi_off(d, 0, b, m);
  const float2* obj0 = (const float2*)d.object;
  float* gpm0 = d.gpm;
  double nrm = 0.0;

  constexpr int E = Plane<N>::E;
  float2 x[E];
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const int py = pl.QL * k + pl.q, px = pl.rc;
    x[k] = cmul(Pk[py * N + px], shift_filter(sy, sx, d.ky_s[py], d.kx_s[px]));
  }
  pl.inverse_regs(x);
  // x now holds psi0 (unscaled) at row rc, columns 4j + q
#pragma unroll
  for (int j = 0; j < E; ++j) {
    const int yy = pl.rc, xx = pl.QL * j + pl.q;
    const float2 v = cscale(x[j], inv_nn);
    psi0[yy * N + xx] = v;
    const float a = abs2f(v);
    nrm += a;
    const bool inb = c.inb(yy, xx);
    if (ILLUM && inb) atomic_add_f32(&gpm0[c.idx(yy, xx)], a);
    x[j] = cmul(v, obj_at(obj0, c, yy, xx));
  }
  const double tot = block_sum_f64(nrm, red_lds_);
  if (threadIdx.x == 0) {
    RedLayout L = red_layout(d.n_slices, bmax, M);
    atomic_add_f64(&d.red[L.norm + b], tot);
  }
  if (!propagate) return;
  const float2* prop = prop_of(d, 0, b);
  float2* psi1 = (float2*)d.psi + psi_off(d, 1, b, m);
  pl.convolve_regs(x, [&](int py, int px, float2 v) { return cmul(v, prop[py * N + px]); }, d.prop_zero_waves);
  pl.store_rows([&](int yy, int xx, float2 v) { psi1[yy * N + xx] = cscale(v, inv_nn); }, x);
}

template <int N, int T, bool ILLUM>
__global__ void __launch_bounds__(T) forward_slice_k(phx_group_desc d, int i) {
  const int m = blockIdx.x, b = blockIdx.y;
  if (!slot_active(d, b)) return;
  PHX_PLANE_SETUP(N, T);
  constexpr int NN = N * N;
  constexpr float inv_nn = 1.0f / (float)NN;
  const Cut c = cut_of(d, d.group_pos[b]);
  const size_t HW = (size_t)d.obj_h * d.obj_w;
  const float2* psi_i = (const float2*)d.psi + psi_off(d, i, b, m);
  float2* psi_o = (float2*)d.psi + psi_off(d, i + 1, b, m);
  const float2* obj_i = (const float2*)d.object + i * HW;
  float* gpm_i = d.gpm + i * HW;
  const float2* prop = prop_of(d, i, b);
  pl.convolve(
      [&](int y, int x) {
        const float2 v = psi_i[y * N + x];
        const bool inb = c.inb(y, x);
        if (ILLUM && inb) atomic_add_f32(&gpm_i[c.idx(y, x)], abs2f(v));
        return cmul(v, obj_at(obj_i, c, y, x));
      },
      [&](int py, int px, float2 v) { return cmul(v, prop[py * N + px]); },
      [&](int y, int x, float2 v) { psi_o[y * N + x] = cscale(v, inv_nn); });
}

// Exit wave spectrum and model intensity (sum over modes).
template <int N, int T, bool ILLUM>
__global__ void __launch_bounds__(T) farfield_fwd_k(phx_group_desc d) {
  const int m = blockIdx.x, b = blockIdx.y;
  if (!slot_active(d, b)) return;
  PHX_PLANE_SETUP(N, T);
  constexpr int NN = N * N;
  constexpr float inv_nn = 1.0f / (float)NN;
  const int M = d.n_modes, s = d.n_slices - 1;
  const Cut c = cut_of(d, d.group_pos[b]);
  const size_t HW = (size_t)d.obj_h * d.obj_w;
  const float2* psi_s = (const float2*)d.psi + psi_off(d, s, b, m);
  const float2* obj_s = (const float2*)d.object + s * HW;
  float* gpm_s = d.gpm + s * HW;
  float2* F = (float2*)d.farfield + ((size_t)b * M + m) * NN;
  float* I = intensity_term(d, b, 0, m);  // this mode's terms, stride M (model_intensity sums them)
  pl.forward(
      [&](int y, int x) {
        const float2 v = psi_s[y * N + x];
        const bool inb = c.inb(y, x);
        if (ILLUM && inb) atomic_add_f32(&gpm_s[c.idx(y, x)], abs2f(v));
        return cmul(v, obj_at(obj_s, c, y, x));
      },
      [&](int py, int px, float2& v) {
        const int e = py * N + px;
        F[e] = v;
        I[(size_t)e * M] = abs2f(v) * inv_nn;
      });
}

// Far-field step + inverse transform into chi_a.
//  FM_MODULUS: amplitude/Anscombe modulus constraint (noise_models.py:47-63) and, when
//    calc_error, errors[b] = sqrt(nansum((I - D)^2)) (solvers.py:237, FP64 sum);
//  FM_GRAD_AMPLITUDE / FM_GRAD_POISSON: gradient of the amplitude (noise_models.py:29-45)
//    or Poisson (:86-102) loss w.r.t. the exit wave, in torch's convention
//    (2 dL/dconj(a)), times params[P_LOSS_SCALE]; the group loss (FP64) goes to red.loss.
// Per-point far-field factor u (the update multiplies F by it) and the point's error or
// loss term t, for the modes above.
template <int MODE>
__device__ __forceinline__ float farfield_factor(float Iv, float Dv, float mk, float off, float eps, float var,
                                                 float gscale, float& t) {
  const float Dp = (Dv > 0.f || isnan(Dv)) ? Dv : 0.f;  // numpy.maximum keeps NaN
  float u;
  if (MODE == FM_MODULUS) {
    const float df = Iv - Dv;
    t = df * df;  // unmasked (solvers.py:237), NaN-skipping (nansum)
    t = isnan(t) ? 0.f : t;
    u = (sqrtf(Dp + off) / (sqrtf(Iv + off) + eps) - 1.0f) * mk;
  } else if (MODE == FM_GRAD_AMPLITUDE) {
    const float si = sqrtf(Iv + off);
    const float sd = sqrtf(Dp + off) - si - eps;
    t = mk * sd * sd;
    // at exactly zero model intensity every Psi_m is 0 and |Psi|/sqrt(I) <= 1: the
    // zero subgradient (autodiff of sqrt at 0 would give inf * 0 = NaN)
    u = si > 0.f ? -4.0f * mk * sd / (var * si) * gscale : 0.f;
  } else {
    t = mk * (Iv + eps + Dp * (logf(Dp + eps) - logf(Iv + eps) - 1.0f));
    u = 2.0f * mk * (1.0f - Dp / (Iv + eps)) * gscale;
  }
  return u;
}

// The group loss of the gradient modes from the block's FP32 partials (mode-0 planes
// only): FP64 block sum, amplitude scaling, one atomic per plane into red.loss.
template <int MODE>
__device__ __forceinline__ void add_group_loss(const phx_group_desc& d, float e2, double* red_lds) {
  double tot = block_sum_f64(e2, red_lds);
  if (MODE == FM_GRAD_AMPLITUDE) tot *= 2.0 / (double)d.params[P_VAR];
  if (threadIdx.x == 0) atomic_add_f64(&d.red[red_layout(d.n_slices, d.bmax, d.n_modes).loss], tot);
}

template <int N, int T, int MODE>
__global__ void __launch_bounds__(T) farfield_update_k(phx_group_desc d, int calc_error) {
  const int m = blockIdx.x, b = blockIdx.y;
  if (!slot_active(d, b)) return;
  PHX_PLANE_SETUP(N, T);
  constexpr int NN = N * N;
  constexpr float inv_nn = 1.0f / (float)NN;
  const int M = d.n_modes;
  const int p = d.group_pos[b];
  const float2* F = (const float2*)d.farfield + ((size_t)b * M + m) * NN;
  const float* D = d.patterns_s + (size_t)p * NN;
  const float* mask = d.mask_s;
  const float off = d.params[P_OFFSET], eps = d.params[P_EPS];
  float2* chi = (float2*)d.chi_a + ((size_t)b * M + m) * NN;
  // the error (modulus mode) or the loss (gradient modes) is summed by mode 0 only
  const bool err = MODE == FM_MODULUS ? (calc_error && m == 0) : (m == 0);
  const float var = d.params[P_VAR];
  const float gscale = MODE != FM_MODULUS ? d.params[P_LOSS_SCALE] : 1.0f;
  float e2 = 0.f;  // per-lane FP32 partial over its n/4 points; block sum in FP64
  pl.inverse(
      [&](int py, int px) {
        const int e = py * N + px;
        const float Iv = model_intensity(d, b, e), Dv = D[e], mk = mask[e];
        const float2 Fv = F[e];
        float t;
        const float u = farfield_factor<MODE>(Iv, Dv, mk, off, eps, var, gscale, t);
        e2 += t;
        return cscale(Fv, u);
      },
      [&](int y, int x, float2 v) { chi[y * N + x] = cscale(v, inv_nn); },
      [&] { __asm__ volatile("" : "+v"(e2)); });
  if (err) {
    if (MODE == FM_MODULUS) {
      const double tot = block_sum_f64(e2, red_lds_);
      if (threadIdx.x == 0) d.errors[b] = (float)sqrt(tot);
    } else {
      add_group_loss<MODE>(d, e2, red_lds_);
    }
  }
}

// chi entering the step of slice i, and the buffer the step writes chi_{i-1} to.  With
// chi_slices every slice keeps its chi (chi_{S-1} from the far field stays in chi_a);
// otherwise chi ping-pongs between chi_a and chi_b.
__device__ __forceinline__ float2* chi_in_buf(const phx_group_desc& d, int i) {
  if (d.chi_slices && i < d.n_slices - 1)
    return (float2*)d.chi_slices + (size_t)i * d.bmax * d.n_modes * d.n * d.n;
  return (float2*)(((d.n_slices - 1 - i) % 2 == 0) ? d.chi_a : d.chi_b);
}
__device__ __forceinline__ float2* chi_out_buf(const phx_group_desc& d, int i) {
  if (d.chi_slices) return (float2*)d.chi_slices + (size_t)(i - 1) * d.bmax * d.n_modes * d.n * d.n;
  return (float2*)(((d.n_slices - 1 - i) % 2 == 0) ? d.chi_b : d.chi_a);
}

// LSQ-ML backward step at slice i (solvers.py:242-271).  The variant is compile-time:
// BK_PROPAGATE (i > 0: back-propagate chi through slice i-1's propagator),
// BK_PROBE (i == 0 with a probe update: per-plane probe spectrum), BK_NONE (i == 0, no
// probe update: step-size sums only); UPD_OBJ adds the object step-size sums.  The
// row source is branch-free, so the unrolled load phase is one basic block.
enum BackKind { BK_PROPAGATE = 0, BK_PROBE = 1, BK_NONE = 2 };
template <int N, int T, int KIND, bool UPD_OBJ>
__global__ void __launch_bounds__(T) lsqml_backward_k(phx_group_desc d, int i) {
  const int m = blockIdx.x, b = blockIdx.y;
  if (!slot_active(d, b)) return;
  PHX_PLANE_SETUP(N, T);
  constexpr int NN = N * N;
  constexpr float inv_nn = 1.0f / (float)NN;
  constexpr bool PROBE = KIND == BK_PROBE;
  const int M = d.n_modes, bmax = d.bmax;
  const int p = d.group_pos[b];
  const Cut c = cut_of(d, p);
  const size_t HW = (size_t)d.obj_h * d.obj_w;
  const float2* chi = chi_in_buf(d, i) + ((size_t)b * M + m) * NN;
  const float2* psi_i = (const float2*)d.psi + psi_off(d, i, b, m);
  const float2* obj_i = (const float2*)d.object + i * HW;
  // per-lane partial sums over its n/4 points in FP32, block and group sums in FP64
  float numO = 0.f, denO = 0.f, numP = 0.f, denP = 0.f;

  // delta_O = chi conj(psi_i) is not scattered here: lsqml_object_update_k gathers it
  // (and the illumination sum |psi_i|^2) per object tile from chi and psi_i.
  auto src = [&](int y, int x) -> float2 {
    const float2 cv = chi[y * N + x];
    const float2 o = obj_at(obj_i, c, y, x);
    if (UPD_OBJ) {
      const float2 ps = psi_i[y * N + x];
      const float2 t = cmul(cmulc(cv, ps), ps);  // (chi conj(psi)) psi
      numO += cv.x * t.x + cv.y * t.y;
      denO += abs2f(t);
    }
    const float2 dP = cmulc(cv, o);  // chi * conj(O)
    if (PROBE) {
      const float2 t2 = cmul(dP, o);
      numP += cv.x * t2.x + cv.y * t2.y;
      denP += abs2f(t2);
    }
    return dP;
  };
  // keeps the sums in the load phase, one group of points at a time: sunk to their use
  // after the transform they would hold every point's inputs live and spill at n = 128
  auto pin = [&] {
    if (UPD_OBJ) __asm__ volatile("" : "+v"(numO), "+v"(denO));
    if (PROBE) __asm__ volatile("" : "+v"(numP), "+v"(denP));
  };

  if (KIND == BK_PROPAGATE) {
    float2* chi_o = chi_out_buf(d, i) + ((size_t)b * M + m) * NN;
    const float2* prop = prop_of(d, i - 1, b);
    pl.convolve(src, [&](int py, int px, float2 v) { return cmulc(v, prop[py * N + px]); },
                [&](int y, int x, float2 v) { chi_o[y * N + x] = cscale(v, inv_nn); }, pin);
  } else if (PROBE) {
    // per-plane fft2(dP) conj(filt_b) into the (free by now) farfield buffer; the sum
    // over positions is probe_spectrum_reduce_k
    const float sy = d.shift[2 * p], sx = d.shift[2 * p + 1];
    float2* Fp = (float2*)d.farfield + ((size_t)b * M + m) * NN;
    pl.forward(src, [&](int py, int px, float2& v) {
      Fp[py * N + px] = cmulc(v, shift_filter(sy, sx, d.ky_s[py], d.kx_s[px]));
    }, pin);
  } else {
    for (int e = threadIdx.x; e < NN; e += T) (void)src(e / N, e % N);
  }

  const RedLayout L = red_layout(d.n_slices, bmax, M);
  if (UPD_OBJ) {
    const double a = block_sum_f64(numO, red_lds_);
    const double bb = block_sum_f64(denO, red_lds_);
    if (threadIdx.x == 0) {
      atomic_add_f64(&d.red[L.num_obj + i * bmax + b], a);
      atomic_add_f64(&d.red[L.den_obj + i], bb);
    }
  }
  if (PROBE) {
    const double a = block_sum_f64(numP, red_lds_);
    const double bb = block_sum_f64(denP, red_lds_);
    if (threadIdx.x == 0) {
      d.red[L.numP + b * M + m] = a;
      atomic_add_f64(&d.red[L.denP], bb);
    }
  }
}

// ---------------------------------------------------------------------------------
// Fused group sweeps.  One workgroup carries one (position slot b, mode m) plane through
// every slice: the wave stays in registers between slice steps, so per step it reads
// only the slice's object cutout and propagator (both L2-resident across the group's
// planes) and writes the stored wave (or chi).  This replaces S+1 per-slice launches and
// the per-slice HBM reload of the wave.

// Forward model of a group (simulation.py:241-285; solvers.py:202-236, 431-457):
//   psi_0 = ifft2(fft2(P_m) filt_b), psi_{i+1} = ifft2(fft2(psi_i O_i) p_i),
//   farfield = FFT(psi_{S-1} O_{S-1}) (unnormalised), intensity += |farfield|^2 / n^2,
// storing every psi_i and the norm sum_m |psi_0|^2 of each slot.
// The body runs plane (b, m) on the workgroup's plane and reduction LDS (PHX_PLANE_LDS).
// At 128^2 it stays within 104 VGPRs: 4 waves per SIMD then leave 96, room for one
// group_obj_mag_k workgroup (8 waves of 48 VGPRs, 2 per SIMD) beside the sweep, which the
// LSQ-ML group step runs on a second stream.  At 112 VGPRs that workgroup waits for the
// sweep's CUs and then delays the backward sweep's workgroups (2.5% of an iteration).
template <int N, int T>
__device__ __forceinline__ void group_forward_plane(const phx_group_desc& d, int b, int m, float2* lds_,
                                                    double* red_lds_) {
  PHX_PLANE_BIND(N);
  constexpr int NN = N * N;
  constexpr int E = Plane<N>::E;
  constexpr float inv_nn = 1.0f / (float)NN;
  const int M = d.n_modes, S = d.n_slices;
  const int p = d.group_pos[b];
  const Cut c = cut_of(d, p);
  const size_t HW = (size_t)d.obj_h * d.obj_w;
  const float sy = d.shift[2 * p], sx = d.shift[2 * p + 1];
  const float2* Pk = (const float2*)d.probe_k + (size_t)m * NN;
  const bool inside = c.interior(N);
  const PropSeq props = prop_seq(d, b);
  float nrm = 0.f;  // per-lane FP32 partial of sum |psi_0|^2, FP64 block sum
  auto no_pin = [] {};

  float2 x[E];
  pl.cols_apply(x, [&](int py, int px, float2) {
    return cmul(Pk[py * N + px], shift_filter(sy, sx, d.ky_s[py], d.kx_s[px]));
  }, no_pin);
  pl.inverse_regs(x);
  for (int i = 0;; ++i) {
    pl.refresh();
    // x: n^2 psi_i in row phase (natural order)
    float2* psi_i = (float2*)d.psi + psi_off(d, i, b, m);
    const float2* obj_i = (const float2*)d.object + i * HW;
    const bool first = i == 0;
    // an interior cutout reads its row at one base plus immediate offsets
    auto step = [&](auto interior_tag) __attribute__((always_inline)) {
      constexpr bool IN = decltype(interior_tag)::value;
      const float2* orow = obj_i + (size_t)(c.y0 + pl.rc) * c.W + c.x0;
      pl.rows_apply(x, [&](int yy, int xx, float2 v) {
        v = cscale(v, inv_nn);
        psi_i[yy * N + xx] = v;
        if (first) nrm += abs2f(v);
        return cmul(v, IN ? orow[xx] : obj_at(obj_i, c, yy, xx));
      }, [&] { __asm__ volatile("" : "+v"(nrm)); });
    };
    if (inside) step(std::true_type{});
    else step(std::false_type{});
    if (i == S - 1) break;
    pl.template convolve_midfetch<false>(x, props[i], d.prop_zero_waves);
  }
  pl.forward_regs(x);
  float2* F = (float2*)d.farfield + ((size_t)b * M + m) * NN;
  float* I = intensity_term(d, b, 0, m);  // this mode's terms, stride M (model_intensity sums them)
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const int e = (pl.QL * k + pl.q) * N + pl.rc;
    F[e] = x[k];
    I[(size_t)e * M] = abs2f(x[k]) * inv_nn;
  }
  const double tot = block_sum_f64(nrm, red_lds_);
  if (threadIdx.x == 0) atomic_add_f64(&d.red[red_layout(S, d.bmax, M).norm + b], tot);
}

template <int N, int T>
__global__ void __launch_bounds__(T) group_forward_k(phx_group_desc d) {
  int b, m;
  sweep_plane(d, b, m);
  if (!slot_active(d, b)) return;
  PHX_PLANE_LDS(N, T);
  group_forward_plane<N, T>(d, b, m, lds_, red_lds_);
}

// LSQ-ML backward sweep of a group (solvers.py:231-271): the modulus update of the far
// field (noise_models.py:47-63, error of mode 0 when calc_error), then for i = S-1 .. 0
// with chi_i in registers: store chi_i (for the object update), the FP64 alpha_O sums
// of slice i, dP = chi_i conj(O_i), and either chi_{i-1} = ifft2(fft2(dP) conj(p_{i-1}))
// or, at slice 0 with a probe update, the slot's probe spectrum fft2(dP) conj(filt_b)
// into the farfield buffer with the alpha_P sums.  Object and probe updates follow in
// their own launches; the sweep reads only the pre-update cutouts (solvers.py:246-260).
// slices whose alpha partials the fused backward sweep keeps per wave in LDS
constexpr int BK_WRED_SLICES = 32;

template <int N, int T, bool UPD_OBJ, bool PROBE>
__global__ void __launch_bounds__(T) lsqml_backward_fused_k(phx_group_desc d, int calc_error) {
  int b, m;
  sweep_plane(d, b, m);
  if (!slot_active(d, b)) return;
  PHX_PLANE_SETUP(N, T);
  constexpr int NN = N * N;
  constexpr int E = Plane<N>::E;
  constexpr float inv_nn = 1.0f / (float)NN;
  const int M = d.n_modes, S = d.n_slices, bmax = d.bmax;
  const int p = d.group_pos[b];
  const Cut c = cut_of(d, p);
  const size_t HW = (size_t)d.obj_h * d.obj_w;
  const RedLayout L = red_layout(S, bmax, M);
  const PropSeq props = prop_seq(d, b);
  // alpha sums: each wave leaves its FP64 partials of a slice step in its own LDS slots
  // (slice i at slot i, the probe sums at slot S) and the slots are summed once, after
  // the sweep, so the slice steps take no block-wide barrier for them.  Deeper stacks
  // than the slots hold sum every step across the block instead.
  constexpr int NW = T / 64;
  __shared__ double wred_lds_[(BK_WRED_SLICES + 1) * 2 * NW];
  const bool deferred = S <= BK_WRED_SLICES;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  auto stash = [&](int slot, float v0, float v1, auto flush) __attribute__((always_inline)) {
    if (deferred) {
      const double a = wave_sum_f64(v0), bb = wave_sum_f64(v1);
      if (lane == 0) {
        wred_lds_[(2 * slot) * NW + wave] = a;
        wred_lds_[(2 * slot + 1) * NW + wave] = bb;
      }
    } else {
      const double a = block_sum_f64(v0, red_lds_);
      const double bb = block_sum_f64(v1, red_lds_);
      if (threadIdx.x == 0) flush(a, bb);
    }
  };
  auto flush_obj = [&](int i) {
    return [&, i](double a, double bb) {
      atomic_add_f64(&d.red[L.num_obj + i * bmax + b], a);
      atomic_add_f64(&d.red[L.den_obj + i], bb);
    };
  };
  auto flush_probe = [&](double a, double bb) {
    d.red[L.numP + b * M + m] = a;
    atomic_add_f64(&d.red[L.denP], bb);
  };

  float2 x[E];
  // far field: chi_hat = (sqrt(D + c)/(sqrt(I + c) + eps) - 1) mask F
  const float2* F = (const float2*)d.farfield + ((size_t)b * M + m) * NN;
  const float* D = d.patterns_s + (size_t)p * NN;
  const float* mask = d.mask_s;
  const float off = d.params[P_OFFSET], eps = d.params[P_EPS];
  float e2 = 0.f;
  // isum(e): the model intensity of point e (mode terms added in mode order)
  auto farfield_phase = [&](auto isum) __attribute__((always_inline)) {
    pl.cols_apply(x, [&](int py, int px, float2) {
      const int e = py * N + px;
      const float Iv = isum(e), Dv = D[e], mk = mask[e];
      const float2 Fv = F[e];
      const float Dp = (Dv > 0.f || isnan(Dv)) ? Dv : 0.f;  // numpy.maximum keeps NaN
      const float df = Iv - Dv;
      const float t = df * df;  // unmasked (solvers.py:237), NaN-skipping (nansum)
      e2 += isnan(t) ? 0.f : t;
      return cscale(Fv, (sqrtf(Dp + off) / (sqrtf(Iv + off) + eps) - 1.0f) * mk);
    }, [&] { __asm__ volatile("" : "+v"(e2)); });
  };
  if (M <= 4) {
    const ModeSum4 ms(d, b);
    farfield_phase([&](int e) { return ms(e); });
  } else {
    float Ik[E];
    model_intensity_cols<N>(d, b, pl.q, pl.rc, Ik);
    int kk = 0;  // register index of the point cols_apply passes (its loop is unrolled)
    farfield_phase([&](int) { return Ik[kk++]; });
  }
  if (calc_error && m == 0) {
    const double tot = block_sum_f64(e2, red_lds_);
    if (threadIdx.x == 0) d.errors[b] = (float)sqrt(tot);
  }
  pl.inverse_regs(x);

  // one slice step on chi_i (row phase, unscaled): store it, the alpha_O sums, and
  // x <- dP = chi_i conj(O_i); at slice 0 also the alpha_P sums
  auto slice_step = [&](int i, auto probe_tag, auto interior_tag) __attribute__((always_inline)) {
    constexpr bool PS = decltype(probe_tag)::value;
    constexpr bool IN = decltype(interior_tag)::value;
    float2* chi_i = chi_in_buf(d, i) + ((size_t)b * M + m) * NN;
    const float2* psi_i = (const float2*)d.psi + psi_off(d, i, b, m);
    const float2* obj_i = (const float2*)d.object + i * HW;
    const float2* orow = obj_i + (size_t)(c.y0 + pl.rc) * c.W + c.x0;  // interior cutouts
    float numO = 0.f, denO = 0.f, numP = 0.f, denP = 0.f;
    pl.rows_apply(x, [&](int yy, int xx, float2 v) {
      const float2 cv = cscale(v, inv_nn);
      chi_i[yy * N + xx] = cv;
      const float a = abs2f(cv);
      if (UPD_OBJ) {
        // t = (chi conj(psi)) psi = chi |psi|^2: Re(conj(chi) t) = |chi|^2 |psi|^2,
        // |t|^2 = |chi|^2 |psi|^4
        const float p2 = abs2f(psi_i[yy * N + xx]);
        const float ap = a * p2;
        numO += ap;
        denO = fmaf(ap, p2, denO);
      }
      const float2 o = IN ? orow[xx] : obj_at(obj_i, c, yy, xx);
      const float2 dP = cmulc(cv, o);  // chi conj(O)
      if (PS) {
        // t2 = dP O = chi |O|^2, as above
        const float o2 = abs2f(o);
        const float ao = a * o2;
        numP += ao;
        denP = fmaf(ao, o2, denP);
      }
      return dP;
    }, [&] {
      if (UPD_OBJ) __asm__ volatile("" : "+v"(numO), "+v"(denO));
      if (PS) __asm__ volatile("" : "+v"(numP), "+v"(denP));
    });
    if (UPD_OBJ) stash(i, numO, denO, flush_obj(i));
    if (PS) stash(S, numP, denP, flush_probe);
  };
  // (cutouts read through obj_at: an interior-only copy of the sweep would need more
  // registers than the plane allows alongside the general one)
  for (int i = S - 1; i > 0; --i) {
    pl.refresh();
    slice_step(i, std::false_type{}, std::false_type{});
    pl.template convolve_prefetched<true>(x, props[i - 1], d.prop_zero_waves);
  }
  pl.refresh();
  slice_step(0, std::integral_constant<bool, PROBE>{}, std::false_type{});
  if (PROBE) {
    // per-plane fft2(dP) conj(filt_b) into this plane's (already consumed) farfield slot
    const float sy = d.shift[2 * p], sx = d.shift[2 * p + 1];
    float2* Fp = (float2*)d.farfield + ((size_t)b * M + m) * NN;
    pl.forward_regs(x);
    const int px = pl.rc;
    const float kx = d.kx_s[px];
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const int py = pl.QL * k + pl.q;
      Fp[py * N + px] = cmulc(x[k], shift_filter(sy, sx, d.ky_s[py], kx));
      if ((k & 3) == 3) __builtin_amdgcn_sched_barrier(0);  // a few phase factors at a time
    }
  }
  if (deferred && (UPD_OBJ || PROBE)) {
    __syncthreads();
    for (int t = threadIdx.x; t < 2 * (S + 1); t += T) {
      const int slot = t >> 1;
      if (slot < S ? !UPD_OBJ : !PROBE) continue;
      double sum = 0.0;
      for (int w = 0; w < NW; ++w) sum += wred_lds_[t * NW + w];
      if (slot == S) {
        if (t & 1) atomic_add_f64(&d.red[L.denP], sum);
        else d.red[L.numP + b * M + m] = sum;
      } else {
        atomic_add_f64(t & 1 ? &d.red[L.den_obj + slot] : &d.red[L.num_obj + slot * bmax + b], sum);
      }
    }
  }
}

// Fused adjoint sweep of the gradient engine (the reverse of group_forward_k; cf. the
// per-slice gradient_backward_k): starting from the exit-wave gradient in chi_a (written
// by the far-field gradient update), for i = S-1 .. 0 with the gradient carried in
// registers: store g as chi_i (grad_O_i = sum_{b,m} scatter(g conj(psi_i)) is gathered
// afterwards by gradient_object_gather_k, without atomics), g <- g conj(O_i), and
// g <- ifft2(fft2(g) conj(p_{i-1})); at slice 0 the probe-gradient spectrum and the
// position gradients.  Needs every psi_i stored and, with S > 1, chi_slices.
// FF = 0 starts from chi_a as phx_farfield_gradient_update left it; FF = FM_GRAD_AMPLITUDE
// or FM_GRAD_POISSON computes that far-field gradient step itself (the same values, the
// group loss included) in the first column pass and transforms it in registers, and
// stores slice S-1's wave gradient into chi_a.
template <int N, int T, bool WANT_OBJ, bool WANT_P0, int FF>
__device__ __forceinline__ void gradient_sweep_plane(const phx_group_desc& d, int b, int m, int want_probe,
                                                     int want_pos, float2* lds_, double* red_lds_) {
  PHX_PLANE_BIND(N);
  constexpr int NN = N * N;
  constexpr int E = Plane<N>::E;
  constexpr int G = E < 8 ? E : 8;  // points per load group
  constexpr float inv_nn = 1.0f / (float)NN;
  const int M = d.n_modes, S = d.n_slices;
  const int p = d.group_pos[b];
  const Cut c = cut_of(d, p);
  const bool inside = c.interior(N);
  const size_t HW = (size_t)d.obj_h * d.obj_w;
  const PropSeq props = prop_seq(d, b);
  float2 x[E];
  if constexpr (FF == 0) {
    // the stored exit-wave gradient, times n^2 (a power of two: exact) to carry it unscaled
    const float2* g0 = (const float2*)d.chi_a + ((size_t)b * M + m) * NN;
    pl.rows_apply(x, [&](int yy, int xx, float2) { return cscale(g0[yy * N + xx], (float)NN); }, [] {});
  } else {
    const float2* F = (const float2*)d.farfield + ((size_t)b * M + m) * NN;
    const float* D = d.patterns_s + (size_t)p * NN;
    const float* mask = d.mask_s;
    const float off = d.params[P_OFFSET], eps = d.params[P_EPS], var = d.params[P_VAR];
    const float gscale = d.params[P_LOSS_SCALE];
    float e2 = 0.f;
    // isum(e): the model intensity of point e (mode terms added in mode order)
    auto farfield_phase = [&](auto isum) __attribute__((always_inline)) {
      pl.cols_apply(x, [&](int py, int px, float2) {
        const int e = py * N + px;
        const float Iv = isum(e), Dv = D[e], mk = mask[e];
        const float2 Fv = F[e];
        float t;
        const float u = farfield_factor<FF>(Iv, Dv, mk, off, eps, var, gscale, t);
        e2 += t;
        return cscale(Fv, u);
      }, [&] { __asm__ volatile("" : "+v"(e2)); });
    };
    if (M <= 4) {
      const ModeSum4 ms(d, b);
      farfield_phase([&](int e) { return ms(e); });
    } else {
      float Ik[E];
      model_intensity_cols<N>(d, b, pl.q, pl.rc, Ik);
      int kk = 0;  // register index of the point cols_apply passes (its loop is unrolled)
      farfield_phase([&](int) { return Ik[kk++]; });
    }
    if (m == 0) add_group_loss<FF>(d, e2, red_lds_);
    pl.inverse_regs(x);
  }
  for (int i = S - 1; i >= 0; --i) {
    pl.refresh();
    const float2* obj_i = (const float2*)d.object + i * HW;
    const int yy = pl.rc;
    // the wave gradient at slice i goes to chi_i (chi_a itself at i = S-1), for
    // gradient_object_gather_k to form grad_O_i without atomics
    float2* gi = chi_in_buf(d, i) + ((size_t)b * M + m) * NN;
    const bool store = WANT_OBJ && (FF != 0 || i < S - 1);
    auto step = [&](auto interior_tag) __attribute__((always_inline)) {
      constexpr bool IN = decltype(interior_tag)::value;
      const float2* orow = obj_i + (size_t)(c.y0 + yy) * c.W + c.x0;  // interior cutouts
#pragma unroll
      for (int j0 = 0; j0 < E; j0 += G) {
        float2 ob[G];
#pragma unroll
        for (int u = 0; u < G; ++u) {
          const int xx = pl.QL * (j0 + u) + pl.q;
          ob[u] = IN ? orow[xx] : obj_at(obj_i, c, yy, xx);
        }
#pragma unroll
        for (int u = 0; u < G; ++u) {
          const int xx = pl.QL * (j0 + u) + pl.q;
          const float2 gv = cscale(x[j0 + u], inv_nn);
          if (store) gi[yy * N + xx] = gv;
          x[j0 + u] = cmulc(gv, ob[u]);
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    };
    if (inside) step(std::true_type{});
    else step(std::false_type{});
    if (i == 0) break;
    const float2* prop = props[i - 1];
    pl.convolve_regs(x, [&](int py, int px, float2 v) { return cmulc(v, prop[py * N + px]); }, d.prop_zero_waves);
  }
  if (!WANT_P0) return;
  // slice 0: probe-gradient spectrum and position gradients (as gradient_backward_k)
  const float sy = d.shift[2 * p], sx = d.shift[2 * p + 1];
  const float2* Pk = (const float2*)d.probe_k + (size_t)m * NN;
  // this plane's probe-gradient spectrum goes to its own farfield slot (read only by this
  // workgroup's far-field pass above); phx_gradient_probe_finish sums the slots over the
  // group in position order, so the probe gradient does not depend on scheduling
  float2* Fb = (float2*)d.farfield + ((size_t)b * M + m) * NN;
  float gy = 0.f, gx = 0.f;  // per-lane FP32 partials, FP64 block sums
  pl.forward_regs(x);
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const int py = pl.QL * k + pl.q, px = pl.rc;
    const int e = py * N + px;
    const float2 f = shift_filter(sy, sx, d.ky_s[py], d.kx_s[px]);
    if (want_probe) Fb[e] = cmulc(x[k], f);
    if (want_pos) {
      const float2 t = cmulc(cmul(Pk[e], f), x[k]);  // conj(v) * fft2(P) filt
      gy += TWO_PI * d.ky_s[py] * t.y;
      gx += TWO_PI * d.kx_s[px] * t.y;
    }
    if ((k & 3) == 3) __builtin_amdgcn_sched_barrier(0);
  }
  if (want_pos) {
    const double ty = block_sum_f64(gy, red_lds_);
    const double tx = block_sum_f64(gx, red_lds_);
    if (threadIdx.x == 0) {
      const RedLayout L = red_layout(S, d.bmax, M);
      atomic_add_f64(&d.red[L.pos_num + 2 * b], ty * (double)inv_nn);
      atomic_add_f64(&d.red[L.pos_num + 2 * b + 1], tx * (double)inv_nn);
    }
  }
}

template <int N, int T, bool WANT_OBJ, bool WANT_P0, int FF>
__global__ void __launch_bounds__(T) gradient_sweep_k(phx_group_desc d, int want_probe, int want_pos) {
  int b, m;
  sweep_plane(d, b, m);
  if (!slot_active(d, b)) return;
  PHX_PLANE_LDS(N, T);
  gradient_sweep_plane<N, T, WANT_OBJ, WANT_P0, FF>(d, b, m, want_probe, want_pos, lds_, red_lds_);
}

// A group's forward and adjoint sweeps as plane tasks in launches of at most one
// workgroup per CU (phx_gradient_group_sweeps): task t < n_fwd is the forward sweep of
// plane fwd_lo + t, the others the adjoint sweep (with the far-field step FF) of plane
// adj_lo + t - n_fwd, where plane q is (b, m) = (q / M, q mod M).  The host places an
// adjoint plane only in a launch after the one holding the forward sweeps of all its
// slot's modes.  Workgroups go to the XCDs round-robin by linear id, so with a multiple
// of 8 of them XCD x takes the contiguous tasks [x n/8, (x+1) n/8): the modes of a slot
// share one L2.
template <int N, int T, bool WANT_OBJ, bool WANT_P0, int FF>
__global__ void __launch_bounds__(T) gradient_tasks_k(phx_group_desc d, int want_probe, int want_pos, int fwd_lo,
                                                      int n_fwd, int adj_lo) {
  const int n = gridDim.x, L = blockIdx.x;
  const int t = (n & 7) == 0 ? (L & 7) * (n >> 3) + (L >> 3) : L;
  const bool fwd = t < n_fwd;
  const int q = fwd ? fwd_lo + t : adj_lo + (t - n_fwd);
  const int b = q / d.n_modes, m = q - b * d.n_modes;
  if (!slot_active(d, b)) return;
  PHX_PLANE_LDS(N, T);
  if (fwd)
    group_forward_plane<N, T>(d, b, m, lds_, red_lds_);
  else
    gradient_sweep_plane<N, T, WANT_OBJ, WANT_P0, FF>(d, b, m, want_probe, want_pos, lds_, red_lds_);
}

// Object tiles of the gathers: one row of UPD_TW pixels (a multiple of 64, so each wave
// owns 64 contiguous pixels of one row).  A wave reads 512 contiguous bytes of a plane
// row per listed position and the block's waves read adjacent spans, so only the block's
// two outer 128-byte lines are partial (a 64 x 4 tile had one partial line per wave
// and row: 1.21x the algorithmic bytes).
#ifndef PHX_UPD_TW
#define PHX_UPD_TW 256
#endif
constexpr int UPD_TW = PHX_UPD_TW, UPD_TH = 256 / PHX_UPD_TW;
constexpr int UPD_LIST = 256;  // positions listed per pass

// The gather shared by the LSQ-ML object update and the gradient engine's object
// gradient: for the pixel (y, x) of this thread in the UPD_TW x UPD_TH tile at (ty0, tx0) of
// slice i, dO = sum_{b,m} chi_i conj(psi_i) and (with WANT_G) g = sum_{b,m} |psi_i|^2
// over the group's cutouts that cover it, positions in group order, modes in order.
// PER_POS_W (ePIE) weights position b's term by 1 / aux[bmax + b] (its max |psi|^2).
// Every thread of the block must call it (it synchronises the block).
template <bool WANT_G, bool PER_POS_W = false>
__device__ __forceinline__ void gather_tile(const phx_group_desc& d, int i, int ty0, int tx0, int y, int x,
                                            float2& dO, float& g) {
  __shared__ int s_b[UPD_LIST], s_y0[UPD_LIST], s_x0[UPD_LIST];
  __shared__ float s_w[PER_POS_W ? UPD_LIST : 1];
  __shared__ int s_cnt;
  const int gs = *d.group_size;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int n = d.n, NN = n * n, M = d.n_modes;
  const float2* chi = chi_in_buf(d, i);
  const float2* psi = (const float2*)d.psi;
  dO = make_float2(0.f, 0.f);
  g = 0.f;
  for (int c0 = 0; c0 < gs; c0 += UPD_LIST) {
    const int c1 = min(gs, c0 + UPD_LIST);
    __syncthreads();  // the previous list is consumed
    if (wave == 0) {
      int cnt = 0;
      for (int b0 = c0; b0 < c1; b0 += 64) {
        const int b = b0 + lane;
        bool ov = false;
        int y0 = 0, x0 = 0;
        if (b < c1) {
          const int p = d.group_pos[b];
          y0 = d.start[2 * p];
          x0 = d.start[2 * p + 1];
          ov = !(y0 >= ty0 + UPD_TH || y0 + n <= ty0 || x0 >= tx0 + UPD_TW || x0 + n <= tx0);
        }
        const uint64_t msk = __ballot(ov);
        if (ov) {
          const int r = cnt + __popcll(msk & ((1ull << lane) - 1ull));
          s_b[r] = b;
          s_y0[r] = y0;
          s_x0[r] = x0;
          if constexpr (PER_POS_W) s_w[r] = 1.0f / d.aux[d.bmax + b];
        }
        cnt += __popcll(msk);
      }
      if (lane == 0) s_cnt = cnt;
    }
    __syncthreads();
    const int cnt = s_cnt;
    // this wave's pixels: row y, columns [wx0, wx0 + 64)
    const int wx0 = x - lane;
    const bool wave_out = y >= d.obj_h || wx0 >= d.obj_w;
#pragma unroll 2
    for (int t = 0; t < cnt; ++t) {
      const int b = s_b[t];
      const int ly = y - s_y0[t], lx = x - s_x0[t];
      // a listed cutout that misses this wave's row or its 64 columns (or a wave wholly
      // outside the object) adds nothing: the wave skips its loads (wave-uniform test)
      const int lxw = wx0 - s_x0[t];
      if (wave_out || (unsigned)ly >= (unsigned)n || lxw >= n || lxw + 64 <= 0) continue;
      // branch-free: clamped (valid) address, contribution weighted by the in-cutout test
      const float w = ((unsigned)ly < (unsigned)n && (unsigned)lx < (unsigned)n) ? 1.f : 0.f;
      const size_t off = (size_t)min(max(ly, 0), n - 1) * n + min(max(lx, 0), n - 1);
      const float2* cb = chi + (size_t)b * M * NN + off;
      const float2* pb = psi + psi_off(d, i, b, 0) + off;
      float2 dOb = make_float2(0.f, 0.f);
      float gb = 0.f;
      int m = 0;
      for (; m + 4 <= M; m += 4) {
        const float2 c0v = cb[(m + 0) * NN], c1v = cb[(m + 1) * NN];
        const float2 c2v = cb[(m + 2) * NN], c3v = cb[(m + 3) * NN];
        const float2 p0 = pb[(m + 0) * NN], p1 = pb[(m + 1) * NN];
        const float2 p2 = pb[(m + 2) * NN], p3 = pb[(m + 3) * NN];
        dOb = cadd(dOb, cmulc(c0v, p0));
        dOb = cadd(dOb, cmulc(c1v, p1));
        dOb = cadd(dOb, cmulc(c2v, p2));
        dOb = cadd(dOb, cmulc(c3v, p3));
        if (WANT_G) {
          gb += abs2f(p0);
          gb += abs2f(p1);
          gb += abs2f(p2);
          gb += abs2f(p3);
        }
      }
      for (; m < M; ++m) {
        const float2 cv = cb[m * NN], ps = pb[m * NN];
        dOb = cadd(dOb, cmulc(cv, ps));
        if (WANT_G) gb += abs2f(ps);
      }
      if constexpr (PER_POS_W) dO = cadd(dO, cscale(dOb, w * s_w[t]));
      else dO = cadd(dO, cscale(dOb, w));
      if (WANT_G) g += gb * w;
    }
  }
}

// Accumulators a group step starts from zero (gom, red, errors); cleared by
// zero4_k, or by the previous group's object update in the same launch.
struct ClearSpan {
  float* a;
  int64_t na;
  float* b;
  int64_t nb;
  double* c;
  int64_t nc;
  float* e;
  int64_t ne;
};

// Grid-stride zeroing of a ClearSpan by every thread of the launch (any grid shape).
__device__ __forceinline__ void clear_span(const ClearSpan& z) {
  const int64_t nthr = (int64_t)gridDim.x * gridDim.y * gridDim.z * blockDim.x;
  const int64_t blk = ((int64_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
  const int64_t tot = z.na + z.nb + z.nc + z.ne;
  for (int64_t k = blk * blockDim.x + threadIdx.x; k < tot; k += nthr) {
    if (k < z.na) z.a[k] = 0.f;
    else if (k < z.na + z.nb) z.b[k - z.na] = 0.f;
    else if (k < z.na + z.nb + z.nc) z.c[k - z.na - z.nb] = 0.0;
    else z.e[k - z.na - z.nb - z.nc] = 0.f;
  }
}

// LSQ-ML object update of slice i (solvers.py:213-215, 248-257):
//   O_i += beta_O (sum_b alpha_O,b) dO/(gpm + reg) * gpm/(gpm + 1e-16),
//   dO  = sum_{b,m} scatter(chi_i conj(psi_i)),   gpm = sum_{b,m} scatter(|psi_i|^2).
// One workgroup owns a UPD_TW x UPD_TH object tile of slice first + blockIdx.z (one pixel
// per thread; see UPD_TW for the line alignment).  Wave 0
// lists the group's positions whose cutout meets the tile, in group order (ballot +
// prefix count); then every thread gathers both sums from those positions' chi_i and
// psi_i planes (gather_tile): no atomics, and the result does not depend on scheduling.
// Pixels no cutout covers are left untouched (their reference update is exactly zero).
// `z` (possibly empty) is the next group's accumulator clear, issued first: the update
// reads none of those buffers (its own red slot is not the next group's).
__global__ void __launch_bounds__(256) lsqml_object_update_k(phx_group_desc d, int first, ClearSpan z) {
  clear_span(z);
  const int i = first + (int)blockIdx.z;
  __shared__ float sA;
  const RedLayout L = red_layout(d.n_slices, d.bmax, d.n_modes);
  const int gs = *d.group_size;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (wave == 0) {
    double s = 0.0;
    for (int b = lane; b < gs; b += 64) s += d.red[L.num_obj + i * d.bmax + b];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) {
      const double den = d.red[L.den_obj + i] + (double)d.params[P_GAMMA];
      sA = (float)((double)d.params[P_BETA_OBJ] * (s / den));
    }
  }
  const float reg = d.params[P_REG_OBJ];
  const int H = d.obj_h, W = d.obj_w;
  const int ty0 = blockIdx.y * UPD_TH, tx0 = blockIdx.x * UPD_TW;
  const int y = ty0 + threadIdx.x / UPD_TW, x = tx0 + threadIdx.x % UPD_TW;
  float2 dO;
  float g;
  gather_tile<true>(d, i, ty0, tx0, y, x, dO, g);  // its barriers also publish sA
  if (y >= H || x >= W || (g == 0.f && dO.x == 0.f && dO.y == 0.f)) return;
  const float A = sA;
  float2* obj_i = (float2*)d.object + (size_t)i * H * W;
  const float2 q = cscale(dO, 1.0f / (g + reg));
  const float2 u = cscale(cscale(cscale(q, A), g), 1.0f / (g + 1e-16f));
  const float2 o = cadd(obj_i[(size_t)y * W + x], u);
  obj_i[(size_t)y * W + x] = o;
  if (!(isfinite(o.x) && isfinite(o.y))) atomicMin(d.finite_flag, d.group_index);
}

// Object gradient of slice first + blockIdx.z for the gradient engine:
// grad_O_i = sum_{b,m} scatter(g_i conj(psi_i)) from the wave gradients g_i that
// gradient_sweep_k stored as chi_i; tiles and order as lsqml_object_update_k.  Writes
// every pixel of its tiles (zero where no cutout covers it).
// `z` (possibly empty) is the next group's accumulator clear, as in lsqml_object_update_k.
__global__ void __launch_bounds__(256) gradient_object_gather_k(phx_group_desc d, int first, ClearSpan z) {
  clear_span(z);
  const int i = first + (int)blockIdx.z;
  const int H = d.obj_h, W = d.obj_w;
  const int ty0 = blockIdx.y * UPD_TH, tx0 = blockIdx.x * UPD_TW;
  const int y = ty0 + threadIdx.x / UPD_TW, x = tx0 + threadIdx.x % UPD_TW;
  float2 dO;
  float g;
  gather_tile<false>(d, i, ty0, tx0, y, x, dO, g);
  if (y >= H || x >= W) return;
  ((float2*)d.grad_obj)[(size_t)i * H * W + (size_t)y * W + x] = dO;
}

// d_probe[m] = sum_b farfield[b][m]: the per-position probe-update spectra written by
// the slice-0 backward step, summed over the group in position order.  Loads are issued
// eight positions ahead of the (ordered) additions.
__global__ void __launch_bounds__(256) probe_spectrum_reduce_k(phx_group_desc d) {
  const int NN = d.n * d.n, M = d.n_modes, m = blockIdx.y;
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= NN) return;
  const int gs = *d.group_size;
  const size_t stride = (size_t)M * NN;
  const float2* F = (const float2*)d.farfield + (size_t)m * NN + e;
  float2 s = make_float2(0.f, 0.f);
  int b = 0;
  for (; b + 8 <= gs; b += 8) {
    float2 v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = F[(size_t)(b + u) * stride];
#pragma unroll
    for (int u = 0; u < 8; ++u) s = cadd(s, v[u]);
  }
  for (; b < gs; ++b) s = cadd(s, F[(size_t)b * stride]);
  ((float2*)d.d_probe)[(size_t)m * NN + e] = s;
}

// P += beta_P (sum_b alpha_P) ifft2(sum_b fft2(dP) conj(filt))/(gom+reg) * gom/(gom+eps),
// then probe_k = fft2(P) of the updated probe (what probe_spectra_k computes), so the
// next group starts from current spectra.  The update's operands (gom, P) are all
// loaded before the first store, not point by point.
template <int N, int T>
__global__ void __launch_bounds__(T) lsqml_probe_apply_k(phx_group_desc d) {
  PHX_PLANE_SETUP(N, T);
  constexpr int NN = N * N;
  constexpr int E = Plane<N>::E;
  constexpr float inv_nn = 1.0f / (float)NN;
  __shared__ float sA;
  const int m = blockIdx.x, M = d.n_modes;
  const RedLayout L = red_layout(d.n_slices, d.bmax, M);
  if (threadIdx.x < 64) {
    const int gs = *d.group_size;
    double s = 0.0;
    for (int b = threadIdx.x; b < gs; b += 64) s += d.red[L.numP + b * M + m];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (threadIdx.x == 0) {
      const double den = d.red[L.denP] + (double)d.params[P_GAMMA];
      sA = (float)((double)d.params[P_BETA_PROBE] * (s / den));
    }
  }
  float2* dPk = (float2*)d.d_probe + (size_t)m * NN;
  float2* P = (float2*)d.probe + (size_t)m * NN;
  float2 x[E];
#pragma unroll
  for (int k = 0; k < E; ++k) x[k] = dPk[(pl.QL * k + pl.q) * N + pl.rc];
#pragma unroll
  for (int k = 0; k < E; ++k) dPk[(pl.QL * k + pl.q) * N + pl.rc] = make_float2(0.f, 0.f);
  __syncthreads();  // sA
  const float A = sA, reg = d.params[P_REG_PROBE];
  pl.inverse_regs(x);
  float2 pv[E];
  float g[E];
#pragma unroll
  for (int j = 0; j < E; ++j) {
    const int e = pl.rc * N + pl.QL * j + pl.q;
    g[j] = d.gom[e];
    pv[j] = P[e];
  }
  __builtin_amdgcn_sched_barrier(0);  // every load above issued before the first store
  bool bad = false;
#pragma unroll
  for (int j = 0; j < E; ++j) {
    const float2 avg = cscale(cscale(x[j], inv_nn), 1.0f / (g[j] + reg));
    const float2 u = cscale(cscale(cscale(avg, A), g[j]), 1.0f / (g[j] + 1e-16f));
    const float2 o = cadd(pv[j], u);
    P[pl.rc * N + pl.QL * j + pl.q] = o;
    bad |= !(isfinite(o.x) && isfinite(o.y));
    x[j] = o;
  }
  if (bad) atomicMin(d.finite_flag, d.group_index);
  pl.refresh();  // no address arithmetic carried over from the update above
  pl.forward_regs(x);
  float2* Pk = (float2*)d.probe_k + (size_t)m * NN;
#pragma unroll
  for (int k = 0; k < E; ++k) Pk[(pl.QL * k + pl.q) * N + pl.rc] = x[k];
}

// The same probe step split over three launches of strips (R lines of a plane per
// workgroup, QL lanes per line: 128 threads at n = 128), each with a footprint like one
// object-update workgroup, so that they run beside the object update as its workgroups
// retire instead of waiting for whole free CUs.  A 1-D transform along a lane group is the
// plane's own fwd/inv (the in-wave exchange uses this wave's slots of a strip-sized LDS
// buffer), and the layouts in tmp are those the plane's transposes produce, so every
// value is computed as in lsqml_probe_apply_k:
//   cols_inv:  tmp[y][px]   = inverse column transform of d_probe (scrambled py -> y)
//   rows:      P += update(inverse row transform of tmp) ; tmp[y][px] = forward row
//              transform of the new P (x -> scrambled px)
//   cols_fwd:  probe_k      = forward column transform of tmp (y -> scrambled py)
template <int N>
struct Strip {
  static constexpr int QL = Plane<N>::QL;
  static constexpr int R = (128 / QL) < N ? (128 / QL) : N;  // lines per workgroup
  static constexpr int T = R * QL;
  static constexpr int LDS_FLOAT2 = QL == 8 ? R * Plane<N>::LD : 0;  // exchange slots
};

#define PHX_STRIP_SETUP(N)                                                                   \
  using SG = Strip<N>;                                                                       \
  __shared__ float2 slds_[SG::LDS_FLOAT2 + N];                                               \
  Plane<N> pl(slds_, slds_ + SG::LDS_FLOAT2);                                                \
  for (int i = threadIdx.x; i < N; i += SG::T)                                               \
    slds_[SG::LDS_FLOAT2 + i] = ((const float2*)d.twiddle)[(i % SG::QL) * (i / SG::QL)];     \
  __syncthreads();                                                                           \
  const int line = blockIdx.x * SG::R + pl.rc;                                               \
  const int m = blockIdx.y;                                                                  \
  constexpr int NN = N * N;                                                                  \
  constexpr int E = Plane<N>::E;                                                             \
  constexpr int QL = SG::QL;

template <int N>
__global__ void __launch_bounds__(Strip<N>::T) probe_apply_cols_inv_k(phx_group_desc d, float2* tmp) {
  PHX_STRIP_SETUP(N);
  float2* dPk = (float2*)d.d_probe + (size_t)m * NN;
  float2* t = tmp + (size_t)m * NN;
  float2 x[E];
#pragma unroll
  for (int k = 0; k < E; ++k) x[k] = dPk[(QL * k + pl.q) * N + line];
#pragma unroll
  for (int k = 0; k < E; ++k) dPk[(QL * k + pl.q) * N + line] = make_float2(0.f, 0.f);
  pl.template inv<false>(x);
#pragma unroll
  for (int j = 0; j < E; ++j) t[(QL * j + pl.q) * N + line] = x[j];
}

template <int N>
__global__ void __launch_bounds__(Strip<N>::T) probe_apply_rows_k(phx_group_desc d, float2* tmp) {
  PHX_STRIP_SETUP(N);
  constexpr float inv_nn = 1.0f / (float)NN;
  __shared__ float sA;
  const int M = d.n_modes;
  if (threadIdx.x < 64) {
    const RedLayout L = red_layout(d.n_slices, d.bmax, M);
    const int gs = *d.group_size;
    double s = 0.0;
    for (int b = threadIdx.x; b < gs; b += 64) s += d.red[L.numP + b * M + m];
    s = wave_sum_f64(s);
    if (threadIdx.x == 0) {
      const double den = d.red[L.denP] + (double)d.params[P_GAMMA];
      sA = (float)((double)d.params[P_BETA_PROBE] * (s / den));
    }
  }
  float2* t = tmp + (size_t)m * NN;
  float2* P = (float2*)d.probe + (size_t)m * NN;
  float2 x[E];
#pragma unroll
  for (int k = 0; k < E; ++k) x[k] = t[line * N + QL * k + pl.q];
  __syncthreads();  // sA
  const float A = sA, reg = d.params[P_REG_PROBE];
  pl.template inv<false>(x);
  float2 pv[E];
  float g[E];
#pragma unroll
  for (int j = 0; j < E; ++j) {
    const int e = line * N + QL * j + pl.q;
    g[j] = d.gom[e];
    pv[j] = P[e];
  }
  __builtin_amdgcn_sched_barrier(0);  // every load above issued before the first store
  bool bad = false;
#pragma unroll
  for (int j = 0; j < E; ++j) {
    const float2 avg = cscale(cscale(x[j], inv_nn), 1.0f / (g[j] + reg));
    const float2 u = cscale(cscale(cscale(avg, A), g[j]), 1.0f / (g[j] + 1e-16f));
    const float2 o = cadd(pv[j], u);
    P[line * N + QL * j + pl.q] = o;
    bad |= !(isfinite(o.x) && isfinite(o.y));
    x[j] = o;
  }
  if (bad) atomicMin(d.finite_flag, d.group_index);
  pl.template fwd<false>(x);
#pragma unroll
  for (int k = 0; k < E; ++k) t[line * N + QL * k + pl.q] = x[k];
}

template <int N>
__global__ void __launch_bounds__(Strip<N>::T) probe_apply_cols_fwd_k(phx_group_desc d, float2* tmp) {
  PHX_STRIP_SETUP(N);
  const float2* t = tmp + (size_t)m * NN;
  float2* Pk = (float2*)d.probe_k + (size_t)m * NN;
  float2 x[E];
#pragma unroll
  for (int j = 0; j < E; ++j) x[j] = t[(QL * j + pl.q) * N + line];
  pl.template fwd<false>(x);
#pragma unroll
  for (int k = 0; k < E; ++k) Pk[(QL * k + pl.q) * N + line] = x[k];
}

// LimitProbeSupport (regularizers.py:45-64): P <- ifft2(fft2(P) mask) per mode, mask the
// host-computed k-disk |k| <= max_angle / lambda in scrambled order (0 or 1); then
// probe_k = fft2 of the new probe, as the probe step leaves it, so the next group's
// forward model starts from current spectra.
template <int N, int T>
__global__ void __launch_bounds__(T) probe_support_k(phx_group_desc d, const float* mask_s) {
  PHX_PLANE_SETUP(N, T);
  constexpr int NN = N * N;
  constexpr int E = Plane<N>::E;
  constexpr float inv_nn = 1.0f / (float)NN;
  const int m = blockIdx.x;
  float2* P = (float2*)d.probe + (size_t)m * NN;
  float2 x[E];
  pl.load_rows([&](int y, int xx) { return P[y * N + xx]; }, x);
  pl.convolve_regs(x, [&](int py, int px, float2 v) { return cscale(v, mask_s[py * N + px]); });
#pragma unroll
  for (int j = 0; j < E; ++j) {
    const float2 v = cscale(x[j], inv_nn);
    P[pl.rc * N + pl.QL * j + pl.q] = v;
    x[j] = v;
  }
  pl.refresh();
  pl.forward_regs(x);
  float2* Pk = (float2*)d.probe_k + (size_t)m * NN;
#pragma unroll
  for (int k = 0; k < E; ++k) Pk[(pl.QL * k + pl.q) * N + pl.rc] = x[k];
}

// ---------------------------------- ePIE ------------------------------------------
// aux[b] = max_px |O_i|^2, aux[bmax + b] = max_px sum_m |psi_i|^2 (solvers.py:462-472)
__global__ void __launch_bounds__(256) epie_maxima_k(phx_group_desc d, int i) {
  __shared__ float scratch[4];
  const int b = blockIdx.x;
  if (!slot_active(d, b)) return;
  const int n = d.n, NN = n * n, M = d.n_modes;
  const Cut c = cut_of(d, d.group_pos[b]);
  const size_t HW = (size_t)d.obj_h * d.obj_w;
  const float2* obj_i = (const float2*)d.object + i * HW;
  const float2* psi_i = (const float2*)d.psi + psi_off(d, i, b, 0);
  float mo = 0.f, mp = 0.f;
  for (int e = threadIdx.x; e < NN; e += 256) {
    const int y = e / n, x = e % n;
    mo = fmaxf(mo, abs2f(obj_at(obj_i, c, y, x)));
    float s = 0.f;
    for (int mm = 0; mm < M; ++mm) s += abs2f(psi_i[(size_t)mm * NN + e]);
    mp = fmaxf(mp, s);
  }
  mo = block_max_f32(mo, scratch);
  mp = block_max_f32(mp, scratch);
  if (threadIdx.x == 0) {
    d.aux[b] = mo;
    d.aux[d.bmax + b] = mp;
  }
}

template <int N, int T>
__global__ void __launch_bounds__(T) epie_backward_k(phx_group_desc d, int i, int upd_obj, int upd_probe) {
  const int m = blockIdx.x, b = blockIdx.y;
  if (!slot_active(d, b)) return;
  PHX_PLANE_SETUP(N, T);
  constexpr int NN = N * N;
  constexpr float inv_nn = 1.0f / (float)NN;
  const int M = d.n_modes, bmax = d.bmax;
  const int p = d.group_pos[b];
  const Cut c = cut_of(d, p);
  const size_t HW = (size_t)d.obj_h * d.obj_w;
  const float2* chi = chi_in_buf(d, i) + ((size_t)b * M + m) * NN;
  float2* chi_o = chi_out_buf(d, i) + ((size_t)b * M + m) * NN;
  const float2* obj_i = (const float2*)d.object + i * HW;
  const float maxO = d.aux[b];
  // (the object update of slice i is gathered afterwards from chi_i and psi_i by
  // epie_object_update_k; upd_obj needs nothing from this kernel)
  (void)upd_obj;

  auto src = [&](int y, int x) -> float2 {
    const float2 cv = chi[y * N + x];
    const bool inb = c.inb(y, x);
    const float2 o = inb ? obj_i[c.idx(y, x)] : make_float2(1.f, 0.f);
    return cscale(cmulc(cv, o), 1.0f / maxO);
  };
  if (i > 0) {
    const float2* prop = prop_of(d, i - 1, b);
    pl.convolve(src, [&](int py, int px, float2 v) { return cmulc(v, prop[py * N + px]); },
                [&](int y, int x, float2 v) { chi_o[y * N + x] = cscale(v, inv_nn); });
  } else if (upd_probe) {
    const float sy = d.shift[2 * p], sx = d.shift[2 * p + 1];
    // per-position spectrum into this plane's farfield slot (the far-field step is done
    // with it); phx_epie_probe_apply sums the slots in position order
    float2* Fb = (float2*)d.farfield + ((size_t)b * M + m) * NN;
    pl.forward(src, [&](int py, int px, float2& v) {
      Fb[py * N + px] = cmulc(v, shift_filter(sy, sx, d.ky_s[py], d.kx_s[px]));
    });
  }
}

// ePIE object update of slice i (solvers.py:466-477):
//   O_i += beta_O / S * sum_b (sum_m chi_i conj(psi_i)) / max_pixels(sum_m |psi_i|^2)_b
// over the group's cutouts, gathered per object tile from the chi_i and psi_i planes in
// group order (gather_tile with the per-position weights epie_maxima_k left in aux):
// no atomics, so the object does not depend on scheduling.  Runs after epie_backward_k
// of slice i and before that of slice i-1 (which overwrites chi_i's buffer).  Pixels no
// cutout covers are left untouched.  Grid (ceil(W/64), ceil(H/4)).
__global__ void __launch_bounds__(256) epie_object_update_k(phx_group_desc d, int i) {
  const int H = d.obj_h, W = d.obj_w;
  const int ty0 = blockIdx.y * UPD_TH, tx0 = blockIdx.x * UPD_TW;
  const int y = ty0 + threadIdx.x / UPD_TW, x = tx0 + threadIdx.x % UPD_TW;
  float2 dO;
  float g;
  gather_tile<false, true>(d, i, ty0, tx0, y, x, dO, g);
  if (y >= H || x >= W || (dO.x == 0.f && dO.y == 0.f)) return;
  const float betaS = d.params[P_BETA_OBJ] / (float)d.n_slices;
  float2* obj_i = (float2*)d.object + (size_t)i * H * W;
  const float2 o = cadd(obj_i[(size_t)y * W + x], cscale(dO, betaS));
  obj_i[(size_t)y * W + x] = o;
  if (!(isfinite(o.x) && isfinite(o.y))) atomicMin(d.finite_flag, d.group_index);
}

template <int N, int T>
__global__ void __launch_bounds__(T) epie_probe_apply_k(phx_group_desc d) {
  PHX_PLANE_SETUP(N, T);
  constexpr int NN = N * N;
  constexpr float inv_nn = 1.0f / (float)NN;
  const int m = blockIdx.x;
  const float scale = d.params[P_BETA_PROBE] * inv_nn / (float)(*d.group_size);
  float2* dPk = (float2*)d.d_probe + (size_t)m * NN;
  float2* P = (float2*)d.probe + (size_t)m * NN;
  bool bad = false;
  pl.inverse(
      [&](int py, int px) {
        const int e = py * N + px;
        const float2 v = dPk[e];
        dPk[e] = make_float2(0.f, 0.f);
        return v;
      },
      [&](int y, int x, float2 v) {
        const int e = y * N + x;
        const float2 o = cadd(P[e], cscale(v, scale));
        P[e] = o;
        bad |= !(isfinite(o.x) && isfinite(o.y));
      });
  if (bad) atomicMin(d.finite_flag, d.group_index);
}

// ----------------------------- dry run / position ---------------------------------
__global__ void __launch_bounds__(256) dry_run_stats_k(phx_group_desc d) {
  __shared__ double scratch[4];
  const int b = blockIdx.x;
  if (!slot_active(d, b)) return;
  const int NN = d.n * d.n, M = d.n_modes, s = d.n_slices - 1;
  const int p = d.group_pos[b];
  const float2* psi_s = (const float2*)d.psi + psi_off(d, s, b, 0);
  const float* D = d.patterns_s + (size_t)p * NN;
  double model = 0.0, ex = 0.0;
  for (int e = threadIdx.x; e < NN; e += 256) {
    for (int mm = 0; mm < M; ++mm) model += (double)abs2f(psi_s[(size_t)mm * NN + e]);
    ex += (double)(D[e] * d.mask_s[e]);
  }
  model = block_sum_f64(model, scratch);
  ex = block_sum_f64(ex, scratch);
  if (threadIdx.x == 0) d.errors[b] = (float)ex / (float)model;
}

template <int N, int T>
__global__ void __launch_bounds__(T) position_gradient_k(phx_group_desc d, int dir) {
  const int m = blockIdx.x, b = blockIdx.y;
  if (!slot_active(d, b)) return;
  PHX_PLANE_SETUP(N, T);
  constexpr int NN = N * N;
  const int M = d.n_modes, bmax = d.bmax;
  const RedLayout L = red_layout(d.n_slices, bmax, M);
  const Cut c = cut_of(d, d.group_pos[b]);
  const float2* psi0 = (const float2*)d.psi + psi_off(d, 0, b, m);
  const float2* chi = chi_in_buf(d, 0) + ((size_t)b * M + m) * NN;  // the final chi (solvers.py:275-288)
  const float2* obj0 = (const float2*)d.object;
  const float scale = (float)(1.0 / ((double)NN * d.red[L.norm + b]));
  float num = 0.f, den = 0.f;  // per-lane FP32 partials, FP64 block sums
  pl.convolve([&](int y, int x) { return psi0[y * N + x]; },
              [&](int py, int px, float2 v) {
                const float k = dir == 0 ? d.ky_s[py] : d.kx_s[px];
                const float w = TWO_PI * k;
                return make_float2(w * v.y, -w * v.x);  // v * (-2 pi i k)
              },
              [&](int y, int x, float2 v) {
                const float2 prod = cmul(cscale(v, scale), obj_at(obj0, c, y, x));
                const float2 cv = chi[y * N + x];
                num += cv.x * prod.x + cv.y * prod.y;
                den += abs2f(prod);
              });
  const double a = block_sum_f64(num, red_lds_);
  const double bb = block_sum_f64(den, red_lds_);
  if (threadIdx.x == 0) {
    atomic_add_f64(&d.red[L.pos_num + 2 * b + dir], a);
    atomic_add_f64(&d.red[L.pos_den + dir], bb);
  }
}

// -------------------------- gradient engine: adjoint ---------------------------------
// Adjoint of the multislice model at slice i (phaser/engines/gradient/run.py:352-405,
// differentiated by hand).  chi_in holds g_a = 2 dL/dconj(a_i) for a_i = psi_i * O_i
// (torch's convention).  Writes g_O += g_a conj(psi_i) into grad_obj[i]; for i > 0 the
// back-propagated g_{a_{i-1}} = ifft2(fft2(g_a conj(O_i)) conj(prop_{i-1})); for i == 0
// the probe-gradient spectrum fft2(g_psi0) conj(filt_b) into the plane's farfield slot
// (summed over b in order by phx_gradient_probe_finish) and the
// position gradients dL/dy_b, dL/dx_b = Re sum conj(fft2 g_psi0) fft2(P) filt (-2 pi i k).
template <int N, int T>
__global__ void __launch_bounds__(T) gradient_backward_k(phx_group_desc d, int i, int want_obj, int want_probe,
                                                         int want_pos) {
  const int m = blockIdx.x, b = blockIdx.y;
  if (!slot_active(d, b)) return;
  PHX_PLANE_SETUP(N, T);
  constexpr int NN = N * N;
  constexpr float inv_nn = 1.0f / (float)NN;
  const int M = d.n_modes;
  const int p = d.group_pos[b];
  const Cut c = cut_of(d, p);
  const size_t HW = (size_t)d.obj_h * d.obj_w;
  const float2* g = chi_in_buf(d, i) + ((size_t)b * M + m) * NN;
  float2* g_o = chi_out_buf(d, i) + ((size_t)b * M + m) * NN;
  const float2* psi_i = (const float2*)d.psi + psi_off(d, i, b, m);
  const float2* obj_i = (const float2*)d.object + i * HW;
  float2* gO = (float2*)d.grad_obj + i * HW;

  auto src = [&](int y, int x) -> float2 {
    const float2 gv = g[y * N + x];
    const bool inb = c.inb(y, x);
    const float2 o = inb ? obj_i[c.idx(y, x)] : make_float2(1.f, 0.f);
    if (want_obj && inb) atomic_add_c64(&gO[c.idx(y, x)], cmulc(gv, psi_i[y * N + x]));
    return cmulc(gv, o);
  };
  if (i > 0) {
    const float2* prop = prop_of(d, i - 1, b);
    pl.convolve(src, [&](int py, int px, float2 v) { return cmulc(v, prop[py * N + px]); },
                [&](int y, int x, float2 v) { g_o[y * N + x] = cscale(v, inv_nn); });
  } else if (want_probe || want_pos) {
    const float sy = d.shift[2 * p], sx = d.shift[2 * p + 1];
    const float2* Pk = (const float2*)d.probe_k + (size_t)m * NN;
    // per-position spectrum into this plane's farfield slot, summed in position order
    // by phx_gradient_probe_finish (no atomics)
    float2* Fb = (float2*)d.farfield + ((size_t)b * M + m) * NN;
    float gy = 0.f, gx = 0.f;  // per-lane FP32 partials, FP64 block sums
    pl.forward(src, [&](int py, int px, float2& v) {
      const int e = py * N + px;
      const float2 f = shift_filter(sy, sx, d.ky_s[py], d.kx_s[px]);
      if (want_probe) Fb[e] = cmulc(v, f);
      if (want_pos) {
        const float2 t = cmulc(cmul(Pk[e], f), v);  // conj(v) * fft2(P) filt
        gy += TWO_PI * d.ky_s[py] * t.y;
        gx += TWO_PI * d.kx_s[px] * t.y;
      }
    });
    if (want_pos) {
      const double ty = block_sum_f64(gy, red_lds_);
      const double tx = block_sum_f64(gx, red_lds_);
      if (threadIdx.x == 0) {
        const RedLayout L = red_layout(d.n_slices, d.bmax, M);
        atomic_add_f64(&d.red[L.pos_num + 2 * b], ty * (double)inv_nn);
        atomic_add_f64(&d.red[L.pos_num + 2 * b + 1], tx * (double)inv_nn);
      }
    }
  } else {
    for (int e = threadIdx.x; e < NN; e += T) (void)src(e / N, e % N);
  }
}

// grad_probe[m] = ifft2(d_probe[m]) (natural order, ortho scale); d_probe = 0.  d_probe
// holds the ordered sum of the per-position spectra (probe_spectrum_reduce_k).
template <int N, int T>
__global__ void __launch_bounds__(T) gradient_probe_finish_k(phx_group_desc d) {
  PHX_PLANE_SETUP(N, T);
  constexpr int NN = N * N;
  constexpr float inv_nn = 1.0f / (float)NN;
  const int m = blockIdx.x;
  float2* dPk = (float2*)d.d_probe + (size_t)m * NN;
  float2* gP = (float2*)d.grad_probe + (size_t)m * NN;
  pl.inverse(
      [&](int py, int px) {
        const int e = py * N + px;
        const float2 v = dPk[e];
        dPk[e] = make_float2(0.f, 0.f);
        return v;
      },
      [&](int y, int x, float2 v) { gP[y * N + x] = cscale(v, inv_nn); });
}

// Gradient-engine dry run (gradient/run.py:408-438): rescale[b] = sum(D) / sum_m
// |psi_{S-1} O_{S-1}|^2 (last slice included, no mask; Parseval for the fft2).
__global__ void __launch_bounds__(256) dry_run_stats_gradient_k(phx_group_desc d) {
  __shared__ double scratch[4];
  const int b = blockIdx.x;
  if (!slot_active(d, b)) return;
  const int n = d.n, NN = n * n, M = d.n_modes, s = d.n_slices - 1;
  const int p = d.group_pos[b];
  const Cut c = cut_of(d, p);
  const float2* psi_s = (const float2*)d.psi + psi_off(d, s, b, 0);
  const float2* obj_s = (const float2*)d.object + (size_t)s * d.obj_h * d.obj_w;
  const float* D = d.patterns_s + (size_t)p * NN;
  double model = 0.0, ex = 0.0;
  for (int e = threadIdx.x; e < NN; e += 256) {
    const float2 o = obj_at(obj_s, c, e / n, e % n);
    for (int mm = 0; mm < M; ++mm) model += (double)abs2f(cmul(psi_s[(size_t)mm * NN + e], o));
    ex += (double)D[e];
  }
  model = block_sum_f64(model, scratch);
  ex = block_sum_f64(ex, scratch);
  if (threadIdx.x == 0) d.errors[b] = (float)ex / (float)model;
}

// The same ratio after group_forward_k: the slot's model intensity sum_m |F_m|^2 / n^2
// summed over the plane equals sum_m |psi_{S-1} O_{S-1}|^2 (Parseval).
__global__ void __launch_bounds__(256) dry_run_stats_intensity_k(phx_group_desc d) {
  __shared__ double scratch[4];
  const int b = blockIdx.x;
  if (!slot_active(d, b)) return;
  const int NN = d.n * d.n;
  const float* D = d.patterns_s + (size_t)d.group_pos[b] * NN;
  double model = 0.0, ex = 0.0;
  for (int e = threadIdx.x; e < NN; e += 256) {
    model += (double)model_intensity(d, b, e);
    ex += (double)D[e];
  }
  model = block_sum_f64(model, scratch);
  ex = block_sum_f64(ex, scratch);
  if (threadIdx.x == 0) d.errors[b] = (float)ex / (float)model;
}

// out[pos[b]][natural k] = intensity[b][scrambled k]  (model intensity of each slot)
template <int N>
__global__ void exit_intensity_copy_k(phx_group_desc d, float* out) {
  const int b = blockIdx.y;
  if (!slot_active(d, b)) return;
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= N * N) return;
  const int py = e / N, px = e % N;
  const int fe = scr_to_freq<N>(py) * N + scr_to_freq<N>(px);
  out[(size_t)d.group_pos[b] * N * N + fe] = model_intensity(d, b, e);
}

// pos_update[p] = num[b] / den (solvers.py:280), scattered to the group's positions.
__global__ void position_finish_k(phx_group_desc d, float* pos_update) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= *d.group_size) return;
  const RedLayout L = red_layout(d.n_slices, d.bmax, d.n_modes);
  const int p = d.group_pos[b];
  for (int dir = 0; dir < 2; ++dir)
    pos_update[2 * p + dir] = (float)(d.red[L.pos_num + 2 * b + dir] / d.red[L.pos_den + dir]);
}

__global__ void zero4_k(ClearSpan z) { clear_span(z); }

#include "phx_split256.h"

// --------------------------------- host side ---------------------------------------
// The accumulators phx_group_begin clears for group descriptor d.
ClearSpan begin_span(const phx_group_desc& d) {
  ClearSpan z;
  z.a = d.gom;
  z.na = (int64_t)d.n * d.n;
  // intensity needs no clear: the forward kernels store every mode term of an active slot
  z.b = nullptr;
  z.nb = 0;
  z.c = d.red;
  z.nc = red_layout(d.n_slices, d.bmax, d.n_modes).total;
  z.e = d.errors;
  z.ne = d.bmax;
  return z;
}

bool supported_n(int n) { return n == 16 || n == 32 || n == 64 || n == 128 || n == 192 || n == 256; }

int check_desc(const phx_group_desc* d) {
  if (!d || !supported_n(d->n) || d->n_modes < 1 || d->n_slices < 1 || d->bmax < 1 || d->obj_h < 1 ||
      d->obj_w < 1 || d->n_pos < 1)
    return PHX_EINVAL;
  if (!d->group_pos || !d->group_size || !d->start || !d->shift || !d->twiddle || !d->object || !d->probe ||
      !d->params || !d->psi || !d->red)
    return PHX_EINVAL;
  if (d->n_slices > 1 && !d->props) return PHX_EINVAL;
  if (d->tilt && (!d->slice_dz || !d->tprops)) return PHX_EINVAL;
  return PHX_OK;
}

#define PHX_PLANE_GRID(d) dim3((unsigned)(d)->n_modes, (unsigned)(d)->bmax)

#define PHX_DISPATCH(n, LAUNCH)       \
  switch (n) {                        \
    case 16: LAUNCH(16, 64); break;   \
    case 32: LAUNCH(32, 128); break;  \
    case 64: LAUNCH(64, 256); break;  \
    case 128: LAUNCH(128, 1024); break;\
    case 192: LAUNCH(192, 768); break; \
    default: return PHX_EUNSUPPORTED; \
  }

}  // namespace

template <int N, int T>
static void launch_lsqml_backward(const phx_group_desc& d, int slice, int upd_obj, int upd_probe,
                                  hipStream_t stream) {
  const dim3 grid = PHX_PLANE_GRID(&d), block(T);
#define K(KIND)                                                                                        \
  do {                                                                                                 \
    if (upd_obj)                                                                                       \
      hipLaunchKernelGGL((lsqml_backward_k<N, T, KIND, true>), grid, block, 0, stream, d, slice);      \
    else                                                                                               \
      hipLaunchKernelGGL((lsqml_backward_k<N, T, KIND, false>), grid, block, 0, stream, d, slice);     \
  } while (0)
  if (slice > 0)
    K(BK_PROPAGATE);
  else if (upd_probe)
    K(BK_PROBE);
  else
    K(BK_NONE);
#undef K
}

// ---- work plane of the no-allocation probe step (phx_lsqml_probe_apply_strips_static) ----
// A file-scope device array: every GPU the library runs on holds one copy at its own
// address until the process ends.  Host code never passes the array's name to a kernel
// (in host code that name is a stand-in, not a device address); it uses the address
// hipGetSymbolAddress returns for the current device, looked up once per device.
constexpr int64_t PROBE_STRIP_SCRATCH_C64 = 131072;  // 1 MiB: n_modes * n * n <= 2^17
static __device__ __align__(256) float2 probe_strip_scratch[PROBE_STRIP_SCRATCH_C64];

namespace {
constexpr int SCRATCH_MAX_DEVICES = 64;
std::once_flag g_scratch_once[SCRATCH_MAX_DEVICES];
void* g_scratch_ptr[SCRATCH_MAX_DEVICES];
hipError_t g_scratch_err[SCRATCH_MAX_DEVICES];

// Device address of probe_strip_scratch on the current device (looked up on the first call
// made on that device, with that device current).  0 or a negative hipError_t.
int probe_strip_scratch_ptr(float2** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return -(int)e;
  if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) return PHX_EINVAL;
  std::call_once(g_scratch_once[dev], [dev] {
    void* p = nullptr;
    g_scratch_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(probe_strip_scratch));
    g_scratch_ptr[dev] = p;
  });
  if (g_scratch_err[dev] != hipSuccess) return -(int)g_scratch_err[dev];
  if (!g_scratch_ptr[dev]) return PHX_EINVAL;
  *out = (float2*)g_scratch_ptr[dev];
  return PHX_OK;
}

std::once_flag g_cu_once[SCRATCH_MAX_DEVICES];
int g_cu_count[SCRATCH_MAX_DEVICES];

// Compute units of the current device (queried once per device), or 0 if unknown.
int device_cu_count() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= SCRATCH_MAX_DEVICES) return 0;
  std::call_once(g_cu_once[dev], [dev] {
    int c = 0;
    if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) c = 0;
    g_cu_count[dev] = c;
  });
  return g_cu_count[dev];
}
}  // namespace

extern "C" {

int64_t phx_red_doubles(int32_t n_slices, int32_t bmax, int32_t n_modes) {
  return red_layout(n_slices, bmax, n_modes).total;
}

int phx_group_begin(const phx_group_desc* d, hipStream_t stream) {
  int r = check_desc(d);
  if (r) return r;
  if (!d->gom || !d->intensity || !d->errors) return PHX_EINVAL;
  const ClearSpan z = begin_span(*d);
  int64_t blocks = (z.na + z.nb + z.nc + z.ne + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(zero4_k, dim3((unsigned)blocks), dim3(256), 0, stream, z);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_tilt_propagators(const phx_group_desc* d, hipStream_t stream) {
  int r = check_desc(d);
  if (r) return r;
  if (!d->tilt || !d->slice_dz || !d->tprops || !d->props || d->n_slices < 2) return PHX_EINVAL;
  const int NN = d->n * d->n;
  hipLaunchKernelGGL(tilt_props_k, dim3((unsigned)((NN + 255) / 256), (unsigned)(d->n_slices - 1), (unsigned)d->bmax),
                     dim3(256), 0, stream, *d);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_probe_spectra(const phx_group_desc* d, hipStream_t stream) {
  int r = check_desc(d);
  if (r) return r;
  if (!d->probe_k) return PHX_EINVAL;
  if (d->n == S256) return s256_probe_spectra(*d, stream);
#define L(N, T) hipLaunchKernelGGL((probe_spectra_k<N, T>), dim3((unsigned)d->n_modes), dim3(T), 0, stream, *d)
  PHX_DISPATCH(d->n, L);
#undef L
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_group_obj_mag(const phx_group_desc* d, hipStream_t stream) {
  int r = check_desc(d);
  if (r) return r;
  if (!d->gom) return PHX_EINVAL;
  const int NN = d->n * d->n;
  hipLaunchKernelGGL(group_obj_mag_k, dim3((unsigned)((NN + 63) / 64)), dim3(64 * GOM_WAVES), 0, stream, *d);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_shift_probes(const phx_group_desc* d, int illum, int propagate, hipStream_t stream) {
  int r = check_desc(d);
  if (r) return r;
  if (!d->probe_k || (illum && !d->gpm)) return PHX_EINVAL;
  if (propagate && d->n_slices < 2) return PHX_EINVAL;
  if (d->n == S256) return illum ? PHX_EUNSUPPORTED : s256_shift_probes(*d, propagate, stream);
#define L(N, T)                                                                                          \
  do {                                                                                                   \
    if (illum)                                                                                           \
      hipLaunchKernelGGL((shift_probes_k<N, T, true>), PHX_PLANE_GRID(d), dim3(T), 0, stream, *d, propagate);  \
    else                                                                                                 \
      hipLaunchKernelGGL((shift_probes_k<N, T, false>), PHX_PLANE_GRID(d), dim3(T), 0, stream, *d, propagate); \
  } while (0)
  PHX_DISPATCH(d->n, L);
#undef L
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_forward_slice(const phx_group_desc* d, int slice, int illum, hipStream_t stream) {
  int r = check_desc(d);
  if (r) return r;
  if (slice < 0 || slice > d->n_slices - 2 || (illum && !d->gpm)) return PHX_EINVAL;
  if (d->n == S256) return illum ? PHX_EUNSUPPORTED : s256_forward_slice(*d, slice, stream);
#define L(N, T)                                                                                          \
  do {                                                                                                   \
    if (illum)                                                                                           \
      hipLaunchKernelGGL((forward_slice_k<N, T, true>), PHX_PLANE_GRID(d), dim3(T), 0, stream, *d, slice);     \
    else                                                                                                 \
      hipLaunchKernelGGL((forward_slice_k<N, T, false>), PHX_PLANE_GRID(d), dim3(T), 0, stream, *d, slice);    \
  } while (0)
  PHX_DISPATCH(d->n, L);
#undef L
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_farfield(const phx_group_desc* d, int illum, int calc_error, hipStream_t stream) {
  int r = check_desc(d);
  if (r) return r;
  if (!d->farfield || !d->intensity || !d->patterns_s || !d->mask_s || !d->chi_a || (illum && !d->gpm) ||
      (calc_error && !d->errors))
    return PHX_EINVAL;
  if (d->n == S256) return illum ? PHX_EUNSUPPORTED : s256_farfield(*d, calc_error, stream);
#define L(N, T)                                                                                          \
  do {                                                                                                   \
    if (illum)                                                                                           \
      hipLaunchKernelGGL((farfield_fwd_k<N, T, true>), PHX_PLANE_GRID(d), dim3(T), 0, stream, *d);             \
    else                                                                                                 \
      hipLaunchKernelGGL((farfield_fwd_k<N, T, false>), PHX_PLANE_GRID(d), dim3(T), 0, stream, *d);            \
  } while (0)
  PHX_DISPATCH(d->n, L);
#undef L
  PHX_CHECK_LAUNCH();
#define L(N, T) \
  hipLaunchKernelGGL((farfield_update_k<N, T, FM_MODULUS>), PHX_PLANE_GRID(d), dim3(T), 0, stream, *d, calc_error)
  PHX_DISPATCH(d->n, L);
#undef L
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_lsqml_backward(const phx_group_desc* d, int slice, int update_object, int update_probe,
                       hipStream_t stream) {
  int r = check_desc(d);
  if (r) return r;
  if (slice < 0 || slice >= d->n_slices || !d->chi_a || !d->chi_b || (update_probe && !d->d_probe))
    return PHX_EINVAL;
  if (d->n == S256) return s256_lsqml_backward(*d, slice, update_object, update_probe, stream);
#define L(N, T) launch_lsqml_backward<N, T>(*d, slice, update_object, update_probe, stream)
  PHX_DISPATCH(d->n, L);
#undef L
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_group_forward(const phx_group_desc* d, hipStream_t stream) {
  int r = check_desc(d);
  if (r) return r;
  if (!d->probe_k || !d->farfield || !d->intensity || (d->n_slices > 1 && !d->props)) return PHX_EINVAL;
  if (d->psi_slots > 0 && d->psi_slots < d->n_slices) return PHX_EINVAL;  // stores every slice
  if (d->n == S256) return s256_group_forward(*d, stream);
#define L(N, T) hipLaunchKernelGGL((group_forward_k<N, T>), PHX_PLANE_GRID(d), dim3(T), 0, stream, *d)
  PHX_DISPATCH(d->n, L);
#undef L
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_farfield_update(const phx_group_desc* d, int calc_error, hipStream_t stream) {
  int r = check_desc(d);
  if (r) return r;
  if (!d->farfield || !d->intensity || !d->patterns_s || !d->mask_s || !d->chi_a || (calc_error && !d->errors))
    return PHX_EINVAL;
  if (d->n == S256) return s256_farfield_update(*d, calc_error, stream);
#define L(N, T) \
  hipLaunchKernelGGL((farfield_update_k<N, T, FM_MODULUS>), PHX_PLANE_GRID(d), dim3(T), 0, stream, *d, calc_error)
  PHX_DISPATCH(d->n, L);
#undef L
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_lsqml_backward_sweep(const phx_group_desc* d, int update_object, int update_probe, int calc_error,
                             hipStream_t stream) {
  int r = check_desc(d);
  if (r) return r;
  if (!d->farfield || !d->intensity || !d->patterns_s || !d->mask_s || !d->chi_a || (calc_error && !d->errors) ||
      (d->n_slices > 1 && (!d->props || !d->chi_slices)))
    return PHX_EINVAL;
  if (d->psi_slots > 0 && d->psi_slots < d->n_slices) return PHX_EINVAL;
  if (d->n == S256) return s256_lsqml_backward_sweep(*d, update_object, update_probe, calc_error, stream);
#define K(N, T, UO, PR) \
  hipLaunchKernelGGL((lsqml_backward_fused_k<N, T, UO, PR>), PHX_PLANE_GRID(d), dim3(T), 0, stream, *d, calc_error)
#define L(N, T)                       \
  do {                                \
    if (update_object && update_probe) \
      K(N, T, true, true);            \
    else if (update_object)           \
      K(N, T, true, false);           \
    else if (update_probe)            \
      K(N, T, false, true);           \
    else                              \
      K(N, T, false, false);          \
  } while (0)
  PHX_DISPATCH(d->n, L);
#undef L
#undef K
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_farfield_gradient_update(const phx_group_desc* d, int poisson, hipStream_t stream) {
  int r = check_desc(d);
  if (r) return r;
  if (!d->farfield || !d->intensity || !d->patterns_s || !d->mask_s || !d->chi_a) return PHX_EINVAL;
  if (d->n == S256) return s256_farfield_gradient_update(*d, poisson, stream);
#define L(N, T)                                                                                                  \
  do {                                                                                                           \
    if (poisson)                                                                                                 \
      hipLaunchKernelGGL((farfield_update_k<N, T, FM_GRAD_POISSON>), PHX_PLANE_GRID(d), dim3(T), 0, stream, *d, 0);   \
    else                                                                                                         \
      hipLaunchKernelGGL((farfield_update_k<N, T, FM_GRAD_AMPLITUDE>), PHX_PLANE_GRID(d), dim3(T), 0, stream, *d, 0); \
  } while (0)
  PHX_DISPATCH(d->n, L);
#undef L
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_gradient_sweep(const phx_group_desc* d, int want_object, int want_probe, int want_positions,
                       hipStream_t stream) {
  int r = check_desc(d);
  if (r) return r;
  if (!d->chi_a || (d->n_slices > 1 && !d->props) || (want_object && d->n_slices > 1 && !d->chi_slices) ||
      ((want_probe || want_positions) && (!d->d_probe || !d->probe_k)) || (want_probe && !d->farfield))
    return PHX_EINVAL;
  if (d->psi_slots > 0 && d->psi_slots < d->n_slices) return PHX_EINVAL;  // needs every psi_i
  const bool p0 = want_probe || want_positions;
#define K(N, T, WO, WP) \
  hipLaunchKernelGGL((gradient_sweep_k<N, T, WO, WP, 0>), PHX_PLANE_GRID(d), dim3(T), 0, stream, *d, want_probe, \
                     want_positions)
#define L(N, T)                 \
  do {                          \
    if (want_object && p0)      \
      K(N, T, true, true);      \
    else if (want_object)       \
      K(N, T, true, false);     \
    else if (p0)                \
      K(N, T, false, true);     \
    else                        \
      K(N, T, false, false);    \
  } while (0)
  PHX_DISPATCH(d->n, L);
#undef L
#undef K
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_gradient_sweep_farfield(const phx_group_desc* d, int poisson, int want_object, int want_probe,
                                int want_positions, hipStream_t stream) {
  int r = check_desc(d);
  if (r) return r;
  if (!d->chi_a || !d->farfield || !d->intensity || !d->patterns_s || !d->mask_s ||
      (d->n_slices > 1 && !d->props) || (want_object && d->n_slices > 1 && !d->chi_slices) ||
      ((want_probe || want_positions) && (!d->d_probe || !d->probe_k)))
    return PHX_EINVAL;
  if (d->psi_slots > 0 && d->psi_slots < d->n_slices) return PHX_EINVAL;  // needs every psi_i
  const bool p0 = want_probe || want_positions;
#define K(N, T, WO, WP, FF) \
  hipLaunchKernelGGL((gradient_sweep_k<N, T, WO, WP, FF>), PHX_PLANE_GRID(d), dim3(T), 0, stream, *d, want_probe, \
                     want_positions)
#define KF(N, T, WO, WP)                 \
  do {                                   \
    if (poisson)                         \
      K(N, T, WO, WP, FM_GRAD_POISSON);  \
    else                                 \
      K(N, T, WO, WP, FM_GRAD_AMPLITUDE);\
  } while (0)
#define L(N, T)                 \
  do {                          \
    if (want_object && p0)      \
      KF(N, T, true, true);     \
    else if (want_object)       \
      KF(N, T, true, false);    \
    else if (p0)                \
      KF(N, T, false, true);    \
    else                        \
      KF(N, T, false, false);   \
  } while (0)
  PHX_DISPATCH(d->n, L);
#undef L
#undef KF
#undef K
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_plane_task_plan(int planes, int modes, int cus, int* plan, int max_launches) {
  const int P = planes, M = modes, C = cus;
  if (P < 1 || M < 1 || P % M != 0 || !plan || max_launches < 1) return PHX_EINVAL;
  if (C < 1 || P <= C || P > 2 * C) return 0;  // one round per sweep, or no round to save
  int launches = 0, fwd = 0, adj = 0;
  for (; adj < P && launches < max_launches; ++launches) {
    const int nf = std::min(C, P - fwd);
    const int na = std::max(0, std::min(C - nf, (fwd / M) * M - adj));
    int* t = plan + 4 * launches;
    t[0] = fwd;
    t[1] = nf;
    t[2] = adj;
    t[3] = na;
    fwd += nf;
    adj += na;
  }
  if (adj < P || launches >= 2 * ((P + C - 1) / C)) return 0;
  return launches;
}

int phx_gradient_group_sweeps(const phx_group_desc* d, int poisson, int want_object, int want_probe,
                              int want_positions, hipStream_t stream) {
  int r = check_desc(d);
  if (r) return r;
  if (!d->probe_k || !d->farfield || !d->intensity || !d->chi_a || !d->patterns_s || !d->mask_s ||
      (d->n_slices > 1 && !d->props) || (want_object && d->n_slices > 1 && !d->chi_slices) ||
      ((want_probe || want_positions) && (!d->d_probe || !d->probe_k)))
    return PHX_EINVAL;
  if (d->psi_slots > 0 && d->psi_slots < d->n_slices) return PHX_EINVAL;  // needs every psi_i
  // Plane tasks: P forward sweeps, then P adjoint sweeps, in launches of at most C (one
  // workgroup per CU at 128^2, so each launch is one round).  Launch k takes the
  // remaining forward planes first, then adjoint planes whose slots' forward sweeps all
  // ran in earlier launches.  For P = 384 planes on 256 CUs that is 3 full launches
  // where the two sweeps take 4 rounds (2 of them half full).
  constexpr int MAX_LAUNCHES = 8;
  int plan[4 * MAX_LAUNCHES];
  const int launches = d->n == 128 ? phx_plane_task_plan(d->bmax * d->n_modes, d->n_modes, device_cu_count(), plan,
                                                         MAX_LAUNCHES)
                                   : 0;
  if (launches <= 0) {
    r = phx_group_forward(d, stream);
    if (r) return r;
    return phx_gradient_sweep_farfield(d, poisson, want_object, want_probe, want_positions, stream);
  }
  const bool p0 = want_probe || want_positions;
  for (int k = 0; k < launches; ++k) {
    const int* t = plan + 4 * k;  // forward planes [t0, t0 + t1), adjoint planes [t2, t2 + t3)
#define K(WO, WP, FF)                                                                                         \
  hipLaunchKernelGGL((gradient_tasks_k<128, 1024, WO, WP, FF>), dim3((unsigned)(t[1] + t[3])), dim3(1024), 0, \
                     stream, *d, want_probe, want_positions, t[0], t[1], t[2])
#define KF(WO, WP)                     \
  do {                                 \
    if (poisson)                       \
      K(WO, WP, FM_GRAD_POISSON);      \
    else                               \
      K(WO, WP, FM_GRAD_AMPLITUDE);    \
  } while (0)
    if (want_object && p0)
      KF(true, true);
    else if (want_object)
      KF(true, false);
    else if (p0)
      KF(false, true);
    else
      KF(false, false);
#undef KF
#undef K
    PHX_CHECK_LAUNCH();
  }
  return PHX_OK;
}

int phx_lsqml_object_update_slices(const phx_group_desc* d, int first, int count, hipStream_t stream) {
  int r = check_desc(d);
  if (r) return r;
  if (first < 0 || count < 1 || first + count > d->n_slices || !d->chi_a || !d->chi_b || !d->finite_flag ||
      (count > 1 && !d->chi_slices))
    return PHX_EINVAL;
  const dim3 grid((unsigned)((d->obj_w + UPD_TW - 1) / UPD_TW), (unsigned)((d->obj_h + UPD_TH - 1) / UPD_TH),
                  (unsigned)count);
  hipLaunchKernelGGL(lsqml_object_update_k, grid, dim3(256), 0, stream, *d, first, ClearSpan{});
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_lsqml_object_update_slices_begin(const phx_group_desc* d, int first, int count, const phx_group_desc* next,
                                         hipStream_t stream) {
  int r = check_desc(d);
  if (r) return r;
  r = check_desc(next);
  if (r) return r;
  if (first < 0 || count < 1 || first + count > d->n_slices || !d->chi_a || !d->chi_b || !d->finite_flag ||
      (count > 1 && !d->chi_slices))
    return PHX_EINVAL;
  if (!next->gom || !next->intensity || !next->errors || next->red == d->red) return PHX_EINVAL;
  const dim3 grid((unsigned)((d->obj_w + UPD_TW - 1) / UPD_TW), (unsigned)((d->obj_h + UPD_TH - 1) / UPD_TH),
                  (unsigned)count);
  hipLaunchKernelGGL(lsqml_object_update_k, grid, dim3(256), 0, stream, *d, first, begin_span(*next));
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_gradient_object_gather(const phx_group_desc* d, int first, int count, hipStream_t stream) {
  int r = check_desc(d);
  if (r) return r;
  if (first < 0 || count < 1 || first + count > d->n_slices || !d->chi_a || !d->grad_obj ||
      (d->n_slices > 1 && !d->chi_slices) || (d->psi_slots > 0 && d->psi_slots < d->n_slices))
    return PHX_EINVAL;
  const dim3 grid((unsigned)((d->obj_w + UPD_TW - 1) / UPD_TW), (unsigned)((d->obj_h + UPD_TH - 1) / UPD_TH),
                  (unsigned)count);
  hipLaunchKernelGGL(gradient_object_gather_k, grid, dim3(256), 0, stream, *d, first, ClearSpan{});
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_gradient_object_gather_begin(const phx_group_desc* d, int first, int count, const phx_group_desc* next,
                                     hipStream_t stream) {
  int r = check_desc(d);
  if (r) return r;
  r = check_desc(next);
  if (r) return r;
  if (first < 0 || count < 1 || first + count > d->n_slices || !d->chi_a || !d->grad_obj ||
      (d->n_slices > 1 && !d->chi_slices) || (d->psi_slots > 0 && d->psi_slots < d->n_slices))
    return PHX_EINVAL;
  if (!next->gom || !next->intensity || !next->errors || next->red == d->red) return PHX_EINVAL;
  const dim3 grid((unsigned)((d->obj_w + UPD_TW - 1) / UPD_TW), (unsigned)((d->obj_h + UPD_TH - 1) / UPD_TH),
                  (unsigned)count);
  hipLaunchKernelGGL(gradient_object_gather_k, grid, dim3(256), 0, stream, *d, first, begin_span(*next));
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_lsqml_object_update(const phx_group_desc* d, int slice, hipStream_t stream) {
  return phx_lsqml_object_update_slices(d, slice, 1, stream);
}

int phx_probe_spectrum_reduce(const phx_group_desc* d, hipStream_t stream) {
  int r = check_desc(d);
  if (r) return r;
  if (!d->farfield || !d->d_probe) return PHX_EINVAL;
  const dim3 grid((unsigned)((d->n * d->n + 255) / 256), (unsigned)d->n_modes);
  hipLaunchKernelGGL(probe_spectrum_reduce_k, grid, dim3(256), 0, stream, *d);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_lsqml_probe_apply(const phx_group_desc* d, hipStream_t stream) {
  int r = check_desc(d);
  if (r) return r;
  if (!d->d_probe || !d->gom || !d->finite_flag) return PHX_EINVAL;
  if (d->n == S256) return s256_lsqml_probe_apply(*d, stream);
#define L(N, T) hipLaunchKernelGGL((lsqml_probe_apply_k<N, T>), dim3((unsigned)d->n_modes), dim3(T), 0, stream, *d)
  PHX_DISPATCH(d->n, L);
#undef L
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_lsqml_probe_apply_strips(const phx_group_desc* d, void* tmp, hipStream_t stream) {
  int r = check_desc(d);
  if (r) return r;
  if (!d->d_probe || !d->gom || !d->finite_flag || !d->probe_k || !tmp) return PHX_EINVAL;
  float2* t = (float2*)tmp;
#define L(N, T_)                                                                                         \
  {                                                                                                      \
    const dim3 grid((unsigned)(N / Strip<N>::R), (unsigned)d->n_modes), block(Strip<N>::T);              \
    hipLaunchKernelGGL((probe_apply_cols_inv_k<N>), grid, block, 0, stream, *d, t);                      \
    PHX_CHECK_LAUNCH();                                                                                  \
    hipLaunchKernelGGL((probe_apply_rows_k<N>), grid, block, 0, stream, *d, t);                          \
    PHX_CHECK_LAUNCH();                                                                                  \
    hipLaunchKernelGGL((probe_apply_cols_fwd_k<N>), grid, block, 0, stream, *d, t);                      \
  }
  switch (d->n) {
    case 16: L(16, 0); break;
    case 32: L(32, 0); break;
    case 64: L(64, 0); break;
    case 128: L(128, 0); break;
    default: return PHX_EUNSUPPORTED;
  }
#undef L
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int64_t phx_probe_strip_scratch_elems(void) { return PROBE_STRIP_SCRATCH_C64; }

int phx_lsqml_probe_apply_strips_static(const phx_group_desc* d, hipStream_t stream) {
  int r = check_desc(d);
  if (r) return r;
  if (d->n > 128) return PHX_EUNSUPPORTED;
  if ((int64_t)d->n_modes * d->n * d->n > PROBE_STRIP_SCRATCH_C64) return PHX_ETOOLARGE;
  float2* t = nullptr;
  r = probe_strip_scratch_ptr(&t);
  if (r) return r;
  // every element of the work plane is written (inverse column pass) before it is read,
  // so nothing left by an earlier call needs clearing
  return phx_lsqml_probe_apply_strips(d, t, stream);
}

int phx_epie_backward(const phx_group_desc* d, int slice, int update_object, int update_probe,
                      hipStream_t stream) {
  int r = check_desc(d);
  if (r) return r;
  if (slice < 0 || slice >= d->n_slices || !d->chi_a || !d->chi_b || !d->aux ||
      (update_probe && (!d->d_probe || !d->farfield)))
    return PHX_EINVAL;
  hipLaunchKernelGGL(epie_maxima_k, dim3((unsigned)d->bmax), dim3(256), 0, stream, *d, slice);
  PHX_CHECK_LAUNCH();
  if (d->n == S256) return s256_epie_backward(*d, slice, update_probe, stream);
#define L(N, T)                                                                                           \
  hipLaunchKernelGGL((epie_backward_k<N, T>), PHX_PLANE_GRID(d), dim3(T), 0, stream, *d, slice, update_object, \
                     update_probe)
  PHX_DISPATCH(d->n, L);
#undef L
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_probe_limit_support(const phx_group_desc* d, const float* mask_s, hipStream_t stream) {
  int r = check_desc(d);
  if (r) return r;
  if (!mask_s || !d->probe_k) return PHX_EINVAL;
#define L(N, T) hipLaunchKernelGGL((probe_support_k<N, T>), dim3((unsigned)d->n_modes), dim3(T), 0, stream, *d, mask_s)
  PHX_DISPATCH(d->n, L);
#undef L
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_epie_object_apply(const phx_group_desc* d, int slice, hipStream_t stream) {
  int r = check_desc(d);
  if (r) return r;
  if (slice < 0 || slice >= d->n_slices || !d->finite_flag || !d->aux || !d->chi_a || !d->chi_b) return PHX_EINVAL;
  const dim3 grid((unsigned)((d->obj_w + UPD_TW - 1) / UPD_TW), (unsigned)((d->obj_h + UPD_TH - 1) / UPD_TH));
  hipLaunchKernelGGL(epie_object_update_k, grid, dim3(256), 0, stream, *d, slice);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_epie_probe_apply(const phx_group_desc* d, hipStream_t stream) {
  int r = check_desc(d);
  if (r) return r;
  if (!d->d_probe || !d->finite_flag || !d->farfield) return PHX_EINVAL;
  // d_probe = the per-position spectra phx_epie_backward left in the farfield slots at
  // slice 0, summed in position order
  r = phx_probe_spectrum_reduce(d, stream);
  if (r) return r;
  if (d->n == S256) return s256_epie_probe_apply(*d, stream);
#define L(N, T) hipLaunchKernelGGL((epie_probe_apply_k<N, T>), dim3((unsigned)d->n_modes), dim3(T), 0, stream, *d)
  PHX_DISPATCH(d->n, L);
#undef L
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_dry_run_stats(const phx_group_desc* d, hipStream_t stream) {
  int r = check_desc(d);
  if (r) return r;
  if (!d->patterns_s || !d->mask_s || !d->errors) return PHX_EINVAL;
  hipLaunchKernelGGL(dry_run_stats_k, dim3((unsigned)d->bmax), dim3(256), 0, stream, *d);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_position_gradient(const phx_group_desc* d, int dir, hipStream_t stream) {
  int r = check_desc(d);
  if (r) return r;
  if ((dir != 0 && dir != 1) || !d->chi_a || !d->chi_b) return PHX_EINVAL;
  if (d->n == S256) return d->farfield ? s256_position_gradient(*d, dir, stream) : PHX_EINVAL;
#define L(N, T) hipLaunchKernelGGL((position_gradient_k<N, T>), PHX_PLANE_GRID(d), dim3(T), 0, stream, *d, dir)
  PHX_DISPATCH(d->n, L);
#undef L
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_farfield_gradient(const phx_group_desc* d, int poisson, hipStream_t stream) {
  int r = check_desc(d);
  if (r) return r;
  if (!d->farfield || !d->intensity || !d->patterns_s || !d->mask_s || !d->chi_a) return PHX_EINVAL;
  if (d->n == S256) return s256_farfield_gradient(*d, poisson, stream);
#define L(N, T) hipLaunchKernelGGL((farfield_fwd_k<N, T, false>), PHX_PLANE_GRID(d), dim3(T), 0, stream, *d)
  PHX_DISPATCH(d->n, L);
#undef L
  PHX_CHECK_LAUNCH();
  const int mode = poisson ? (int)FM_GRAD_POISSON : (int)FM_GRAD_AMPLITUDE;
#define L(N, T)                                                                                                  \
  do {                                                                                                           \
    if (mode == FM_GRAD_POISSON)                                                                                 \
      hipLaunchKernelGGL((farfield_update_k<N, T, FM_GRAD_POISSON>), PHX_PLANE_GRID(d), dim3(T), 0, stream, *d, 0);   \
    else                                                                                                         \
      hipLaunchKernelGGL((farfield_update_k<N, T, FM_GRAD_AMPLITUDE>), PHX_PLANE_GRID(d), dim3(T), 0, stream, *d, 0); \
  } while (0)
  PHX_DISPATCH(d->n, L);
#undef L
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_gradient_backward(const phx_group_desc* d, int slice, int want_object, int want_probe, int want_positions,
                          hipStream_t stream) {
  int r = check_desc(d);
  if (r) return r;
  if (slice < 0 || slice >= d->n_slices || !d->chi_a || !d->chi_b || (want_object && !d->grad_obj) ||
      ((want_probe || want_positions) && (!d->d_probe || !d->probe_k)) || (want_probe && !d->farfield))
    return PHX_EINVAL;
  // 256 x 256: the object gradient is only gathered (phx_gradient_object_gather), not scattered
  if (d->n == S256) return want_object ? PHX_EUNSUPPORTED : s256_gradient_backward(*d, slice, want_probe, want_positions, stream);
#define L(N, T)                                                                                              \
  hipLaunchKernelGGL((gradient_backward_k<N, T>), PHX_PLANE_GRID(d), dim3(T), 0, stream, *d, slice, want_object, \
                     want_probe, want_positions)
  PHX_DISPATCH(d->n, L);
#undef L
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_gradient_probe_finish(const phx_group_desc* d, hipStream_t stream) {
  int r = check_desc(d);
  if (r) return r;
  if (!d->d_probe || !d->grad_probe || !d->farfield) return PHX_EINVAL;
  // d_probe = the per-position spectra the adjoint left in the farfield slots, summed in
  // position order
  r = phx_probe_spectrum_reduce(d, stream);
  if (r) return r;
  if (d->n == S256) return s256_gradient_probe_finish(*d, stream);
#define L(N, T) \
  hipLaunchKernelGGL((gradient_probe_finish_k<N, T>), dim3((unsigned)d->n_modes), dim3(T), 0, stream, *d)
  PHX_DISPATCH(d->n, L);
#undef L
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_dry_run_stats_gradient(const phx_group_desc* d, hipStream_t stream) {
  int r = check_desc(d);
  if (r) return r;
  if (!d->patterns_s || !d->errors) return PHX_EINVAL;
  hipLaunchKernelGGL(dry_run_stats_gradient_k, dim3((unsigned)d->bmax), dim3(256), 0, stream, *d);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_dry_run_stats_intensity(const phx_group_desc* d, hipStream_t stream) {
  int r = check_desc(d);
  if (r) return r;
  if (!d->patterns_s || !d->errors || !d->intensity) return PHX_EINVAL;
  hipLaunchKernelGGL(dry_run_stats_intensity_k, dim3((unsigned)d->bmax), dim3(256), 0, stream, *d);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_exit_intensity(const phx_group_desc* d, float* out, hipStream_t stream) {
  int r = check_desc(d);
  if (r) return r;
  if (!out || !d->farfield || !d->intensity) return PHX_EINVAL;
  const dim3 grid((unsigned)((d->n * d->n + 255) / 256), (unsigned)d->bmax);
  if (d->n == S256) {
    r = s256_exit_spectra(*d, stream);
    if (r) return r;
    hipLaunchKernelGGL((exit_intensity_copy_k<S256>), grid, dim3(256), 0, stream, *d, out);
    PHX_CHECK_LAUNCH();
    return PHX_OK;
  }
#define L(N, T) hipLaunchKernelGGL((farfield_fwd_k<N, T, false>), PHX_PLANE_GRID(d), dim3(T), 0, stream, *d)
  PHX_DISPATCH(d->n, L);
#undef L
  PHX_CHECK_LAUNCH();
#define L(N, T) hipLaunchKernelGGL((exit_intensity_copy_k<N>), grid, dim3(256), 0, stream, *d, out)
  PHX_DISPATCH(d->n, L);
#undef L
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_position_finish(const phx_group_desc* d, void* pos_update, hipStream_t stream) {
  int r = check_desc(d);
  if (r) return r;
  if (!pos_update) return PHX_EINVAL;
  hipLaunchKernelGGL(position_finish_k, dim3((unsigned)((d->bmax + 63) / 64)), dim3(64), 0, stream, *d,
                     (float*)pos_update);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

}  // extern "C"
