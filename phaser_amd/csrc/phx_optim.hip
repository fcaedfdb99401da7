// Elementwise optimizer steps of the gradient engine on complex64 parameters.
//
// The reference wraps optax (phaser/engines/gradient/solvers.py:88-138): SGD =
// [trace(momentum, nesterov)] + scale_by_learning_rate(lr, flip_sign=False); Adam =
// scale_by_adam(b1, b2, eps, eps_root, nesterov) + scale_by_learning_rate; Polyak =
// scale_by_learning_rate(scaling) + scale_by_polyak(max_lr, f_min, eps).  The engine feeds
// them the steepest-descent direction d = -conj(g_jax) / scale (gradient/run.py:325-330),
// which in torch's gradient convention is d = -g / scale.  These kernels take g (torch
// convention) and `scale`, form d, apply the transform and add the update to the
// parameter in place; a non-finite result records `flag_value` in finite_flag.
#include "phx_common.h"
#include "../../include/phaser_mi355x.h"

using namespace phx;

namespace {

__device__ __forceinline__ void flag_bad(bool bad, int32_t* flag, int32_t v) {
  if (bad && flag) atomicMin(flag, v);
}

__global__ void sgd_k(float2* __restrict__ p, const float2* __restrict__ g, float2* __restrict__ tr, int64_t n,
                      float inv_scale, float lr, float momentum, int nesterov, int use_momentum, int32_t* flag,
                      int32_t fv) {
  bool bad = false;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float2 dv = cscale(g[i], -inv_scale);
    float2 u = dv;
    if (use_momentum) {
      const float2 t = cadd(dv, cscale(tr[i], momentum));  // optax.trace: g + decay * t
      tr[i] = t;
      u = nesterov ? cadd(dv, cscale(t, momentum)) : t;
    }
    const float2 np = cadd(p[i], cscale(u, lr));
    p[i] = np;
    bad |= !(isfinite(np.x) && isfinite(np.y));
  }
  flag_bad(bad, flag, fv);
}

__global__ void adam_k(float2* __restrict__ p, const float2* __restrict__ g, float2* __restrict__ mu,
                       float* __restrict__ nu, int64_t n, float inv_scale, float lr, float b1, float b2, float eps,
                       float eps_root, float bc1, float bc1_next, float bc2, int nesterov, int32_t* flag,
                       int32_t fv) {
  bool bad = false;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float2 dv = cscale(g[i], -inv_scale);
    const float2 m1 = cadd(cscale(dv, 1.0f - b1), cscale(mu[i], b1));
    const float v1 = (1.0f - b2) * abs2f(dv) + b2 * nu[i];
    mu[i] = m1;
    nu[i] = v1;
    float2 mh;
    if (nesterov) mh = cadd(cscale(m1, b1 / bc1_next), cscale(dv, (1.0f - b1) / bc1));
    else mh = cscale(m1, 1.0f / bc1);
    const float vh = v1 / bc2;
    const float2 u = cscale(mh, 1.0f / (sqrtf(vh + eps_root) + eps));
    const float2 np = cadd(p[i], cscale(u, lr));
    p[i] = np;
    bad |= !(isfinite(np.x) && isfinite(np.y));
  }
  flag_bad(bad, flag, fv);
}

__global__ void sqnorm_k(const float2* __restrict__ g, int64_t n, float inv_scale, double* acc) {
  __shared__ double scratch[4];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    s += (double)abs2f(cscale(g[i], inv_scale));
  s = block_sum_f64(s, scratch);
  if (threadIdx.x == 0) atomic_add_f64(acc, s);
}

__global__ void polyak_k(float2* __restrict__ p, const float2* __restrict__ g, int64_t n, float inv_scale,
                         float scaling, float max_lr, float f_min, float eps, const double* sqnorm,
                         const double* loss, int32_t* flag, int32_t fv) {
  // sqnorm holds sum |d|^2 of the unscaled directions; the Polyak step sees scaling * d
  const double sq = (double)scaling * (double)scaling * sqnorm[0];
  // optax: where(sq + eps <= finfo(float).eps, 0, min((f - f_min)/(sq + eps), max_lr)),
  // with the FP64 machine epsilon (the reference enables x64)
  const double f64eps = 2.220446049250313e-16;
  const double step = (sq + eps <= f64eps) ? 0.0 : fmin((loss[0] - (double)f_min) / (sq + eps), (double)max_lr);
  const float k = (float)step * scaling;
  bool bad = false;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float2 np = cadd(p[i], cscale(g[i], -inv_scale * k));
    p[i] = np;
    bad |= !(isfinite(np.x) && isfinite(np.y));
  }
  flag_bad(bad, flag, fv);
}

unsigned blocks_for(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > 4096) b = 4096;
  if (b < 1) b = 1;
  return (unsigned)b;
}

}  // namespace

extern "C" {

int phx_sgd_step(void* param, const void* grad, void* trace, int64_t n, float scale, float lr, float momentum,
                 int nesterov, int use_momentum, int32_t* finite_flag, int32_t flag_value, hipStream_t stream) {
  if (!param || !grad || n < 0 || scale == 0.f || (use_momentum && !trace)) return PHX_EINVAL;
  if (n == 0) return PHX_OK;
  hipLaunchKernelGGL(sgd_k, dim3(blocks_for(n)), dim3(256), 0, stream, (float2*)param, (const float2*)grad,
                     (float2*)trace, n, 1.0f / scale, lr, momentum, nesterov, use_momentum, finite_flag, flag_value);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_adam_step(void* param, const void* grad, void* mu, float* nu, int64_t n, float scale, float lr, float b1,
                  float b2, float eps, float eps_root, int32_t count, int nesterov, int32_t* finite_flag,
                  int32_t flag_value, hipStream_t stream) {
  if (!param || !grad || !mu || !nu || n < 0 || scale == 0.f || count < 1) return PHX_EINVAL;
  if (n == 0) return PHX_OK;
  // bias corrections 1 - b^count (optax bias_correction), in FP64 on the host
  const float bc1 = (float)(1.0 - pow((double)b1, (double)count));
  const float bc1n = (float)(1.0 - pow((double)b1, (double)count + 1.0));
  const float bc2 = (float)(1.0 - pow((double)b2, (double)count));
  hipLaunchKernelGGL(adam_k, dim3(blocks_for(n)), dim3(256), 0, stream, (float2*)param, (const float2*)grad,
                     (float2*)mu, nu, n, 1.0f / scale, lr, b1, b2, eps, eps_root, bc1, bc1n, bc2, nesterov,
                     finite_flag, flag_value);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_sqnorm_accumulate(const void* grad, int64_t n, float scale, double* acc, hipStream_t stream) {
  if (!grad || !acc || n < 0 || scale == 0.f) return PHX_EINVAL;
  if (n == 0) return PHX_OK;
  hipLaunchKernelGGL(sqnorm_k, dim3(blocks_for(n)), dim3(256), 0, stream, (const float2*)grad, n, 1.0f / scale, acc);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

int phx_polyak_step(void* param, const void* grad, int64_t n, float scale, float scaling, float max_lr, float f_min,
                    float eps, const double* sqnorm, const double* loss, int32_t* finite_flag, int32_t flag_value,
                    hipStream_t stream) {
  if (!param || !grad || !sqnorm || !loss || n < 0 || scale == 0.f) return PHX_EINVAL;
  if (n == 0) return PHX_OK;
  hipLaunchKernelGGL(polyak_k, dim3(blocks_for(n)), dim3(256), 0, stream, (float2*)param, (const float2*)grad, n,
                     1.0f / scale, scaling, max_lr, f_min, eps, sqnorm, loss, finite_flag, flag_value);
  PHX_CHECK_LAUNCH();
  return PHX_OK;
}

}  // extern "C"
