// Optimizer step of a training iteration, the step after the backward: Adam on the
// splat parameters (vertices, opacity, sigma, SH), fused over up to kMaxGroups parameter
// groups in one launch, optionally zeroing the gradients it consumed. The reference has
// no optimizer (its caller runs torch.optim.Adam on the tensors its autograd Function
// returns, diff_tritracer/__init__.py:205-227); the update follows torch.optim.Adam's
// single-tensor path (no weight decay, no amsgrad) step for step.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "../../include/tritracer_hip.h"
#include "tt_internal.h"

namespace {

constexpr int kMaxGroups = TT_ADAM_MAX_GROUPS;

struct AdamGroups {
  float *p[kMaxGroups];
  float *g[kMaxGroups];
  float *m[kMaxGroups];
  float *v[kMaxGroups];
  long long n[kMaxGroups];
  float step_size[kMaxGroups];  // lr / (1 - beta1^t)
  int vec4;                     // every group's pointers are 16-byte aligned and n % 4 == 0
};

struct AdamScalars {
  float w1;         // 1 - beta1 (the lerp weight)
  float beta2, one_minus_beta2;
  float bc2_sqrt;   // sqrt(1 - beta2^t)
  float eps;
  int zero_grads;
};

// One element: exp_avg.lerp_(grad, 1 - beta1) (weight < 0.5: m + w (g - m)),
// exp_avg_sq.mul_(beta2).addcmul_(grad, grad, 1 - beta2),
// denom = sqrt(v) / sqrt(bc2) + eps, param.addcdiv_(m, denom, -step_size).
__device__ __forceinline__ void adam1(float &p, float g, float &m, float &v, float step_size, const AdamScalars &s) {
  m = s.w1 < 0.5f ? m + s.w1 * (g - m) : g - (g - m) * (1.0f - s.w1);
  v = v * s.beta2 + (s.one_minus_beta2 * g) * g;
  const float denom = sqrtf(v) / s.bc2_sqrt + s.eps;
  p = p + (-step_size) * (m / denom);
}

__global__ __launch_bounds__(256) void k_adam(AdamGroups G, AdamScalars s) {
  const int gi = blockIdx.y;
  float *p = G.p[gi], *g = G.g[gi], *m = G.m[gi], *v = G.v[gi];
  const long long n = G.n[gi];
  const float ss = G.step_size[gi];
  const long long stride = (long long)gridDim.x * blockDim.x;
  if (G.vec4) {
    float4 *p4 = reinterpret_cast<float4 *>(p), *g4 = reinterpret_cast<float4 *>(g);
    float4 *m4 = reinterpret_cast<float4 *>(m), *v4 = reinterpret_cast<float4 *>(v);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n / 4; i += stride) {
      float4 P = p4[i], Gr = g4[i], M = m4[i], V = v4[i];
      adam1(P.x, Gr.x, M.x, V.x, ss, s);
      adam1(P.y, Gr.y, M.y, V.y, ss, s);
      adam1(P.z, Gr.z, M.z, V.z, ss, s);
      adam1(P.w, Gr.w, M.w, V.w, ss, s);
      p4[i] = P;
      m4[i] = M;
      v4[i] = V;
      if (s.zero_grads) g4[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
  } else {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
      float P = p[i], M = m[i], V = v[i];
      adam1(P, g[i], M, V, ss, s);
      p[i] = P;
      m[i] = M;
      v[i] = V;
      if (s.zero_grads) g[i] = 0.0f;
    }
  }
}

bool aligned16(const void *q) { return (reinterpret_cast<uintptr_t>(q) & 15u) == 0; }

}  // namespace

extern "C" int tt_adam_step(const tt_adam_group *groups, int n_groups, int step, double beta1, double beta2,
                            double eps, int zero_grads, void *stream) {
  if (n_groups < 0 || n_groups > kMaxGroups || (n_groups > 0 && !groups)) {
    tt::set_last_error("n_groups must be in [0, " + std::to_string(kMaxGroups) + "] with a group array");
    return TT_ERR_INVALID_ARG;
  }
  if (step < 1 || !(beta1 >= 0.0 && beta1 < 1.0) || !(beta2 >= 0.0 && beta2 < 1.0) || !(eps >= 0.0)) {
    tt::set_last_error("step must be >= 1, beta1 and beta2 in [0, 1), eps >= 0");
    return TT_ERR_INVALID_ARG;
  }
  AdamGroups G{};
  long long nmax = 0;
  bool vec4 = true;
  for (int i = 0; i < n_groups; i++) {
    const tt_adam_group &a = groups[i];
    if (a.n < 0 || (a.n > 0 && (!a.param_dev || !a.grad_dev || !a.exp_avg_dev || !a.exp_avg_sq_dev))) {
      tt::set_last_error("adam group " + std::to_string(i) + ": NULL buffer or negative size");
      return TT_ERR_INVALID_ARG;
    }
    G.p[i] = a.param_dev;
    G.g[i] = a.grad_dev;
    G.m[i] = a.exp_avg_dev;
    G.v[i] = a.exp_avg_sq_dev;
    G.n[i] = a.n;
    // torch: bias_correction1 = 1 - beta1 ** step; step_size = lr / bias_correction1 (host doubles)
    G.step_size[i] = (float)(a.lr / (1.0 - std::pow(beta1, (double)step)));
    nmax = a.n > nmax ? a.n : nmax;
    vec4 = vec4 && a.n % 4 == 0 && aligned16(a.param_dev) && aligned16(a.grad_dev) && aligned16(a.exp_avg_dev) &&
           aligned16(a.exp_avg_sq_dev);
  }
  if (n_groups == 0 || nmax == 0) return TT_OK;
  G.vec4 = vec4 ? 1 : 0;
  AdamScalars s;
  // scalars as torch forms them: in double on the host, rounded once to float
  s.w1 = (float)(1.0 - beta1);
  s.beta2 = (float)beta2;
  s.one_minus_beta2 = (float)(1.0 - beta2);
  s.bc2_sqrt = (float)std::sqrt(1.0 - std::pow(beta2, (double)step));
  s.eps = (float)eps;
  s.zero_grads = zero_grads ? 1 : 0;
  const long long work = vec4 ? nmax / 4 : nmax;
  const long long blocks = (work + 255) / 256;
  const unsigned gx = (unsigned)(blocks < 65535 ? (blocks > 0 ? blocks : 1) : 65535);
  hipLaunchKernelGGL(k_adam, dim3(gx, (unsigned)n_groups), dim3(256), 0, static_cast<hipStream_t>(stream), G, s);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    tt::set_last_error(std::string("adam launch failed: ") + hipGetErrorString(e));
    return TT_ERR_HIP;
  }
  return TT_OK;
}
