// Learned depth-interval proposal: a 102-64-64-64-2 MLP with bf16 weights and
// activations and fp32 accumulation on the gfx950 matrix cores
// (v_mfma_f32_16x16x32_bf16), trained on the GPU with Adam on the per-ray
// intervals the tracer emits (tt_trace_fwd's interval_out). No counterpart in the
// reference, whose OptiX traversal has no proposal stage.
//
// Network: x = enc(o, d) in R^102 -> relu(W1 x + b1) -> relu(W2 . + b2) ->
// relu(W3 . + b3) -> W4 . + b4 = (y_start, y_end), distances / t_scale.
// enc: (o / enc_scale, d), then sin and cos of 2^f pi v for f = 0..7 and each of
// the six values v (6 + 96 = 102 features, zero-padded to 128).
//
// MFMA tiling (one wave64, 16x16x32 bf16): A = activations, 16 rays x 32
// features, lane l holds row l%16, features 8(l/16)..+7; B = weights, 32 x 16,
// lane l holds rows 8(l/16)..+7 of column l%16; D = 16x16 fp32, lane l holds rows
// 4(l/16)..+3 of column l%16. Weights are kept in global memory pre-packed in
// that B-fragment order (16 contiguous bytes per lane), once as W^T for the
// forward and once as W for the backward products dH * W. Activations go
// through LDS between layers (D layout -> A layout).
//
// Training step (batch B, B % 32 == 0): one wave per 32-ray tile runs the
// forward, the pinball loss, and the backward; the weight gradients
// dW = dH^T H_prev are MFMAs with the 32 rays as the K dimension. Each tile
// writes its gradient to its own row of a partials buffer; one thread per
// parameter then sums the rows in a fixed order and applies Adam (deterministic).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/tritracer_hip.h"
#include "tt_internal.h"

namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kIn = 102, kInPad = 128, kHid = 64, kOut = 2;
constexpr int oW1 = 0, oB1 = oW1 + kHid * kIn, oW2 = oB1 + kHid, oB2 = oW2 + kHid * kHid, oW3 = oB2 + kHid,
              oB3 = oW3 + kHid * kHid, oW4 = oB3 + kHid, oB4 = oW4 + kOut * kHid, kParams = oB4 + kOut;
static_assert(kParams == 15042, "parameter count");
// fragment slots (64 lanes x 8 bf16 each)
constexpr int fL1 = 0, fL2 = 16, fL3 = 24, fL4 = 32, fB3 = 34, fB2 = 42, kFrags = 50;
constexpr int kPartStride = kParams + 1;  // + the tile's loss sum

struct Hyper {
  float enc_scale, t_scale, t_skip, tau_s, tau_e;
};

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// sin / cos of 2 pi u. v_sin / v_cos take revolutions and are accurate for |u| <= 256
// only, so the whole revolutions are removed first (u - rint(u) is exact in fp32).
__device__ __forceinline__ float sin_rev(float u) { return __builtin_amdgcn_sinf(u - rintf(u)); }
__device__ __forceinline__ float cos_rev(float u) { return __builtin_amdgcn_cosf(u - rintf(u)); }

// Feature e of the encoding of v[6] (e >= 102: padding).
__device__ __forceinline__ float feature(const float v[6], int e) {
  if (e < 6) return v[e];
  if (e >= kIn) return 0.0f;
  const int q = e - 6, dim = q >> 4, f = (q & 15) >> 1;
  const float u = ldexpf(v[dim], f - 1);  // 2^f pi v = 2 pi u
  return (q & 1) ? cos_rev(u) : sin_rev(u);
}

__device__ __forceinline__ bf16x8 frag(const bf16x8 *F, int slot) {
  return F[slot * 64 + (threadIdx.x & 63)];
}

// A fragment of rows r0..r0+15, columns c0..c0+31 of a row-major bf16 LDS matrix.
__device__ __forceinline__ bf16x8 lds_a(const bf16 *M, int stride, int r0, int c0) {
  const int l = threadIdx.x & 63;
  return *reinterpret_cast<const bf16x8 *>(M + (r0 + (l & 15)) * stride + c0 + 8 * (l >> 4));
}
// A fragment of the transpose: A[m][k] = M[k0 + k][m0 + m] (k = rays).
__device__ __forceinline__ bf16x8 lds_at(const bf16 *M, int stride, int k0, int m0) {
  const int l = threadIdx.x & 63;
  bf16x8 v;
#pragma unroll
  for (int j = 0; j < 8; j++) v[j] = M[(k0 + 8 * (l >> 4) + j) * stride + m0 + (l & 15)];
  return v;
}

// Forward of one dense layer for 16 rays (rows r0..r0+15): acc = bias + A * W^T.
template <int KT, int NT>
__device__ __forceinline__ void dense(const bf16 *Ain, int sin, int r0, const bf16x8 *F, int fslot,
                                      const float *bias, int nvalid, f32x4 acc[NT]) {
  const int l = threadIdx.x & 63;
#pragma unroll
  for (int nt = 0; nt < NT; nt++) {
    const int col = 16 * nt + (l & 15);
    const float b = col < nvalid ? bias[col] : 0.0f;
    acc[nt] = f32x4{b, b, b, b};
  }
#pragma unroll
  for (int kt = 0; kt < KT; kt++) {
    const bf16x8 a = lds_a(Ain, sin, r0, 32 * kt);
#pragma unroll
    for (int nt = 0; nt < NT; nt++) acc[nt] = mfma(a, frag(F, fslot + kt * NT + nt), acc[nt]);
  }
}

// relu(acc) -> bf16 rows r0.. of H (D layout -> row-major).
template <int NT>
__device__ __forceinline__ void store_relu(const f32x4 acc[NT], bf16 *H, int sh, int r0) {
  const int l = threadIdx.x & 63;
#pragma unroll
  for (int nt = 0; nt < NT; nt++)
#pragma unroll
    for (int i = 0; i < 4; i++) H[(r0 + 4 * (l >> 4) + i) * sh + 16 * nt + (l & 15)] = (bf16)fmaxf(acc[nt][i], 0.0f);
}

constexpr int SX = kInPad + 8, SH = kHid + 8;  // LDS row strides (bf16), 16-byte multiples

// ---------------------------------------------------------------- inference
// Forward of one dense layer for 16 rays (rows 0..15 of Ain) with the layer's weight
// fragments and this lane's bias values held in registers: acc = bias + A * W^T.
template <int KT, int NT>
__device__ __forceinline__ void dense_reg(const bf16 *Ain, int sin, const bf16x8 (&w)[KT * NT], const float (&b)[NT],
                                          f32x4 acc[NT]) {
#pragma unroll
  for (int nt = 0; nt < NT; nt++) acc[nt] = f32x4{b[nt], b[nt], b[nt], b[nt]};
#pragma unroll
  for (int kt = 0; kt < KT; kt++) {
    const bf16x8 a = lds_a(Ain, sin, 0, 32 * kt);
#pragma unroll
    for (int nt = 0; nt < NT; nt++) acc[nt] = mfma(a, w[kt * NT + nt], acc[nt]);
  }
}
template <int NT>
__device__ __forceinline__ void load_bias(const float *bias, int nvalid, float (&b)[NT]) {
#pragma unroll
  for (int nt = 0; nt < NT; nt++) {
    const int col = 16 * nt + (threadIdx.x & 15);
    b[nt] = col < nvalid ? bias[col] : 0.0f;
  }
}

// Persistent waves (one per block, two per SIMD): each wave loads the network's 34 weight
// fragments (136 registers a lane) and its bias values once and then runs 16-ray tiles
// until the rays are done, so the layers read only LDS (the activations) and registers.
// The products, their order and the roundings are those of the training kernel's forward.
constexpr int kInferWaves = 2048;  // 256 CUs x 4 SIMDs x 2 waves
__global__ __launch_bounds__(64, 2) void k_prop_infer(const float *ro, const float *rd, int R, const bf16x8 *F,
                                                      const float *W, Hyper hp, float ms, float me, int mode,
                                                      float *out) {
  __shared__ __attribute__((aligned(16))) bf16 X[16 * SX];
  __shared__ __attribute__((aligned(16))) bf16 Ha[16 * SH];
  __shared__ __attribute__((aligned(16))) bf16 Hb[16 * SH];
  const int l = threadIdx.x;
  bf16x8 w1[16], w2[8], w3[8], w4[2];
#pragma unroll
  for (int i = 0; i < 16; i++) w1[i] = frag(F, fL1 + i);
#pragma unroll
  for (int i = 0; i < 8; i++) w2[i] = frag(F, fL2 + i), w3[i] = frag(F, fL3 + i);
#pragma unroll
  for (int i = 0; i < 2; i++) w4[i] = frag(F, fL4 + i);
  float b1[4], b2[4], b3[4], b4[1];
  load_bias(W + oB1, kHid, b1);
  load_bias(W + oB2, kHid, b2);
  load_bias(W + oB3, kHid, b3);
  load_bias(W + oB4, kOut, b4);
  const long long tiles = ((long long)R + 15) / 16;
  for (long long tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    const long long base = tile * 16;
    {  // encoding: lane l fills features 32 g .. 32 g + 31 of ray l % 16, g = l / 16 -- feature()
       // unrolled: feature 32 g + i is raw input i (g = 0, i < 6) or a sinusoid of input
       // 2g - 1 (i < 6), 2g (6 <= i < 22) or 2g + 1 (i >= 22), octave ((i + 10) & 15) / 2,
       // cos for odd i; the three inputs are picked once per lane, the row stored 16 B at a time
      const long long ray = base + (l & 15);
      const int g = l >> 4;
      float v[6] = {0, 0, 0, 0, 0, 0};
      if (ray < R) {
        const float inv = 1.0f / hp.enc_scale;
        for (int c = 0; c < 3; c++) v[c] = ro[3 * ray + c] * inv, v[3 + c] = rd[3 * ray + c];
      }
      float vd[3];
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const int dim = 2 * g - 1 + k;
        float x = 0.0f;
#pragma unroll
        for (int c = 0; c < 6; c++) x = dim == c ? v[c] : x;
        vd[k] = x;
      }
      bf16x8 pk[4];
#pragma unroll
      for (int i = 0; i < 32; i++) {
        const float u = ldexpf(vd[i < 6 ? 0 : (i < 22 ? 1 : 2)], (((i + 10) & 15) >> 1) - 1);
        float val = (i & 1) ? cos_rev(u) : sin_rev(u);
        if (i < 6 && g == 0) val = v[i];
        if (32 * g + i >= kIn || ray >= R) val = 0.0f;
        pk[i >> 3][i & 7] = (bf16)val;
      }
#pragma unroll
      for (int j = 0; j < 4; j++) *reinterpret_cast<bf16x8 *>(X + (l & 15) * SX + 32 * g + 8 * j) = pk[j];
    }
    __syncthreads();
    f32x4 acc[4];
    dense_reg<4, 4>(X, SX, w1, b1, acc);
    store_relu<4>(acc, Ha, SH, 0);
    __syncthreads();
    dense_reg<2, 4>(Ha, SH, w2, b2, acc);
    store_relu<4>(acc, Hb, SH, 0);
    __syncthreads();
    dense_reg<2, 4>(Hb, SH, w3, b3, acc);
    store_relu<4>(acc, Ha, SH, 0);
    __syncthreads();
    f32x4 y[1];
    dense_reg<2, 1>(Ha, SH, w4, b4, y);
    // lane l holds y[row 4(l/16)+i][col l%16]; columns 0 and 1 are the outputs
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const float other = __shfl_down(y[0][i], 1);
      const long long ray = base + 4 * (l >> 4) + i;
      if ((l & 15) == 0 && ray < R) {
        if (mode == 0) {
          out[2 * ray] = y[0][i];
          out[2 * ray + 1] = other;
        } else {
          out[2 * ray] = mode == 2 ? 0.0f : fmaxf(0.0f, (y[0][i] - ms) * hp.t_scale);
          out[2 * ray + 1] = (other + me) * hp.t_scale;
        }
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------- training
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Pinball loss of quantile tau at residual y - u (weight 1): value and derivative.
__device__ __forceinline__ float pinball(float y, float u, float tau, float &g) {
  const float r = y - u;
  g = r > 0.0f ? (1.0f - tau) : (r < 0.0f ? -tau : 0.0f);
  return r > 0.0f ? (1.0f - tau) * r : -tau * r;
}

// dW = dH^T Hp over the 32 rays of the tile: M = outputs (MT tiles), N = inputs (NT
// tiles, nin valid), K = rays. Writes rows of the tile's partial gradient.
template <int MT, int NT>
__device__ void weight_grad(const bf16 *dH, int sdh, const bf16 *Hp, int shp, int nin, float *part, int off) {
  const int l = threadIdx.x & 63;
  bf16x8 b[NT];
#pragma unroll
  for (int nt = 0; nt < NT; nt++) b[nt] = lds_at(Hp, shp, 0, 16 * nt);
#pragma unroll
  for (int mt = 0; mt < MT; mt++) {
    const bf16x8 a = lds_at(dH, sdh, 0, 16 * mt);
#pragma unroll
    for (int nt = 0; nt < NT; nt++) {
      f32x4 d = mfma(a, b[nt], f32x4{0.f, 0.f, 0.f, 0.f});
      const int col = 16 * nt + (l & 15);
      if (col < nin)
#pragma unroll
        for (int i = 0; i < 4; i++) part[off + (16 * mt + 4 * (l >> 4) + i) * nin + col] = d[i];
    }
  }
}

// dHp = (dH * W) masked by Hp > 0, for the 32 rays (two row tiles): fp32 copy in
// dHf (for the bias gradient) and bf16 copy in dHo.
__device__ void back_layer(const bf16 *dH, const bf16x8 *F, int fslot, const bf16 *Hp, float *dHf, bf16 *dHo) {
  const int l = threadIdx.x & 63;
  for (int mt = 0; mt < 2; mt++) {
    f32x4 acc[4];
#pragma unroll
    for (int nt = 0; nt < 4; nt++) acc[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < 2; kt++) {
      const bf16x8 a = lds_a(dH, SH, 16 * mt, 32 * kt);
#pragma unroll
      for (int nt = 0; nt < 4; nt++) acc[nt] = mfma(a, frag(F, fslot + kt * 4 + nt), acc[nt]);
    }
#pragma unroll
    for (int nt = 0; nt < 4; nt++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int r = 16 * mt + 4 * (l >> 4) + i, c = 16 * nt + (l & 15);
        const float g = (float)Hp[r * SH + c] > 0.0f ? acc[nt][i] : 0.0f;
        dHf[r * kHid + c] = g;
        dHo[r * SH + c] = (bf16)g;
      }
  }
}

// Bias gradient: column sums of dHf (lane l -> column l).
__device__ void bias_grad(const float *dHf, float *part, int off) {
  const int l = threadIdx.x & 63;
  float s = 0.0f;
  for (int r = 0; r < 32; r++) s += dHf[r * kHid + l];
  part[off + l] = s;
}

// One 32-ray tile per block (one wave). sample != 0: rays drawn by hash from
// [0, n_rays); else rays blockIdx.x*32 + i in order.
__global__ __launch_bounds__(64) void k_prop_train(const float *ro, const float *rd, const float4 *ival,
                                                   long long n_rays, int sample, unsigned long long seed,
                                                   const bf16x8 *F, const float *W, Hyper hp, float inv_batch,
                                                   float *partials) {
  __shared__ __attribute__((aligned(16))) bf16 X[32 * SX];
  __shared__ __attribute__((aligned(16))) bf16 H1[32 * SH];
  __shared__ __attribute__((aligned(16))) bf16 H2[32 * SH];
  __shared__ __attribute__((aligned(16))) bf16 H3[32 * SH];
  __shared__ __attribute__((aligned(16))) bf16 dA[32 * SH];
  __shared__ __attribute__((aligned(16))) bf16 dB[32 * SH];
  __shared__ float dHf[32 * kHid];
  __shared__ float V[32 * 6];
  __shared__ float Tg[32 * 2];
  __shared__ float Y[32 * 2];
  __shared__ float dY[32 * 2];
  __shared__ float Ls[32];
  const int l = threadIdx.x;
  float *part = partials + (size_t)blockIdx.x * kPartStride;
  if (l < 32) {
    long long ray = (long long)blockIdx.x * 32 + l;
    if (sample) ray = (long long)(mix64(seed ^ mix64((unsigned long long)ray)) % (unsigned long long)n_rays);
    const float inv = 1.0f / hp.enc_scale;
    for (int c = 0; c < 3; c++) V[l * 6 + c] = ro[3 * ray + c] * inv, V[l * 6 + 3 + c] = rd[3 * ray + c];
    const float4 iv = ival[ray];
    const bool has = iv.x < INFINITY;
    const float stop = iv.z < INFINITY ? iv.z : iv.y;
    Tg[2 * l] = (has ? iv.x : hp.t_skip) / hp.t_scale;
    Tg[2 * l + 1] = (has ? stop : hp.t_skip) / hp.t_scale;
  }
  __syncthreads();
  {  // encoding: lane l -> ray l%32, features 64*(l/32) .. +63
    const int r = l & 31;
    float v[6];
    for (int c = 0; c < 6; c++) v[c] = V[r * 6 + c];
    for (int e = 64 * (l >> 5); e < 64 * (l >> 5) + 64; e++) X[r * SX + e] = (bf16)feature(v, e);
  }
  __syncthreads();
  // forward, both row tiles per layer
  for (int mt = 0; mt < 2; mt++) {
    f32x4 acc[4];
    dense<4, 4>(X, SX, 16 * mt, F, fL1, W + oB1, kHid, acc);
    store_relu<4>(acc, H1, SH, 16 * mt);
  }
  __syncthreads();
  for (int mt = 0; mt < 2; mt++) {
    f32x4 acc[4];
    dense<2, 4>(H1, SH, 16 * mt, F, fL2, W + oB2, kHid, acc);
    store_relu<4>(acc, H2, SH, 16 * mt);
  }
  __syncthreads();
  for (int mt = 0; mt < 2; mt++) {
    f32x4 acc[4];
    dense<2, 4>(H2, SH, 16 * mt, F, fL3, W + oB3, kHid, acc);
    store_relu<4>(acc, H3, SH, 16 * mt);
  }
  __syncthreads();
  for (int mt = 0; mt < 2; mt++) {
    f32x4 y[1];
    dense<2, 1>(H3, SH, 16 * mt, F, fL4, W + oB4, kOut, y);
    if ((l & 15) < 2)
      for (int i = 0; i < 4; i++) Y[(16 * mt + 4 * (l >> 4) + i) * 2 + (l & 15)] = y[0][i];
  }
  __syncthreads();
  if (l < 32) {  // loss of ray l (mean over the batch)
    float g0, g1;
    const float L = pinball(Y[2 * l], Tg[2 * l], hp.tau_s, g0) + pinball(Y[2 * l + 1], Tg[2 * l + 1], hp.tau_e, g1);
    dY[2 * l] = g0 * inv_batch;
    dY[2 * l + 1] = g1 * inv_batch;
    Ls[l] = L * inv_batch;
  }
  __syncthreads();
  // layer 4 (VALU): dH3 = dY W4 masked; dW4, db4, db3
  for (int idx = l; idx < 32 * kHid; idx += 64) {
    const int r = idx / kHid, c = idx % kHid;
    const float w0 = (float)(bf16)W[oW4 + c], w1 = (float)(bf16)W[oW4 + kHid + c];
    const float g = (float)H3[r * SH + c] > 0.0f ? dY[2 * r] * w0 + dY[2 * r + 1] * w1 : 0.0f;
    dHf[r * kHid + c] = g;
    dA[r * SH + c] = (bf16)g;
  }
  for (int idx = l; idx < kOut * kHid; idx += 64) {
    const int o = idx / kHid, c = idx % kHid;
    float s = 0.0f;
    for (int r = 0; r < 32; r++) s += dY[2 * r + o] * (float)H3[r * SH + c];
    part[oW4 + idx] = s;
  }
  if (l < kOut) {
    float s = 0.0f;
    for (int r = 0; r < 32; r++) s += dY[2 * r + l];
    part[oB4 + l] = s;
  }
  if (l == 0) {
    float s = 0.0f;
    for (int r = 0; r < 32; r++) s += Ls[r];
    part[kParams] = s;
  }
  __syncthreads();
  bias_grad(dHf, part, oB3);
  weight_grad<4, 4>(dA, SH, H2, SH, kHid, part, oW3);
  __syncthreads();
  back_layer(dA, F, fB3, H2, dHf, dB);  // dH2
  __syncthreads();
  bias_grad(dHf, part, oB2);
  weight_grad<4, 4>(dB, SH, H1, SH, kHid, part, oW2);
  __syncthreads();
  back_layer(dB, F, fB2, H1, dHf, dA);  // dH1
  __syncthreads();
  bias_grad(dHf, part, oB1);
  weight_grad<4, 8>(dA, SH, X, SX, kIn, part, oW1);
}

// Sum of the partial rows per parameter (fixed order); mode 0: write the gradient;
// mode 1: Adam step on W (m, v state) at step t (1-based). Thread kParams sums the loss.
__global__ void k_prop_reduce(const float *partials, int tiles, int mode, float *grad, float *W, float *m, float *v,
                              float lr, float b1, float b2, float eps, int t, float *loss_out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > kParams) return;
  // the tiles' rows summed in tile order; 16 rows loaded ahead of their adds (same order, so
  // the same result bit for bit, with 16 loads in flight per thread instead of one)
  float g = 0.0f;
  int i = 0;
  for (; i + 16 <= tiles; i += 16) {
    float x[16];
#pragma unroll
    for (int j = 0; j < 16; j++) x[j] = partials[(size_t)(i + j) * kPartStride + p];
#pragma unroll
    for (int j = 0; j < 16; j++) g += x[j];
  }
  for (; i < tiles; i++) g += partials[(size_t)i * kPartStride + p];
  if (p == kParams) {
    if (loss_out) *loss_out = g;
    return;
  }
  if (mode == 0) {
    grad[p] = g;
    return;
  }
  const float mi = b1 * m[p] + (1.0f - b1) * g;
  const float vi = b2 * v[p] + (1.0f - b2) * g * g;
  m[p] = mi;
  v[p] = vi;
  const float mh = mi / (1.0f - powf(b1, (float)t)), vh = vi / (1.0f - powf(b2, (float)t));
  W[p] -= lr * mh / (sqrtf(vh) + eps);
}

// bf16 fragments from the fp32 parameters: forward slots hold W^T, backward slots W.
__global__ void k_prop_pack(const float *W, bf16 *F) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= kFrags * 64 * 8) return;
  const int slot = e / 512, lane = (e / 8) % 64, j = e % 8;
  const int kk = 8 * (lane >> 4) + j, nn = lane & 15;
  float val = 0.0f;
  auto fwd = [&](int s0, int NT, int wo, int kin, int nout) {
    const int s = slot - s0, kt = s / NT, nt = s % NT;
    const int k = 32 * kt + kk, n = 16 * nt + nn;
    val = (k < kin && n < nout) ? W[wo + n * kin + k] : 0.0f;
  };
  auto bwd = [&](int s0, int wo) {  // B[k = out][n = in] = W[out][in], 64 x 64
    const int s = slot - s0, kt = s / 4, nt = s % 4;
    val = W[wo + (32 * kt + kk) * kHid + 16 * nt + nn];
  };
  if (slot < fL2)
    fwd(fL1, 4, oW1, kIn, kHid);
  else if (slot < fL3)
    fwd(fL2, 4, oW2, kHid, kHid);
  else if (slot < fL4)
    fwd(fL3, 4, oW3, kHid, kHid);
  else if (slot < fB3)
    fwd(fL4, 1, oW4, kHid, kOut);
  else if (slot < fB2)
    bwd(fB3, oW3);
  else
    bwd(fB2, oW2);
  F[e] = (bf16)val;
}

int pfail(int code, const std::string &m) {
  tt::set_last_error(m);
  return code;
}
#define PCHK(x)                                                                          \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) return pfail(TT_ERR_HIP, std::string(#x ": ") + hipGetErrorString(e_)); \
  } while (0)
#define PARG(c, m)                                  \
  do {                                              \
    if (!(c)) return pfail(TT_ERR_INVALID_ARG, m); \
  } while (0)

}  // namespace

struct tt_proposal {
  int device = 0;
  tt_proposal_hparams hp{};
  float *W = nullptr, *m = nullptr, *v = nullptr;
  bf16 *F = nullptr;
  float *partials = nullptr;
  int part_tiles = 0;
  int t = 0;  // Adam steps taken
  unsigned long long seed = 0;
};

namespace {
Hyper hyper(const tt_proposal *p) {
  return Hyper{p->hp.enc_scale, p->hp.t_scale, p->hp.t_skip, p->hp.tau_start, p->hp.tau_end};
}
int ensure_partials(tt_proposal *p, int tiles) {
  if (tiles <= p->part_tiles) return TT_OK;
  if (p->partials) PCHK(hipFree(p->partials));
  p->partials = nullptr;
  PCHK(hipMalloc(&p->partials, (size_t)tiles * kPartStride * sizeof(float)));
  p->part_tiles = tiles;
  return TT_OK;
}
int pack(tt_proposal *p, hipStream_t s) {
  hipLaunchKernelGGL(k_prop_pack, dim3((kFrags * 512 + 255) / 256), dim3(256), 0, s, p->W, p->F);
  PCHK(hipGetLastError());
  return TT_OK;
}
}  // namespace

extern "C" {

int tt_proposal_num_params(void) {
  return kParams;
}

void tt_proposal_default_hparams(tt_proposal_hparams *hp) {
  if (!hp) return;
  hp->enc_scale = 8.0f;
  hp->t_scale = 8.0f;
  hp->t_skip = 16.0f;
  hp->tau_start = 0.1f;
  hp->tau_end = 0.9f;
  hp->lr = 1e-3f;
  hp->beta1 = 0.9f;
  hp->beta2 = 0.999f;
  hp->eps = 1e-8f;
}

int tt_proposal_create(tt_proposal **out, const tt_proposal_hparams *hp, unsigned long long seed) {
  PARG(out, "out is NULL");
  tt_proposal *p = new tt_proposal();
  if (hp)
    p->hp = *hp;
  else
    tt_proposal_default_hparams(&p->hp);
  p->seed = seed;
  auto bad = [&](int rc) {
    tt_proposal_destroy(p);
    return rc;
  };
  if (!(p->hp.enc_scale > 0.0f && p->hp.t_scale > 0.0f)) return bad(pfail(TT_ERR_INVALID_ARG, "scales must be > 0"));
  if (hipGetDevice(&p->device) != hipSuccess) return bad(pfail(TT_ERR_HIP, "hipGetDevice failed"));
  if (hipMalloc(&p->W, kParams * sizeof(float)) != hipSuccess || hipMalloc(&p->m, kParams * sizeof(float)) != hipSuccess ||
      hipMalloc(&p->v, kParams * sizeof(float)) != hipSuccess ||
      hipMalloc(&p->F, (size_t)kFrags * 512 * sizeof(bf16)) != hipSuccess)
    return bad(pfail(TT_ERR_HIP, "hipMalloc failed"));
  // He-uniform weights from a seeded generator, zero biases
  std::vector<float> h(kParams, 0.0f);
  unsigned long long z = seed * 0x9E3779B97F4A7C15ull + 1;
  auto uni = [&]() {
    z = z * 6364136223846793005ull + 1442695040888963407ull;
    return (float)((z >> 11) * (1.0 / 9007199254740992.0)) * 2.0f - 1.0f;
  };
  auto init = [&](int off, int nout, int nin) {
    const float a = std::sqrt(6.0f / (float)nin);
    for (int i = 0; i < nout * nin; i++) h[off + i] = a * uni();
  };
  init(oW1, kHid, kIn);
  init(oW2, kHid, kHid);
  init(oW3, kHid, kHid);
  init(oW4, kOut, kHid);
  const int rc = tt_proposal_set_params(p, h.data());
  if (rc != TT_OK) return bad(rc);
  *out = p;
  return TT_OK;
}

int tt_proposal_destroy(tt_proposal *p) {
  if (!p) return TT_OK;
  (void)hipFree(p->W);
  (void)hipFree(p->m);
  (void)hipFree(p->v);
  (void)hipFree(p->F);
  (void)hipFree(p->partials);
  delete p;
  return TT_OK;
}

int tt_proposal_set_params(tt_proposal *p, const float *host) {
  PARG(p && host, "NULL argument");
  PCHK(hipMemcpy(p->W, host, kParams * sizeof(float), hipMemcpyHostToDevice));
  std::vector<float> zero(kParams, 0.0f);
  PCHK(hipMemcpy(p->m, zero.data(), kParams * sizeof(float), hipMemcpyHostToDevice));
  PCHK(hipMemcpy(p->v, zero.data(), kParams * sizeof(float), hipMemcpyHostToDevice));
  p->t = 0;
  int rc = pack(p, nullptr);
  if (rc != TT_OK) return rc;
  PCHK(hipDeviceSynchronize());
  return TT_OK;
}

int tt_proposal_get_params(tt_proposal *p, float *host) {
  PARG(p && host, "NULL argument");
  PCHK(hipDeviceSynchronize());
  PCHK(hipMemcpy(host, p->W, kParams * sizeof(float), hipMemcpyDeviceToHost));
  return TT_OK;
}

int tt_proposal_forward(tt_proposal *p, const float *ray_o_dev, const float *ray_d_dev, long long R, float margin_start,
                        float margin_end, int mode, float *out_dev, void *stream) {
  PARG(p, "proposal is NULL");
  PARG(R >= 0, "R must be >= 0");
  PARG(mode == TT_PROPOSAL_RAW || mode == TT_PROPOSAL_INTERVAL || mode == TT_PROPOSAL_CAP, "unknown mode");
  if (R == 0) return TT_OK;
  PARG(ray_o_dev && ray_d_dev && out_dev, "NULL buffer");
  PARG(R <= (1ll << 31) - 16, "too many rays");
  const long long tiles = (R + 15) / 16;
  const long long blocks = tiles < kInferWaves ? tiles : kInferWaves;
  hipLaunchKernelGGL(k_prop_infer, dim3((unsigned)blocks), dim3(64), 0, static_cast<hipStream_t>(stream), ray_o_dev,
                     ray_d_dev, (int)R, reinterpret_cast<const bf16x8 *>(p->F), p->W, hyper(p), margin_start,
                     margin_end, mode, out_dev);
  PCHK(hipGetLastError());
  return TT_OK;
}

int tt_proposal_loss_grad(tt_proposal *p, const float *ray_o_dev, const float *ray_d_dev, const float *interval_dev,
                          int B, float *grad_dev, float *loss_dev, void *stream) {
  PARG(p, "proposal is NULL");
  PARG(B > 0 && B % 32 == 0, "B must be a positive multiple of 32");
  PARG(ray_o_dev && ray_d_dev && interval_dev && grad_dev, "NULL buffer");
  int rc = ensure_partials(p, B / 32);
  if (rc != TT_OK) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_prop_train, dim3(B / 32), dim3(64), 0, s, ray_o_dev, ray_d_dev,
                     reinterpret_cast<const float4 *>(interval_dev), (long long)B, 0, 0ull,
                     reinterpret_cast<const bf16x8 *>(p->F), p->W, hyper(p), 1.0f / (float)B, p->partials);
  PCHK(hipGetLastError());
  hipLaunchKernelGGL(k_prop_reduce, dim3((kParams + 64) / 64), dim3(64), 0, s, p->partials, B / 32, 0, grad_dev,
                     p->W, p->m, p->v, 0.0f, 0.0f, 0.0f, 0.0f, 0, loss_dev);
  PCHK(hipGetLastError());
  return TT_OK;
}

int tt_proposal_train(tt_proposal *p, const float *ray_o_dev, const float *ray_d_dev, const float *interval_dev,
                      long long n_rays, int batch, int steps, float *loss_hist_dev, void *stream) {
  PARG(p, "proposal is NULL");
  PARG(batch > 0 && batch % 32 == 0, "batch must be a positive multiple of 32");
  PARG(n_rays > 0 && steps >= 0, "n_rays must be > 0 and steps >= 0");
  PARG(ray_o_dev && ray_d_dev && interval_dev, "NULL buffer");
  int rc = ensure_partials(p, batch / 32);
  if (rc != TT_OK) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  for (int i = 0; i < steps; i++) {
    p->t++;
    const unsigned long long sd = p->seed * 0xD1B54A32D192ED03ull + (unsigned long long)p->t;
    hipLaunchKernelGGL(k_prop_train, dim3(batch / 32), dim3(64), 0, s, ray_o_dev, ray_d_dev,
                       reinterpret_cast<const float4 *>(interval_dev), n_rays, 1, sd,
                       reinterpret_cast<const bf16x8 *>(p->F), p->W, hyper(p), 1.0f / (float)batch, p->partials);
    hipLaunchKernelGGL(k_prop_reduce, dim3((kParams + 64) / 64), dim3(64), 0, s, p->partials, batch / 32, 1,
                       nullptr, p->W, p->m, p->v, p->hp.lr, p->hp.beta1, p->hp.beta2, p->hp.eps, p->t,
                       loss_hist_dev ? loss_hist_dev + i : nullptr);
    hipLaunchKernelGGL(k_prop_pack, dim3((kFrags * 512 + 255) / 256), dim3(256), 0, s, p->W, p->F);
    PCHK(hipGetLastError());
  }
  return TT_OK;
}

}  // extern "C"
