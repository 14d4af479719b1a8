// Training-statistics consumer: prune and split triangles from the forward's per-triangle
// statistics (max_blending, density_factor, scaling: reference forward.cu:34-74 and
// :134-143 produce them; the reference has no consumer, its callers prune and densify in
// Python). Two HBM-bound steps:
//   plan  : classify every triangle (drop / keep / split candidate), grant splits in index
//           order within the output cap, and write the exclusive prefix sum of the output
//           counts (three launches: per-tile counts, one-block scan of the tile counts,
//           per-tile rescan and write);
//   apply : every block takes 256 consecutive input triangles, whose outputs are one
//           contiguous range (at most 1024 rows) because the offsets are monotone, maps
//           output rows to (parent, child) in LDS and streams each array through with
//           consecutive lanes on consecutive floats of the output (coalesced writes, reads
//           in parent order).
// Everything is integer work or exact float copies and midpoints ((x + y) * 0.5f), so the
// output is bit-identical to a CPU restatement of the same rules.
#include <hip/hip_runtime.h>

#include <climits>
#include <string>

#include "../../include/tritracer_hip.h"
#include "tt_internal.h"

namespace {

constexpr int kPlanThreads = 256, kPlanPer = 8, kPlanTile = kPlanThreads * kPlanPer;  // 2048 triangles per tile
constexpr int kApplyTris = 256;                                                       // input triangles per block
constexpr unsigned short kNoRow = 0xFFFF;
// LDS map entry of an output row: parent (index within the block's input range, 10 bits),
// child << 10 (0..3), split << 12
constexpr unsigned kMapParent = 0x3FF, kMapChildShift = 10, kMapSplit = 1u << 12;
static_assert(kApplyTris <= 1024, "parent index must fit the map entry");

struct PlanIn {
  const float *mb, *op, *dens, *scal, *score;
  const unsigned char *act_in;
  float min_blending, min_opacity, split_scaling, split_score;
  int P;
};

// keep (bit 0) and split candidate (bit 16) of triangle i, packed so that sums of up to
// 65535 triangles add in one int.
__device__ __forceinline__ int classify(const PlanIn &q, int i) {
  if (i >= q.P) return 0;
  if (q.act_in) {
    const unsigned a = q.act_in[i];
    return (a == TT_DENSIFY_KEEP || a == TT_DENSIFY_SPLIT ? 1 : 0) | (a == TT_DENSIFY_SPLIT ? 1 << 16 : 0);
  }
  const bool keep = !(q.mb[i] < q.min_blending) && !(q.op[i] < q.min_opacity);
  const bool cand = keep && q.scal && (!q.dens || q.dens[i] != 0.0f) && q.scal[i] > q.split_scaling &&
                    (!q.score || q.score[i] > q.split_score);
  return (keep ? 1 : 0) | (cand ? 1 << 16 : 0);
}

__device__ __forceinline__ int wave_sum(int v) {
  for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d);
  return v;
}

// Per-tile counts: work[tile] = kept triangles, work[nt + tile] = split candidates.
__global__ __launch_bounds__(kPlanThreads) void k_plan_count(PlanIn q, int *work, int nt) {
  __shared__ int s_w[kPlanThreads / 64];
  const int base = blockIdx.x * kPlanTile + threadIdx.x * kPlanPer;
  int s = 0;
#pragma unroll
  for (int e = 0; e < kPlanPer; e++) s += classify(q, base + e);
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < kPlanThreads / 64; w++) t += s_w[w];
    work[blockIdx.x] = t & 0xFFFF;
    work[nt + blockIdx.x] = t >> 16;
  }
}

// One block: exclusive prefix sums of the tile counts (in place), the number of splits
// granted (work[2 nt]) and the new count (offset[P]).
__global__ __launch_bounds__(1024) void k_plan_scan(int *work, int nt, long long cap, int *offset, int P) {
  __shared__ long long s_k[1024], s_c[1024];
  const int t = threadIdx.x;
  const int per = (nt + 1023) / 1024, b = min(nt, t * per), e = min(nt, b + per);
  long long sk = 0, sc = 0;
  for (int i = b; i < e; i++) {
    sk += work[i];
    sc += work[nt + i];
  }
  s_k[t] = sk;
  s_c[t] = sc;
  __syncthreads();
  for (int d = 1; d < 1024; d <<= 1) {  // Hillis-Steele inclusive scan
    const long long vk = t >= d ? s_k[t - d] : 0, vc = t >= d ? s_c[t - d] : 0;
    __syncthreads();
    s_k[t] += vk;
    s_c[t] += vc;
    __syncthreads();
  }
  long long rk = s_k[t] - sk, rc = s_c[t] - sc;
  for (int i = b; i < e; i++) {
    const int k = work[i], c = work[nt + i];
    work[i] = (int)rk;
    work[nt + i] = (int)rc;
    rk += k;
    rc += c;
  }
  if (t == 1023) {
    const long long keep = s_k[1023], cand = s_c[1023];
    long long allowed = cap > keep ? (cap - keep) / 3 : 0;
    allowed = allowed < cand ? allowed : cand;
    work[2 * nt] = (int)allowed;
    offset[P] = (int)(keep + 3 * allowed);
  }
}

// Per tile: rescan the flags, grant splits (candidate rank < allowed), write action and offset.
__global__ __launch_bounds__(kPlanThreads) void k_plan_write(PlanIn q, const int *work, int nt,
                                                              unsigned char *action, int *offset) {
  __shared__ int s_w[kPlanThreads / 64];
  const int base = blockIdx.x * kPlanTile + threadIdx.x * kPlanPer;
  int f[kPlanPer];
  int s = 0;
#pragma unroll
  for (int e = 0; e < kPlanPer; e++) {
    f[e] = classify(q, base + e);
    s += f[e];
  }
  // exclusive scan of the packed per-lane sums over the block
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int inc = s;
  for (int d = 1; d < 64; d <<= 1) {
    const int v = __shfl_up(inc, d);
    if (lane >= d) inc += v;
  }
  if (lane == 63) s_w[wv] = inc;
  __syncthreads();
  int run = inc - s;
  for (int w = 0; w < wv; w++) run += s_w[w];
  long long kp = (long long)work[blockIdx.x] + (run & 0xFFFF);
  long long sp = (long long)work[nt + blockIdx.x] + (run >> 16);
  const long long allowed = work[2 * nt];
#pragma unroll
  for (int e = 0; e < kPlanPer; e++) {
    const int i = base + e;
    if (i >= q.P) break;
    const bool keep = f[e] & 1, cand = f[e] >> 16;
    offset[i] = (int)(kp + 3 * (sp < allowed ? sp : allowed));
    action[i] = keep ? (cand && sp < allowed ? TT_DENSIFY_SPLIT : TT_DENSIFY_KEEP) : TT_DENSIFY_DROP;
    kp += keep ? 1 : 0;
    sp += cand ? 1 : 0;
  }
}

struct ApplyIn {
  tt_densify_array arr[TT_DENSIFY_MAX_ARRAYS];
  int n_arrays, P, P_out;
  const int *offset, *cumsum;
  int *npts_out, *cumsum_out;
  unsigned vec4;  // bit k: array k takes the float4 path
};

// Points of child k (0..3), corner v (0..2) as a pair (u, w) of parent corners, the point
// being (p_u + p_w) / 2 (u == w: the corner itself); 4 bits per entry, u in the low 2.
//   child 0 (a, ab, ca)   child 1 (ab, b, bc)   child 2 (ca, bc, c)   child 3 (ab, bc, ca)
constexpr unsigned long long sub_code(int u0, int w0, int u1, int w1, int u2, int w2) {
  return (unsigned long long)(u0 | w0 << 2) | (unsigned long long)(u1 | w1 << 2) << 4 |
         (unsigned long long)(u2 | w2 << 2) << 8;
}
constexpr unsigned long long kSubTable = sub_code(0, 0, 0, 1, 2, 0) | sub_code(0, 1, 1, 1, 1, 2) << 12 |
                                         sub_code(2, 0, 1, 2, 2, 2) << 24 | sub_code(0, 1, 1, 2, 2, 0) << 36;

constexpr int kApplyUnroll = 4;  // loads in flight per lane before the stores

// The outputs are written once and not read again by this launch: non-temporal vector
// stores (0.215 vs 0.223 ms against plain stores in an A/B run of the 1M-triangle
// prune + split; within run-to-run noise otherwise).
typedef float f4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void store_out(float *p, float v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void store_out4(float4 *p, float4 v) {
  const f4v x = {v.x, v.y, v.z, v.w};
  __builtin_nontemporal_store(x, reinterpret_cast<f4v *>(p));
}

// One array, one float per lane and step (any width, vertex or triangle rows). Each lane
// issues kApplyUnroll independent loads before its stores.
__device__ __forceinline__ void copy_rows(const tt_densify_array &d, unsigned w, unsigned rows,
                                          const unsigned short *s_map, const long long *s_vbase, int i0,
                                          long long o0, long long P_out) {
  const unsigned total = rows * w;
  for (unsigned b = threadIdx.x; b < total; b += 256 * kApplyUnroll) {
    float v[kApplyUnroll];
    long long dj[kApplyUnroll];
#pragma unroll
    for (int u = 0; u < kApplyUnroll; u++) {
      const unsigned f = b + u * 256;
      dj[u] = -1;
      v[u] = 0.0f;
      if (f >= total) continue;
      const unsigned r = f / w, c = f - r * w;
      const unsigned m = s_map[r];
      const long long j = o0 + r;
      if (m == kNoRow || j >= P_out) continue;
      dj[u] = j * w + c;
      const int pl = m & kMapParent, child = (m >> kMapChildShift) & 3;
      const bool split = m & kMapSplit;
      if (split && d.on_split == TT_SPLIT_ZERO) continue;
      if (d.rows == TT_ROWS_VERTEX) {
        const float *src = d.src_dev + s_vbase[pl];
        if (split && d.on_split == TT_SPLIT_SUBDIVIDE) {
          const int vtx = (int)c / 3, comp = (int)c - 3 * vtx;
          const unsigned code = (unsigned)(kSubTable >> (4 * (3 * child + vtx))) & 15u;
          const int cu = code & 3, cw = code >> 2;
          v[u] = cu == cw ? src[3 * cu + comp] : (src[3 * cu + comp] + src[3 * cw + comp]) * 0.5f;
        } else {
          v[u] = src[c];
        }
      } else {
        v[u] = d.src_dev[(size_t)(i0 + pl) * w + c];
      }
    }
#pragma unroll
    for (int u = 0; u < kApplyUnroll; u++)
      if (dj[u] >= 0) store_out(d.dst_dev + dj[u], v[u]);
  }
}

// Triangle rows of a width divisible by 4 with 16-byte aligned buffers: one float4 per lane
// and step.
__device__ __forceinline__ void copy_rows4(const tt_densify_array &d, unsigned w4, unsigned rows,
                                           const unsigned short *s_map, int i0, long long o0, long long P_out) {
  const float4 *src = reinterpret_cast<const float4 *>(d.src_dev);
  float4 *dst = reinterpret_cast<float4 *>(d.dst_dev);
  const unsigned total = rows * w4;
  for (unsigned b = threadIdx.x; b < total; b += 256 * kApplyUnroll) {
    float4 v[kApplyUnroll];
    long long dj[kApplyUnroll];
#pragma unroll
    for (int u = 0; u < kApplyUnroll; u++) {
      const unsigned f = b + u * 256;
      dj[u] = -1;
      v[u] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (f >= total) continue;
      const unsigned r = f / w4, c = f - r * w4;
      const unsigned m = s_map[r];
      const long long j = o0 + r;
      if (m == kNoRow || j >= P_out) continue;
      dj[u] = j * w4 + c;
      if ((m & kMapSplit) && d.on_split == TT_SPLIT_ZERO) continue;
      v[u] = src[(size_t)(i0 + (m & kMapParent)) * w4 + c];
    }
#pragma unroll
    for (int u = 0; u < kApplyUnroll; u++)
      if (dj[u] >= 0) store_out4(dst + dj[u], v[u]);
  }
}

__global__ __launch_bounds__(256) void k_apply(ApplyIn A) {
  __shared__ unsigned short s_map[4 * kApplyTris];  // output row -> parent | child | split (kMap*)
  __shared__ long long s_vbase[kApplyTris];        // first float of the parent's three points
  const int t = threadIdx.x;
  const int i0 = blockIdx.x * kApplyTris;
  const int i1 = min(A.P, i0 + kApplyTris);
  const int n_in = i1 - i0;
  const long long o0 = A.offset[i0];
  long long rows = (long long)A.offset[i1] - o0;
  rows = o0 < 0 ? 0 : (rows < 0 ? 0 : (rows > 4 * kApplyTris ? 4 * kApplyTris : rows));
  for (int r = t; r < 4 * kApplyTris; r += 256) s_map[r] = kNoRow;
  __syncthreads();
  for (int q = t; q < n_in; q += 256) {
    const int i = i0 + q;
    const long long a = (long long)A.offset[i] - o0;
    const int n = A.offset[i + 1] - A.offset[i];
    const int split = n == 4 ? 1 : 0;
    for (int c = 0; c < n && c < 4; c++)
      if (a + c >= 0 && a + c < rows) s_map[a + c] = (unsigned short)(q | c << kMapChildShift | (split ? kMapSplit : 0));
    if (A.cumsum) s_vbase[q] = 3ll * A.cumsum[i];
  }
  __syncthreads();
  const long long P_out = A.P_out;
  for (int k = 0; k < A.n_arrays; k++) {
    const tt_densify_array &d = A.arr[k];
    if (A.vec4 >> k & 1)
      copy_rows4(d, (unsigned)d.width / 4, (unsigned)rows, s_map, i0, o0, P_out);
    else
      copy_rows(d, (unsigned)d.width, (unsigned)rows, s_map, s_vbase, i0, o0, P_out);
  }
  if (A.npts_out || A.cumsum_out) {
    for (int r = t; r < rows; r += 256) {
      const long long j = o0 + r;
      if (s_map[r] == kNoRow || j >= P_out) continue;
      if (A.npts_out) A.npts_out[j] = 3;
      if (A.cumsum_out) A.cumsum_out[j] = (int)(3 * j);
    }
  }
}

bool aligned16(const void *q) { return (reinterpret_cast<uintptr_t>(q) & 15u) == 0; }

int fail(const std::string &m) {
  tt::set_last_error(m);
  return TT_ERR_INVALID_ARG;
}

int launch_status(const char *what) {
  const hipError_t e = hipGetLastError();
  if (e == hipSuccess) return TT_OK;
  tt::set_last_error(std::string(what) + " launch failed: " + hipGetErrorString(e));
  return TT_ERR_HIP;
}

}  // namespace

extern "C" int tt_densify_plan(const tt_densify_plan_args *a, void *stream) {
  if (!a) return fail("tt_densify_plan: args is NULL");
  if (a->P < 0) return fail("tt_densify_plan: P must be non-negative");
  if (!a->offset_dev || !a->work_dev || (a->P > 0 && !a->action_dev))
    return fail("tt_densify_plan: action, offset and work buffers are required");
  if (a->P > 0 && !a->action_in_dev && (!a->max_blending_dev || !a->opacity_dev))
    return fail("tt_densify_plan: max_blending and opacity are required (or a caller action array)");
  hipStream_t s = static_cast<hipStream_t>(stream);
  PlanIn q{a->max_blending_dev, a->opacity_dev, a->density_dev, a->scaling_dev, a->score_dev, a->action_in_dev,
           a->min_blending, a->min_opacity, a->split_scaling, a->split_score, a->P};
  const int nt = (a->P + kPlanTile - 1) / kPlanTile;
  const long long cap = a->max_out <= 0 || a->max_out > INT_MAX ? (long long)INT_MAX : a->max_out;
  if (nt > 0) {
    hipLaunchKernelGGL(k_plan_count, dim3(nt), dim3(kPlanThreads), 0, s, q, a->work_dev, nt);
    if (int rc = launch_status("densify count")) return rc;
  }
  hipLaunchKernelGGL(k_plan_scan, dim3(1), dim3(1024), 0, s, a->work_dev, nt, cap, a->offset_dev, a->P);
  if (int rc = launch_status("densify scan")) return rc;
  if (nt > 0) {
    hipLaunchKernelGGL(k_plan_write, dim3(nt), dim3(kPlanThreads), 0, s, q, a->work_dev, nt, a->action_dev,
                       a->offset_dev);
    if (int rc = launch_status("densify write")) return rc;
  }
  return TT_OK;
}

extern "C" int tt_densify_apply(const tt_densify_apply_args *a, void *stream) {
  if (!a) return fail("tt_densify_apply: args is NULL");
  if (a->P < 0 || a->P_out < 0) return fail("tt_densify_apply: P and P_out must be non-negative");
  if (a->P_out > 4ll * a->P) return fail("tt_densify_apply: P_out exceeds 4 P");
  if (a->n_arrays < 0 || a->n_arrays > TT_DENSIFY_MAX_ARRAYS)
    return fail("tt_densify_apply: n_arrays must be in [0, " + std::to_string(TT_DENSIFY_MAX_ARRAYS) + "]");
  if (a->P > 0 && !a->offset_dev) return fail("tt_densify_apply: offset is NULL");
  if (a->cumsum_out_dev && 3ll * a->P_out > INT_MAX) return fail("tt_densify_apply: 3 P_out overflows cumsum");
  ApplyIn A{};
  for (int k = 0; k < a->n_arrays; k++) {
    const tt_densify_array &d = a->arrays[k];
    const std::string at = "tt_densify_apply: array " + std::to_string(k) + ": ";
    if (d.width < 1) return fail(at + "width must be positive");
    if (d.rows != TT_ROWS_TRIANGLE && d.rows != TT_ROWS_VERTEX) return fail(at + "unknown row layout");
    if (d.on_split != TT_SPLIT_COPY && d.on_split != TT_SPLIT_ZERO && d.on_split != TT_SPLIT_SUBDIVIDE)
      return fail(at + "unknown split policy");
    if (d.rows == TT_ROWS_VERTEX && (d.width != 9 || !a->cumsum_dev))
      return fail(at + "vertex rows are 9 floats and need cumsum");
    if (d.on_split == TT_SPLIT_SUBDIVIDE && d.rows != TT_ROWS_VERTEX)
      return fail(at + "subdivision applies to vertex rows only");
    if ((long long)4 * kApplyTris * d.width > UINT_MAX) return fail(at + "width too large");
    if (a->P > 0 && a->P_out > 0 && (!d.src_dev || !d.dst_dev)) return fail(at + "NULL buffer");
    A.arr[k] = d;
    if (d.rows == TT_ROWS_TRIANGLE && d.width % 4 == 0 && aligned16(d.src_dev) && aligned16(d.dst_dev))
      A.vec4 |= 1u << k;
  }
  if (a->P == 0 || a->P_out == 0) return TT_OK;
  A.n_arrays = a->n_arrays;
  A.P = a->P;
  A.P_out = a->P_out;
  A.offset = a->offset_dev;
  A.cumsum = a->cumsum_dev;
  A.npts_out = a->npts_out_dev;
  A.cumsum_out = a->cumsum_out_dev;
  const int blocks = (a->P + kApplyTris - 1) / kApplyTris;
  hipLaunchKernelGGL(k_apply, dim3(blocks), dim3(256), 0, static_cast<hipStream_t>(stream), A);
  return launch_status("densify apply");
}
