// LBVH builder and refit for gfx950.
//
// Replaces optixAccelBuild (reference tritracer/pipeline.cpp:206-265), which is
// closed driver code. Pipeline (Karras, "Maximizing Parallelism in the
// Construction of BVHs, Octrees, and k-d Trees", HPG 2012):
//   1. per-triangle AABB centroids + scene centroid bounds (wave/LDS reduction,
//      ordered-uint atomics)
//   2. 30-bit Morton codes over the centroid bounds
//   3. LSD radix sort, 4 x 8-bit passes: LDS digit histograms, one-block
//      exclusive scan, stable LDS-ranked scatter (ballot peer masks, wave64)
//   4. Karras hierarchy (index tie-break for equal codes), parent links
//   5. node heights bottom-up (one 64-bit atomic OR per node, no fences), nodes
//      bucketed by height, then child boxes level by level, one launch per height
//      (the kernel boundary orders the levels; per-XCD L2s need no fences).
//      Refit = re-gather leaves + the level launches (topology and levels kept).
// Child boxes live in the parent node, so traversal reads one 64 B node to
// test both children.
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "tt_internal.h"
#include "tt_sort.h"

namespace tt {
namespace {

__device__ __forceinline__ uint32_t f2ord(float f) {
  uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

__global__ void k_init_bounds(uint32_t *bounds) {
  int i = threadIdx.x;
  if (i < 3) bounds[i] = 0xffffffffu;        // min
  else if (i < 6) bounds[i] = 0u;            // max
}

__global__ __launch_bounds__(256) void k_centroids(const float *__restrict__ verts, const int *__restrict__ cumsum,
                                                   int P, float4 *__restrict__ cen, uint32_t *bounds) {
  __shared__ float s_lo[3][4], s_hi[3][4];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  float lo[3] = {__builtin_huge_valf(), __builtin_huge_valf(), __builtin_huge_valf()};
  float hi[3] = {-__builtin_huge_valf(), -__builtin_huge_valf(), -__builtin_huge_valf()};
  if (i < P) {
    const float *v = verts + 3 * (size_t)cumsum[i];
    float c[3];
    for (int a = 0; a < 3; a++) {
      float mn = fminf(v[a], fminf(v[3 + a], v[6 + a]));
      float mx = fmaxf(v[a], fmaxf(v[3 + a], v[6 + a]));
      c[a] = 0.5f * (mn + mx);
      lo[a] = c[a];
      hi[a] = c[a];
    }
    cen[i] = make_float4(c[0], c[1], c[2], 0.0f);
  }
  for (int a = 0; a < 3; a++) {
    for (int off = 32; off > 0; off >>= 1) {
      lo[a] = fminf(lo[a], __shfl_xor(lo[a], off));
      hi[a] = fmaxf(hi[a], __shfl_xor(hi[a], off));
    }
  }
  int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0)
    for (int a = 0; a < 3; a++) s_lo[a][wave] = lo[a], s_hi[a][wave] = hi[a];
  __syncthreads();
  if (threadIdx.x < 3) {
    int a = threadIdx.x;
    float l = s_lo[a][0], h = s_hi[a][0];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) l = fminf(l, s_lo[a][w]), h = fmaxf(h, s_hi[a][w]);
    if (l <= h) {
      atomicMin(bounds + a, f2ord(l));
      atomicMax(bounds + 3 + a, f2ord(h));
    }
  }
}

__device__ __forceinline__ uint32_t expand10(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

__global__ __launch_bounds__(256) void k_morton(const float4 *__restrict__ cen, int P, const uint32_t *bounds,
                                                uint32_t *keys, uint32_t *vals) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  float4 c = cen[i];
  float cc[3] = {c.x, c.y, c.z};
  uint32_t q[3];
  for (int a = 0; a < 3; a++) {
    float lo = ord2f(bounds[a]), hi = ord2f(bounds[3 + a]);
    float ext = hi - lo;
    float x = ext > 0.0f ? (cc[a] - lo) / ext : 0.5f;
    x = fminf(fmaxf(x * 1024.0f, 0.0f), 1023.0f);
    q[a] = (uint32_t)x;
  }
  keys[i] = (expand10(q[0]) << 2) | (expand10(q[1]) << 1) | expand10(q[2]);
  vals[i] = (uint32_t)i;
}

// ---------------------------------------------------------------- hierarchy
__device__ __forceinline__ int delta(const uint32_t *keys, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  uint32_t a = keys[i], b = keys[j];
  if (a == b) return 32 + __clz((int)((uint32_t)i ^ (uint32_t)j));
  return __clz((int)(a ^ b));
}

__global__ __launch_bounds__(256) void k_karras(const uint32_t *__restrict__ keys, int n, int4 *links,
                                                int *leaf_parent) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
  int dmin = delta(keys, n, i, i - d);
  int lmax = 2;
  while (delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
  int l = 0;
  for (int t = lmax / 2; t >= 1; t /= 2)
    if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
  int j = i + l * d;
  int dnode = delta(keys, n, i, j);
  int s = 0;
  int t = l;
  do {
    t = (t + 1) / 2;
    if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
  } while (t > 1);
  int gamma = i + s * d + min(d, 0);
  int left = (min(i, j) == gamma) ? ~gamma : gamma;
  int right = (max(i, j) == gamma + 1) ? ~(gamma + 1) : gamma + 1;
  int *li = reinterpret_cast<int *>(links + 4 * (size_t)i + 3);
  li[0] = left;
  li[1] = right;
  li[3] = max(i, j) - min(i, j) + 1;  // leaves under the node
  if (i == 0) li[2] = -1;
  if (left < 0)
    leaf_parent[~left] = i;
  else
    reinterpret_cast<int *>(links + 4 * (size_t)left + 3)[2] = i;
  if (right < 0)
    leaf_parent[~right] = i;
  else
    reinterpret_cast<int *>(links + 4 * (size_t)right + 3)[2] = i;
}

// Two-leaf tree for P == 1: leaf 0 is the triangle, leaf 1 an unhittable dummy.
__global__ void k_single(int4 *links, int *leaf_parent, uint32_t *vals) {
  int *li = reinterpret_cast<int *>(links + 3);
  li[0] = ~0;
  li[1] = ~1;
  li[2] = -1;
  li[3] = 2;
  leaf_parent[0] = 0;
  leaf_parent[1] = 0;
  vals[0] = 0u;
  vals[1] = 0xFFFFFFFFu;
}

__global__ void k_zero_u64(unsigned long long *p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}
__global__ void k_zero_u32(uint32_t *p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// Leaves (re)gathered from the vertex buffer in Morton order.
__global__ __launch_bounds__(256) void k_leaves(const float *__restrict__ verts, const int *__restrict__ cumsum,
                                                const uint32_t *__restrict__ order, int n_leaves, float4 *leaves) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_leaves) return;
  uint32_t prim = order[s];
  float A[3], B[3], C[3];
  if (prim != 0xFFFFFFFFu) {
    const float *v = verts + 3 * (size_t)cumsum[prim];
    for (int a = 0; a < 3; a++) A[a] = v[a], B[a] = v[3 + a], C[a] = v[6 + a];
  } else {
    for (int a = 0; a < 3; a++) A[a] = B[a] = C[a] = 1e30f;
  }
  leaves[3 * (size_t)s + 0] = make_float4(A[0], A[1], A[2], __uint_as_float(prim));
  leaves[3 * (size_t)s + 1] = make_float4(B[0], B[1], B[2], 0.0f);
  leaves[3 * (size_t)s + 2] = make_float4(C[0], C[1], C[2], 0.0f);
}

// Node heights (leaf 0, internal 1 + max over children), bottom-up. Both children
// report into one 64-bit word of the parent (one 32-bit half each) with a single
// atomic OR: the second arrival sees the sibling's value in the returned word, so
// no other memory has to be made visible across XCDs and no fence is needed.
__global__ __launch_bounds__(256) void k_heights(const float4 *__restrict__ nodes, const int *__restrict__ leaf_parent,
                                                 int n_leaves, unsigned long long *hword, int *height,
                                                 uint32_t *maxh) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_leaves) return;
  int child = ~s;
  int p = leaf_parent[s];
  uint32_t h = 1;
  while (true) {
    const int *li = reinterpret_cast<const int *>(nodes + 4 * (size_t)p + 3);
    const int slot = (li[0] == child) ? 0 : 1;
    const unsigned long long old = atomicOr(hword + p, (unsigned long long)h << (32 * slot));
    const uint32_t other = (uint32_t)(old >> (32 * (1 - slot)));
    if (other == 0u) return;
    const uint32_t hp = max(h, other);
    height[p] = (int)hp;
    const int parent = li[2];
    if (parent < 0) {
      atomicMax(maxh, hp);
      return;
    }
    child = p;
    p = parent;
    h = hp + 1;
  }
}

// Height histogram: per-block LDS counts, one global add per (block, height).
__global__ __launch_bounds__(256) void k_level_hist(const int *__restrict__ height, int ni, uint32_t *hist) {
  __shared__ uint32_t h[kMaxLevels];
  for (int i = threadIdx.x; i < kMaxLevels; i += blockDim.x) h[i] = 0u;
  __syncthreads();
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n < ni) atomicAdd(&h[height[n]], 1u);
  __syncthreads();
  for (int i = threadIdx.x; i < kMaxLevels; i += blockDim.x)
    if (h[i]) atomicAdd(hist + i, h[i]);
}

// Nodes into height order: each block reserves one range per height with a single
// global atomic, then places its nodes by their LDS rank inside that range.
__global__ __launch_bounds__(256) void k_level_scatter(const int *__restrict__ height, int ni,
                                                       const uint32_t *__restrict__ offs, uint32_t *cursor,
                                                       int *level_nodes) {
  __shared__ uint32_t cnt[kMaxLevels];
  __shared__ uint32_t base[kMaxLevels];
  for (int i = threadIdx.x; i < kMaxLevels; i += blockDim.x) cnt[i] = 0u;
  __syncthreads();
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  int h = 0;
  uint32_t rank = 0;
  if (n < ni) {
    h = height[n];
    rank = atomicAdd(&cnt[h], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kMaxLevels; i += blockDim.x)
    if (cnt[i]) base[i] = offs[i] + atomicAdd(cursor + i, cnt[i]);
  __syncthreads();
  if (n < ni) level_nodes[base[h] + rank] = n;
}

// One height level: write both child boxes of every node of the level into the
// node. Children are leaves (box from the leaf triangle, padded) or nodes of lower
// height, completed by an earlier launch.
__global__ __launch_bounds__(256) void k_level_boxes(const int *__restrict__ level_nodes, int count,
                                                     const float4 *__restrict__ leaves, float4 *nodes) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int n = level_nodes[i];
  float4 *nd = nodes + 4 * (size_t)n;
  const int4 li = *reinterpret_cast<const int4 *>(nd + 3);
  float lo[2][3], hi[2][3];
  for (int slot = 0; slot < 2; slot++) {
    const int c = slot ? li.y : li.x;
    if (c < 0) {
      const int s = ~c;
      const float4 A = leaves[3 * (size_t)s], B = leaves[3 * (size_t)s + 1], C = leaves[3 * (size_t)s + 2];
      const float av[3] = {A.x, A.y, A.z}, bv[3] = {B.x, B.y, B.z}, cv[3] = {C.x, C.y, C.z};
      for (int a = 0; a < 3; a++) {
        float l = fminf(av[a], fminf(bv[a], cv[a])), h = fmaxf(av[a], fmaxf(bv[a], cv[a]));
        const float pad = 1e-5f * (1.0f + fmaxf(fabsf(l), fabsf(h)));
        lo[slot][a] = l - pad;
        hi[slot][a] = h + pad;
      }
    } else {
      const float4 f0 = nodes[4 * (size_t)c], f1 = nodes[4 * (size_t)c + 1], f2 = nodes[4 * (size_t)c + 2];
      lo[slot][0] = fminf(f0.x, f1.x);
      hi[slot][0] = fmaxf(f0.y, f1.y);
      lo[slot][1] = fminf(f0.z, f1.z);
      hi[slot][1] = fmaxf(f0.w, f1.w);
      lo[slot][2] = fminf(f2.x, f2.z);
      hi[slot][2] = fmaxf(f2.y, f2.w);
    }
  }
  nd[0] = make_float4(lo[0][0], hi[0][0], lo[0][1], hi[0][1]);
  nd[1] = make_float4(lo[1][0], hi[1][0], lo[1][1], hi[1][1]);
  nd[2] = make_float4(lo[0][2], hi[0][2], lo[1][2], hi[1][2]);
}

// 4-wide nodes: node4[n] holds up to four disjoint subtrees that cover n's leaves:
// starting from n's two children, the internal entry with the largest surface area is
// replaced by its two children, twice (the greedy rule for collapsing a binary tree into
// a wide one: it opens the boxes a ray is most likely to hit). A 4-wide step descends at
// least one binary level, so a walk's path holds at most `height` 4-wide nodes.
//
// Stored quantized, 80 B per node (5 x 16 B; 128 B as floats):
//   [0] float origin.x, origin.y, origin.z; exponents e_x, e_y, e_z (bytes, biased by 128)
//   [1] x planes: step counts q in 0..255, each an f16 integer (exact): .x = lo of children
//       0, 1 (low half, high half), .y = lo of children 2, 3, .z / .w = their hi
//   [2] y planes, [3] z planes, laid out as [1]
//   [4] int child[4] (>= 0 binary node, < 0 leaf ~s, kEmptyChild unused)
// and, in counts4, the leaves under each child (16 bits, saturated): .x children 0, 1, .y 2, 3.
// f16 step counts let the walk turn a plane into a float inside its fma (v_fma_mix_f32).
// Child c's box decodes per axis as fmaf(q, 2^e, origin) (q * 2^e is exact: one rounding).
// origin is the least lo of the node's children, and every q is chosen here with that same
// decode so that the decoded box contains the child's box (lo rounded down, hi up, the
// scale grown when 255 steps do not reach the top after rounding): a walk on the decoded
// boxes culls no less conservatively than on the exact ones, so hit lists do not change.
__device__ __forceinline__ float q_scale(int e) {
  return __int_as_float((e + 127) << 23);
}
__global__ __launch_bounds__(256) void k_collapse4(const float4 *__restrict__ nodes, int ni, float4 *nodes4,
                                                    uint2 *counts4) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= ni) return;
  float lo[3][4], hi[3][4];
  int ref[4], cnt[4] = {0, 0, 0, 0};  // child, leaves under it
  int e = 0;
  auto slot_box = [&](int node, int slot, float l[3], float h[3]) {
    const float4 a = nodes[4 * (size_t)node + slot], z = nodes[4 * (size_t)node + 2];
    l[0] = a.x, h[0] = a.y, l[1] = a.z, h[1] = a.w;
    l[2] = slot ? z.z : z.x;
    h[2] = slot ? z.w : z.y;
  };
  auto set = [&](int i, int r, const float l[3], const float h[3]) {
    for (int a = 0; a < 3; a++) lo[a][i] = l[a], hi[a][i] = h[a];
    ref[i] = r;
    cnt[i] = r < 0 ? 1 : reinterpret_cast<const int4 *>(nodes + 4 * (size_t)r + 3)->w;
  };
  const int4 li = *reinterpret_cast<const int4 *>(nodes + 4 * (size_t)n + 3);
  float l[3], h[3];
  slot_box(n, 0, l, h);
  set(e++, li.x, l, h);
  slot_box(n, 1, l, h);
  set(e++, li.y, l, h);
  for (int open = 0; open < 2; open++) {
    int best = -1;
    float best_area = -1.0f;
    for (int i = 0; i < e; i++) {
      if (ref[i] < 0) continue;
      const float dx = hi[0][i] - lo[0][i], dy = hi[1][i] - lo[1][i], dz = hi[2][i] - lo[2][i];
      const float area = dx * dy + dy * dz + dz * dx;
      if (area > best_area) best_area = area, best = i;
    }
    if (best < 0) break;
    const int c = ref[best];
    const int4 lc = *reinterpret_cast<const int4 *>(nodes + 4 * (size_t)c + 3);
    slot_box(c, 0, l, h);
    set(best, lc.x, l, h);
    slot_box(c, 1, l, h);
    set(e++, lc.y, l, h);
  }
  for (; e < 4; e++) {
    for (int a = 0; a < 3; a++) lo[a][e] = 1.0f, hi[a][e] = 0.0f;
    ref[e] = kEmptyChild;
    cnt[e] = 0;
  }
  uint32_t qlo[3][4], qhi[3][4], eb = 0;
  float org[3];
  for (int a = 0; a < 3; a++) {
    float o = __builtin_huge_valf(), top = -__builtin_huge_valf();
    for (int c = 0; c < 4; c++)
      if (ref[c] != kEmptyChild) o = fminf(o, lo[a][c]), top = fmaxf(top, hi[a][c]);
    const double ext = (double)top - (double)o;
    int sc_e = -120;
    if (ext > 0.0) sc_e = max(-120, min(120, (int)ceil(log2(ext / 254.0))));
    uint32_t wl[4], wh[4];
    for (; sc_e <= 120; sc_e++) {  // the smallest scale whose steps reach every child's top after rounding
      const float sc = q_scale(sc_e);
      bool ok = true;
      for (int c = 0; c < 4 && ok; c++) {
        uint32_t ql = 255, qh = 0;
        if (ref[c] != kEmptyChild) {
          int l = (int)floor(((double)lo[a][c] - (double)o) / (double)sc);
          int h = (int)ceil(((double)hi[a][c] - (double)o) / (double)sc);
          l = max(0, min(255, l));
          h = max(0, min(255, h));
          while (l > 0 && fmaf((float)l, sc, o) > lo[a][c]) l--;
          while (h < 255 && fmaf((float)h, sc, o) < hi[a][c]) h++;
          if (fmaf((float)h, sc, o) < hi[a][c]) ok = false;
          ql = (uint32_t)l;
          qh = (uint32_t)h;
        }
        wl[c] = ql;
        wh[c] = qh;
      }
      if (ok) break;
    }
    org[a] = o;
    for (int c = 0; c < 4; c++) qlo[a][c] = wl[c], qhi[a][c] = wh[c];
    eb |= (uint32_t)(sc_e + 128) << (8 * a);
  }
  auto c16 = [&](int i) { return (uint32_t)min(cnt[i], 0xffff); };
  auto h2 = [](uint32_t a, uint32_t b) {  // two step counts as f16 integers (low half, high half)
    return __uint_as_float((uint32_t)__builtin_bit_cast(uint16_t, (_Float16)a) |
                           (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)b) << 16);
  };
  float4 *out = nodes4 + 5 * (size_t)n;
  out[0] = make_float4(org[0], org[1], org[2], __uint_as_float(eb));
  for (int a = 0; a < 3; a++)
    out[1 + a] = make_float4(h2(qlo[a][0], qlo[a][1]), h2(qlo[a][2], qlo[a][3]), h2(qhi[a][0], qhi[a][1]),
                             h2(qhi[a][2], qhi[a][3]));
  *reinterpret_cast<int4 *>(out + 4) = make_int4(ref[0], ref[1], ref[2], ref[3]);
  counts4[n] = make_uint2(c16(0) | c16(1) << 16, c16(2) | c16(3) << 16);
}

template <class T>
hipError_t grow(T *&ptr, size_t &cap, size_t need) {
  if (cap >= need && ptr) return hipSuccess;
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  cap = 0;
  hipError_t e = hipMalloc(reinterpret_cast<void **>(&ptr), need * sizeof(T));
  if (e == hipSuccess) cap = need;
  return e;
}

// Surface area of a box (in double: the sums below add ~2P terms).
__device__ __forceinline__ double box_area(float lx, float hx, float ly, float hy, float lz, float hz) {
  const double dx = fmax(0.0, (double)hx - lx), dy = fmax(0.0, (double)hy - ly), dz = fmax(0.0, (double)hz - lz);
  return 2.0 * ((dx * dy + dy * dz) + dz * dx);
}

// Per block: the areas of the children of its 256 internal nodes (every non-root node is
// the child of exactly one internal node), summed in a fixed order.
__global__ __launch_bounds__(256) void k_sah_part(const float4 *__restrict__ nodes, int ni, double *part) {
  __shared__ double s[256];
  const int i = blockIdx.x * 256 + threadIdx.x;
  double v = 0.0;
  if (i < ni) {
    const float4 a = nodes[4 * (size_t)i], b = nodes[4 * (size_t)i + 1], z = nodes[4 * (size_t)i + 2];
    v = box_area(a.x, a.y, a.z, a.w, z.x, z.y) + box_area(b.x, b.y, b.z, b.w, z.z, z.w);
  }
  s[threadIdx.x] = v;
  __syncthreads();
  for (int d = 128; d > 0; d >>= 1) {
    if (threadIdx.x < d) s[threadIdx.x] += s[threadIdx.x + d];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// 1 + (sum of the partial sums) / SA(root), the root box being the union of node 0's children.
__global__ __launch_bounds__(1024) void k_sah_total(const double *part, int nb, const float4 *nodes, float *out) {
  __shared__ double s[1024];
  const int t = threadIdx.x;
  double v = 0.0;
  for (int i = t; i < nb; i += 1024) v += part[i];
  s[t] = v;
  __syncthreads();
  for (int d = 512; d > 0; d >>= 1) {
    if (t < d) s[t] += s[t + d];
    __syncthreads();
  }
  if (t == 0) {
    const float4 a = nodes[0], b = nodes[1], z = nodes[2];
    const double root = box_area(fminf(a.x, b.x), fmaxf(a.y, b.y), fminf(a.z, b.z), fmaxf(a.w, b.w),
                                 fminf(z.x, z.z), fmaxf(z.y, z.w));
    out[0] = root > 0.0 ? (float)(1.0 + s[0] / root) : 1.0f;
  }
}

}  // namespace

std::string bvh_cost(Bvh &b, float *out, hipStream_t stream) {
  if (b.P <= 0 || b.n_internal <= 0) {
    const float one = 1.0f;
    hipError_t e = hipMemcpyAsync(out, &one, sizeof(float), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);  // `one` must outlive the copy
    return e == hipSuccess ? "" : std::string("SAH cost copy failed: ") + hipGetErrorString(e);
  }
  const int nb = (b.n_internal + 255) / 256;
  hipError_t e = grow(b.sah_part, b.cap_sah, (size_t)nb);
  if (e != hipSuccess) return std::string("hipMalloc of the SAH partial sums failed: ") + hipGetErrorString(e);
  hipLaunchKernelGGL(k_sah_part, dim3(nb), dim3(256), 0, stream, b.nodes, b.n_internal, b.sah_part);
  hipLaunchKernelGGL(k_sah_total, dim3(1), dim3(1024), 0, stream, b.sah_part, nb, b.nodes, out);
  e = hipGetLastError();
  return e == hipSuccess ? "" : std::string("SAH cost launch failed: ") + hipGetErrorString(e);
}

void bvh_free(Bvh &b) {
  void *ps[] = {b.nodes, b.nodes4, b.counts4, b.leaves, b.leaf_parent, b.keys[0], b.keys[1], b.vals[0], b.vals[1], b.hist,
                b.bounds, b.centroids, b.hword, b.height, b.level_nodes, b.level_meta, b.sah_part};
  for (void *p : ps)
    if (p) (void)hipFree(p);
  b = Bvh();
}

#define TT_TRY(x)                                                                          \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess) return std::string(#x " failed: ") + hipGetErrorString(e_);     \
  } while (0)

static std::string refit_levels(Bvh &b, const float *verts, const int *cumsum, hipStream_t stream) {
  const int nl = b.n_leaves;
  hipLaunchKernelGGL(k_leaves, dim3((nl + 255) / 256), dim3(256), 0, stream, verts, cumsum, b.vals[0], nl, b.leaves);
  for (int h = 1; h <= b.height_max; h++) {
    const int off = b.level_off[h], cnt = b.level_off[h + 1] - b.level_off[h];
    if (cnt == 0) continue;
    hipLaunchKernelGGL(k_level_boxes, dim3((cnt + 255) / 256), dim3(256), 0, stream, b.level_nodes + off, cnt,
                       b.leaves, b.nodes);
  }
  hipLaunchKernelGGL(k_collapse4, dim3((b.n_internal + 255) / 256), dim3(256), 0, stream, b.nodes, b.n_internal,
                     b.nodes4, b.counts4);
  TT_TRY(hipGetLastError());
  return "";
}

std::string bvh_build(Bvh &b, const float *verts, int num_vertices, const int *cumsum, int P, int mode,
                      hipStream_t stream) {
  (void)num_vertices;
  if (mode == 1) {
    if (b.P < 0) return "refit requested but no acceleration structure has been built";
    if (b.P != P) return "refit requires the same number of triangles as the last build";
    if (P == 0) return "";
    return refit_levels(b, verts, cumsum, stream);
  }
  if (P == 0) {
    b.P = 0;
    b.n_leaves = 0;
    b.n_internal = 0;
    b.height_max = 0;
    return "";
  }
  const int nl = P == 1 ? 2 : P;
  const int ni = nl - 1;
  const size_t need = (size_t)nl;
  if (b.cap_leaves < need || !b.nodes) {
    void *ps[] = {b.nodes, b.nodes4, b.counts4, b.leaves, b.leaf_parent, b.keys[0], b.keys[1], b.vals[0], b.vals[1],
                  b.centroids, b.hword, b.height, b.level_nodes};
    for (void *p : ps)
      if (p) (void)hipFree(p);
    b.nodes = nullptr, b.nodes4 = nullptr, b.counts4 = nullptr, b.leaves = nullptr, b.leaf_parent = nullptr;
    b.keys[0] = b.keys[1] = b.vals[0] = b.vals[1] = nullptr;
    b.centroids = nullptr, b.hword = nullptr, b.height = nullptr, b.level_nodes = nullptr;
    b.cap_leaves = 0;
    TT_TRY(hipMalloc(&b.nodes, sizeof(float4) * 4 * need));
    TT_TRY(hipMalloc(&b.nodes4, sizeof(float4) * 5 * need));
    TT_TRY(hipMalloc(&b.counts4, sizeof(uint2) * need));
    TT_TRY(hipMalloc(&b.leaves, sizeof(float4) * 3 * need));
    TT_TRY(hipMalloc(&b.leaf_parent, sizeof(int) * need));
    TT_TRY(hipMalloc(&b.keys[0], sizeof(uint32_t) * need));
    TT_TRY(hipMalloc(&b.keys[1], sizeof(uint32_t) * need));
    TT_TRY(hipMalloc(&b.vals[0], sizeof(uint32_t) * need));
    TT_TRY(hipMalloc(&b.vals[1], sizeof(uint32_t) * need));
    TT_TRY(hipMalloc(&b.centroids, sizeof(float4) * need));
    TT_TRY(hipMalloc(&b.hword, sizeof(unsigned long long) * need));
    TT_TRY(hipMalloc(&b.height, sizeof(int) * need));
    TT_TRY(hipMalloc(&b.level_nodes, sizeof(int) * need));
    b.cap_leaves = need;
  }
  if (!b.bounds) TT_TRY(hipMalloc(&b.bounds, sizeof(uint32_t) * 8));
  if (!b.level_meta) TT_TRY(hipMalloc(&b.level_meta, sizeof(uint32_t) * 3 * kMaxLevels));
  uint32_t *hist = b.level_meta, *offs = b.level_meta + kMaxLevels, *cursor = b.level_meta + 2 * kMaxLevels;
  if (P == 1) {
    hipLaunchKernelGGL(k_single, dim3(1), dim3(1), 0, stream, reinterpret_cast<int4 *>(b.nodes), b.leaf_parent,
                       b.vals[0]);
  } else {
    const int nblk = (P + 255) / 256;
    hipLaunchKernelGGL(k_init_bounds, dim3(1), dim3(64), 0, stream, b.bounds);
    hipLaunchKernelGGL(k_centroids, dim3(nblk), dim3(256), 0, stream, verts, cumsum, P, b.centroids, b.bounds);
    hipLaunchKernelGGL(k_morton, dim3(nblk), dim3(256), 0, stream, b.centroids, P, b.bounds, b.keys[0], b.vals[0]);
    const int sblk = (P + kSortTile - 1) / kSortTile;
    TT_TRY(grow(b.hist, b.cap_hist, (size_t)256 * sblk));
    int cur = 0;
    for (int shift = 0; shift < 32; shift += 8) {
      hipLaunchKernelGGL(k_hist, dim3(sblk), dim3(kSortThreads), 0, stream, b.keys[cur], P, nullptr, shift, b.hist,
                         sblk);
      hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, stream, b.hist, 256 * sblk);
      hipLaunchKernelGGL(k_scatter, dim3(sblk), dim3(kSortThreads), 0, stream, b.keys[cur], b.vals[cur],
                         b.keys[cur ^ 1], b.vals[cur ^ 1], P, nullptr, shift, b.hist, sblk);
      cur ^= 1;
    }
    // four passes: sorted data is back in buffer 0
    hipLaunchKernelGGL(k_karras, dim3((ni + 255) / 256), dim3(256), 0, stream, b.keys[0], P,
                       reinterpret_cast<int4 *>(b.nodes), b.leaf_parent);
  }
  // heights and levels (the level sizes come back to the host once per build)
  hipLaunchKernelGGL(k_zero_u64, dim3((ni + 255) / 256), dim3(256), 0, stream, b.hword, ni);
  hipLaunchKernelGGL(k_zero_u32, dim3((3 * kMaxLevels + 255) / 256), dim3(256), 0, stream, b.level_meta,
                     3 * kMaxLevels);
  hipLaunchKernelGGL(k_heights, dim3((nl + 255) / 256), dim3(256), 0, stream, b.nodes, b.leaf_parent, nl, b.hword,
                     b.height, cursor + kMaxLevels - 1);
  hipLaunchKernelGGL(k_level_hist, dim3((ni + 255) / 256), dim3(256), 0, stream, b.height, ni, hist);
  std::vector<uint32_t> h_hist(kMaxLevels);
  TT_TRY(hipMemcpyAsync(h_hist.data(), hist, sizeof(uint32_t) * kMaxLevels, hipMemcpyDeviceToHost, stream));
  uint32_t hmax = 0;
  TT_TRY(hipMemcpyAsync(&hmax, cursor + kMaxLevels - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
  TT_TRY(hipStreamSynchronize(stream));
  if (hmax == 0 || hmax >= (uint32_t)kMaxLevels - 2) return "LBVH height out of range";
  b.height_max = (int)hmax;
  b.level_off.assign(kMaxLevels + 1, 0);
  for (int h = 0; h < kMaxLevels; h++) b.level_off[h + 1] = b.level_off[h] + (int)h_hist[h];
  if (b.level_off[hmax + 1] != ni) return "LBVH level bucketing lost nodes";
  std::vector<uint32_t> h_offs(b.level_off.begin(), b.level_off.begin() + kMaxLevels);
  TT_TRY(hipMemcpyAsync(offs, h_offs.data(), sizeof(uint32_t) * kMaxLevels, hipMemcpyHostToDevice, stream));
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(256), 0, stream, cursor, kMaxLevels - 1);
  hipLaunchKernelGGL(k_level_scatter, dim3((ni + 255) / 256), dim3(256), 0, stream, b.height, ni, offs, cursor,
                     b.level_nodes);
  TT_TRY(hipStreamSynchronize(stream));  // h_offs must outlive the copy
  b.P = P;
  b.n_leaves = nl;
  b.n_internal = ni;
  b.cumsum_seen = cumsum;
  return refit_levels(b, verts, cumsum, stream);
}

}  // namespace tt
