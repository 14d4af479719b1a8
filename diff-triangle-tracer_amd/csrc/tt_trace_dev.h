// Device side of the traversal kernels: k-closest gather over the LBVH,
// front-to-back compositing, and the re-traversing backward. Included by
// tt_trace.hip (one-shot launches) and tt_multi.hip (persistent multi-camera).
//
// Replaces the OptiX programs of the reference:
//   __raygen__rg forward ........ tritracer/forward.cu:76-215
//   __raygen__rg backward ....... tritracer/backward.cu:191-391
//   reflect / refract / envlight  tritracer/reflect.cu, refract.cu, envlight.cu
//   optixTrace + __anyhit__ah ... tritracer/closestk.h:17-149 (k-buffer insertion)
//
// Mapping: one ray per lane; each wave64 takes an 8x8 pixel tile, a block of
// 128 lanes a 16x8 tile. Blocks are remapped so that the tiles one XCD works
// on form a contiguous band of the image (workgroups are dispatched to the 8
// XCDs round-robin and each XCD has its own L2). Traversal walks the 4-wide
// view of the LBVH (every second binary level: half the dependent node fetches
// of a binary walk for the same box tests). The stack keeps its top 16 entries
// per lane in LDS ([slot][lane] layout, bank-conflict free) and spills deeper
// entries to a per-lane global area sized from the tree height: each visited
// 4-wide node defers at most 3 children and a root-to-leaf path visits at most
// ceil(height/2) of them.
//
// Backward gradients are first summed per block in an LDS table keyed by
// triangle (one slot per lane of the block x 59 floats, LDS float atomics; linear probing, global
// atomics on overflow) and scattered once per block: on the 1080p / 1M-splat
// workload a 128-ray block sees ~11x fewer distinct triangles than blended
// hits, and the global FP32 atomics (which execute past the per-XCD L2) were
// the backward's bound.
#pragma once
#include <type_traits>
#include <hip/hip_runtime.h>

#include "tt_device.h"
#include "tt_internal.h"

namespace tt {
namespace {

constexpr int kStack = kLdsStack;  // LDS entries per lane (deferred node index)
constexpr int kLeafQ = 12;         // LDS queue of postponed leaves per lane
constexpr int kLdsInts = kStack + kLeafQ;  // ints of LDS per lane the gather uses

// Pixel tile of a block: one 8x8 quadrant per wave; 64 lanes -> 8x8, 128 -> 16x8, 256 -> 16x16.
template <int BLOCK>
struct TileShape {
  static constexpr int W = BLOCK == 64 ? 8 : 16;
  static constexpr int H = BLOCK == 256 ? 16 : 8;
};

template <int K>
struct KBuf {
  uint64_t k[K];
  TT_D float t(int i) const { return __uint_as_float((uint32_t)(k[i] >> 32)); }
  TT_D uint32_t p(int i) const { return (uint32_t)k[i]; }
};
TT_D uint64_t hit_key(float t, uint32_t p) { return ((uint64_t)__float_as_uint(t) << 32) | p; }

template <int K>
TT_D void kinit(KBuf<K> &b) {
#pragma unroll
  for (int i = 0; i < K; i++) b.k[i] = hit_key(kInfDist, kInvalid);
}

// Sorted insertion (the anyhit compare-and-swap chain of closestk.h:111-147) with a
// deterministic (t, prim) order instead of traversal order for equal t. A key is the 64-bit
// (t bits, prim): for t >= 0 (every gathered hit: intersect returns -0.0 as +0.0 and the
// gathers start at t >= 0) unsigned key order is the (t, prim) order.
// Slot i takes slot i-1 if the key is less than slot i-1, the key if it is less than slot
// i only, and keeps its value otherwise; slots are rewritten from the last down, each from
// the old contents, so the steps do not depend on one another.
template <int K>
TT_D void kinsert(KBuf<K> &b, float t, uint32_t p) {
  const uint64_t key = hit_key(t, p);
  bool lt = key < b.k[K - 1];  // key < slot i
  if (!lt) return;
#pragma unroll
  for (int i = K - 1; i > 0; i--) {
    const bool lt_prev = key < b.k[i - 1];
    b.k[i] = lt_prev ? b.k[i - 1] : (lt ? key : b.k[i]);
    lt = lt_prev;
  }
  b.k[0] = lt ? key : b.k[0];
}
template <int K>
TT_D void kpop(KBuf<K> &b) {
#pragma unroll
  for (int i = 0; i < K - 1; i++) b.k[i] = b.k[i + 1];
  b.k[K - 1] = hit_key(kInfDist, kInvalid);
}

// Optional traversal counters (debug/profiling): nodes visited and triangles tested.
struct TravCount {
  int nodes = 0, tris = 0;
};

TT_D void cswap(float &da, int &ca, float &db, int &cb) {
  const bool sw = db < da;
  const float td = sw ? db : da;
  const int tc = sw ? cb : ca;
  db = sw ? da : db;
  cb = sw ? ca : cb;
  da = td;
  ca = tc;
}

// One chunk: the K smallest (t, prim) with t <= far_t and (t, prim) after the key.
// Ordered traversal of the 4-wide nodes: the up-to-4 child boxes are tested from
// the node itself, the hit internal children are sorted by entry distance
// (5-comparator network), the nearest is descended and the others are deferred
// farthest-first. Hit leaves are not tested at once: they go to a per-lane queue in
// LDS and a lane tests one queued leaf per step when at least half of the wave's
// active lanes have one queued (or its queue is nearly full, or its walk is over),
// so the intersection code runs with most lanes busy instead of once per leaf-holding
// lane. The K-buffer keeps the K smallest keys whatever the insertion order, so
// postponing a leaf only weakens the culling bound for a while.
// Two f16 integers (low half, high half) of a quantized node's plane word.
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
TT_D half2_t plane_pair(float w) { return __builtin_bit_cast(half2_t, w); }

// A 4-wide node (quantized, layout at k_collapse4) decoded to its child boxes' planes: per
// axis fmaf(q, 2^e, origin), the builder's own decode, so the boxes contain the exact ones.
TT_D void load_node4q(const float4 *q, float4 &lx, float4 &hx, float4 &ly, float4 &hy, float4 &lz, float4 &hz,
                      int4 &ref) {
  const float4 a = q[0];
  ref = *reinterpret_cast<const int4 *>(q + 4);
  const uint32_t eb = __float_as_uint(a.w);
  auto dec4 = [](float w0, float w1, float s, float o) {
    const half2_t u = plane_pair(w0), v = plane_pair(w1);
    return make_float4(fmaf((float)u.x, s, o), fmaf((float)u.y, s, o), fmaf((float)v.x, s, o),
                       fmaf((float)v.y, s, o));
  };
  const float4 px = q[1], py = q[2], pz = q[3];
  const float sx = __uint_as_float(((eb & 0xffu) - 1u) << 23), sy = __uint_as_float((((eb >> 8) & 0xffu) - 1u) << 23),
              sz = __uint_as_float((((eb >> 16) & 0xffu) - 1u) << 23);
  lx = dec4(px.x, px.y, sx, a.x);
  hx = dec4(px.z, px.w, sx, a.x);
  ly = dec4(py.x, py.y, sy, a.y);
  hy = dec4(py.z, py.w, sy, a.y);
  lz = dec4(pz.x, pz.y, sz, a.z);
  hz = dec4(pz.z, pz.w, sz, a.z);
}

// The walk's slab test reads a quantized node in ray space instead of decoding its planes:
// the plane at step q of axis a lies at t = q * (2^e / d_a) + (origin_a - o_a) / d_a, where
// 2^e / d_a is exact, so a plane costs one fma (the decode-then-slab form costs four).
// Against the decoded planes the build verified, such a t errs by one rounding of the
// decode (u |plane|, u = 2^-24), two of (origin - o) / d and one of the pad below, at most
// 4.1 u (C + |o|) / |d_a| together (C bounds every coordinate of the scene), plus one
// rounding of t itself, which the interval widening of box_ok_wide covers. Lower planes
// are moved down and upper planes up by pad = 2^-19 (C + |o|), 32 u, in world space: the
// walk stays conservative, and hit lists do not change.
struct NodePad {
  v3f pinv;                   // pad / |d| per axis
  bool neg_x, neg_y, neg_z;   // d < 0 per axis: the upper plane is the entry plane
};
TT_D NodePad node_pad(const TraceParams &P, const RayPre &r) {
  // the root's origin and 256 steps of its scale bound every coordinate of the scene
  const float4 root = P.nodes4[0];
  const uint32_t eb = __float_as_uint(root.w);
  auto reach = [](float org, uint32_t e) { return fmaf(256.0f, __uint_as_float(((e & 0xffu) - 1u) << 23), fabsf(org)); };
  const float C = fmaxf(fmaxf(reach(root.x, eb), reach(root.y, eb >> 8)), reach(root.z, eb >> 16));
  const float pad = 0x1p-19f * (C + fmaxf(fmaxf(fabsf(r.o.x), fabsf(r.o.y)), fabsf(r.o.z)));
  NodePad np;
  np.pinv = mk(pad * fabsf(r.inv.x), pad * fabsf(r.inv.y), pad * fabsf(r.inv.z));
  np.neg_x = r.inv.x < 0.0f, np.neg_y = r.inv.y < 0.0f, np.neg_z = r.inv.z < 0.0f;
  return np;
}
// Entry and exit distances of the 4 children of a quantized node (unwidened, as slab_raw).
TT_D void slab4q(const RayPre &r, const NodePad &np, const float4 a, const float4 px, const float4 py,
                 const float4 pz, float4 &tn, float4 &tf) {
  const uint32_t eb = __float_as_uint(a.w);
  constexpr float kBig = 3.0e38f;  // 2^e / d overflows only for e near 120 and d near 1e-30
  auto step = [&](uint32_t e, float inv) {  // 2^(e - 128) / d, exact below overflow
    return fminf(fmaxf(__builtin_amdgcn_ldexpf(inv, (int)(e & 0xffu) - 128), -kBig), kBig);
  };
  const float sx = step(eb, r.inv.x), sy = step(eb >> 8, r.inv.y), sz = step(eb >> 16, r.inv.z);
  const float bx = (a.x - r.o.x) * r.inv.x, by = (a.y - r.o.y) * r.inv.y, bz = (a.z - r.o.z) * r.inv.z;
  auto t4 = [](float w0, float w1, float s, float o) {  // f16 step counts converted inside the fma
    const half2_t u = plane_pair(w0), v = plane_pair(w1);
    return make_float4(fmaf((float)u.x, s, o), fmaf((float)u.y, s, o), fmaf((float)v.x, s, o),
                       fmaf((float)v.y, s, o));
  };
  // entry / exit plane of each axis by the sign of d: fma is monotone in q and in the offset,
  // so these are exactly the min / max of the lower and upper plane distances
  const float4 nx = t4(np.neg_x ? px.z : px.x, np.neg_x ? px.w : px.y, sx, bx - np.pinv.x);
  const float4 fx = t4(np.neg_x ? px.x : px.z, np.neg_x ? px.y : px.w, sx, bx + np.pinv.x);
  const float4 ny = t4(np.neg_y ? py.z : py.x, np.neg_y ? py.w : py.y, sy, by - np.pinv.y);
  const float4 fy = t4(np.neg_y ? py.x : py.z, np.neg_y ? py.y : py.w, sy, by + np.pinv.y);
  const float4 nz = t4(np.neg_z ? pz.z : pz.x, np.neg_z ? pz.w : pz.y, sz, bz - np.pinv.z);
  const float4 fz = t4(np.neg_z ? pz.x : pz.z, np.neg_z ? pz.y : pz.w, sz, bz + np.pinv.z);
  tn = make_float4(fmaxf(fmaxf(nx.x, ny.x), nz.x), fmaxf(fmaxf(nx.y, ny.y), nz.y), fmaxf(fmaxf(nx.z, ny.z), nz.z),
                   fmaxf(fmaxf(nx.w, ny.w), nz.w));
  tf = make_float4(fminf(fminf(fx.x, fy.x), fz.x), fminf(fminf(fx.y, fy.y), fz.y), fminf(fminf(fx.z, fy.z), fz.z),
                   fminf(fminf(fx.w, fy.w), fz.w));
}

template <int K, int BLOCK, bool COUNT>
TT_D void gather(const TraceParams &P, const RayPre &r, float tlo, long long plo, float far_t, KBuf<K> &b,
                 int *s_node, TravCount &tc) {
  kinit(b);
  if (P.n_leaves == 0) return;
  const NodePad np = node_pad(P, r);
  const int lane = threadIdx.x;
  // the data of `node`, fetched right after the node step that chose it and after the loads
  // of the leaf tested in the same iteration, so that the leaf test's work (and the rest of
  // the iteration) runs while the node's data is in flight
  float4 qa, qx, qy, qz;
  int4 qref;
  auto fetch = [&](int n) {
    const float4 *q = P.nodes4 + 5 * (size_t)n;
    qa = q[0], qx = q[1], qy = q[2], qz = q[3];
    qref = *reinterpret_cast<const int4 *>(q + 4);
  };
  fetch(0);
  int *ov = P.ov_stack + ((size_t)blockIdx.x * BLOCK + lane) * P.ov_depth;
  int *s_leaf = s_node + kStack * BLOCK;
  const float lo_w = tlo - fabsf(tlo) * kWideRel - kWideAbs;
  // the upper culling bound (the K-th key's t, or far_t), widened; it changes only when a hit
  // is inserted, so it is formed there rather than in every node step
  auto upper = [&]() {
    const float thi = fminf(far_t, b.t(K - 1));
    return thi + fabsf(thi) * kWideRel + kWideAbs;
  };
  float hi_w = upper();
  int node = 0;
  int sp = 0;
  int nq = 0;  // queued leaves of this lane
  auto push = [&](int n) {
    if (sp < kStack)
      s_node[sp * BLOCK + lane] = n;
    else
      ov[sp - kStack] = n;
    sp++;
  };
  while (true) {
    if (node >= 0 && nq <= kLeafQ - 4) {
      if (COUNT) tc.nodes++;
      float4 tn, tf;
      const int4 ref = qref;
      slab4q(r, np, qa, qx, qy, qz, tn, tf);
      const float tn0 = tn.x, tf0 = tf.x, tn1 = tn.y, tf1 = tf.y, tn2 = tn.z, tf2 = tf.z, tn3 = tn.w, tf3 = tf.w;
      // boxes hit (kEmptyChild: no child); a hit child is a leaf (ref < 0) or internal
      const bool h0 = (ref.x != kEmptyChild) & box_ok_wide(tn0, tf0, lo_w, hi_w);
      const bool h1 = (ref.y != kEmptyChild) & box_ok_wide(tn1, tf1, lo_w, hi_w);
      const bool h2 = (ref.z != kEmptyChild) & box_ok_wide(tn2, tf2, lo_w, hi_w);
      const bool h3 = (ref.w != kEmptyChild) & box_ok_wide(tn3, tf3, lo_w, hi_w);
      // hit leaves to the queue without branches: slots nq .. nq + 3 are free (nq <= kLeafQ - 4);
      // a write that is not counted is overwritten by the next one or lies past the queue's end.
      // The queue holds the child references (~leaf), complemented when a leaf is tested.
      s_leaf[nq * BLOCK + lane] = ref.x;
      nq += h0 & (ref.x < 0);
      s_leaf[nq * BLOCK + lane] = ref.y;
      nq += h1 & (ref.y < 0);
      s_leaf[nq * BLOCK + lane] = ref.z;
      nq += h2 & (ref.z < 0);
      s_leaf[nq * BLOCK + lane] = ref.w;
      nq += h3 & (ref.w < 0);
      // internal hit children by entry distance (the others at +inf, which no hit box reaches:
      // tn <= hi_w, finite), so after the sort child i is to be visited iff d_i < inf
      const float inf = __builtin_huge_valf();
      float d0 = (h0 & (ref.x >= 0)) ? tn0 : inf, d1 = (h1 & (ref.y >= 0)) ? tn1 : inf,
            d2 = (h2 & (ref.z >= 0)) ? tn2 : inf, d3 = (h3 & (ref.w >= 0)) ? tn3 : inf;
      int c0 = ref.x, c1 = ref.y, c2 = ref.z, c3 = ref.w;
      cswap(d0, c0, d1, c1);
      cswap(d2, c2, d3, c3);
      cswap(d0, c0, d2, c2);
      cswap(d1, c1, d3, c3);
      cswap(d1, c1, d2, c2);
      if (sp + 3 <= kStack) {  // the usual case: LDS slots sp .. sp + 2, written without branches
        s_node[sp * BLOCK + lane] = c3;
        sp += d3 < inf;
        s_node[sp * BLOCK + lane] = c2;
        sp += d2 < inf;
        s_node[sp * BLOCK + lane] = c1;
        sp += d1 < inf;
      } else {
        if (d3 < inf) push(c3);
        if (d2 < inf) push(c2);
        if (d1 < inf) push(c1);
      }
      node = d0 < inf ? c0 : -1;
      if (node < 0 && sp > 0) {
        sp--;
        node = s_node[min(sp, kStack - 1) * BLOCK + lane];
        if (sp >= kStack) node = ov[sp - kStack];
      }
    }
    const bool walked = node < 0;
    const int busy = __popcll(__ballot(nq > 0)), active = __popcll(__ballot(1));
    const bool test = nq > 0 && (walked || nq > kLeafQ - 4 || 2 * busy >= active);
    float4 a, bb, c;
    if (test) {
      const int s = ~s_leaf[--nq * BLOCK + lane];
      a = P.leaves[3 * s], bb = P.leaves[3 * s + 1], c = P.leaves[3 * s + 2];
    }
    fetch(node >= 0 ? node : 0);  // every iteration: the root once the walk is over (never read)
    if (test) {
      if (COUNT) tc.tris++;
      const uint32_t prim = __float_as_uint(a.w);
      float t;
      if (intersect(r, mk(a.x, a.y, a.z), mk(bb.x, bb.y, bb.z), mk(c.x, c.y, c.z), t) && t <= far_t &&
          after_key(t, prim, tlo, plo)) {
        kinsert(b, t, prim);
        hi_w = upper();
      }
    }
    if (walked && nq == 0) return;
  }
}

// The ray of pixel pix: generated from the view's pinhole camera (fused ray generation)
// or read from the ray buffers.
TT_D void pixel_ray(const TraceParams &P, size_t pix, v3f &o, v3f &d) {
  if (P.cam) {
    float fo[3], fd[3];
    pinhole_ray(P.cam, P.W, P.H, (int)(pix % (size_t)P.W), (int)(pix / (size_t)P.W), fo, fd);
    o = mk(fo[0], fo[1], fo[2]);
    d = mk(fd[0], fd[1], fd[2]);
  } else {
    o = ld3(P.ray_o + 3 * pix);
    d = ld3(P.ray_d + 3 * pix);
  }
}

// Cost estimate of tile `tile` for scheduling: the tile's centre ray walks the 4-wide
// BVH down to subtrees of at most kProbeCut leaves and sums the leaf counts of the
// subtrees and leaves whose boxes its line crosses -- how much geometry the tile's rays
// pass, at a few dozen node visits (a few percent of a full trace). One wave probes one
// tile, level by level: the nodes of a level are spread over the lanes and the children
// to open form the next level, so the chain of dependent node fetches is the tree's
// depth rather than the walk's length. Called by all 64 lanes of a one-wave block with
// two LDS frontiers of kProbeFront entries; nodes past a full frontier are dropped (the
// estimate then undercounts; it only orders tiles). Every lane returns the estimate.
constexpr int kProbeCut = 32, kProbeFront = 512;
template <int BLOCK>
TT_D float probe_tile(const TraceParams &P, int tile, int *s_front, int *s_count) {
  const int tx = tile % P.tiles_x, ty = tile / P.tiles_x;
  constexpr int TW = TileShape<BLOCK>::W, TH = TileShape<BLOCK>::H;
  const int x = min(tx * TW + TW / 2, P.W - 1), y = min(ty * TH + TH / 2, P.H - 1);
  const size_t pix = (size_t)y * P.W + x;
  const int lane = threadIdx.x;
  v3f ro, rd;
  pixel_ray(P, pix, ro, rd);
  const RayPre r = prep_ray(ro, rd);
  const float inf = __builtin_huge_valf();
  uint32_t est = 0;
  int *cur = s_front, *nxt = s_front + kProbeFront;
  if (lane == 0) cur[0] = 0, *s_count = 0;
  int n = P.n_leaves > 0 ? 1 : 0;
  __syncthreads();
  while (n > 0) {
    for (int i = lane; i < n; i += 64) {
      const int node = cur[i];
      float4 lx, hx, ly, hy, lz, hz;
      int4 ref;
      load_node4q(P.nodes4 + 5 * (size_t)node, lx, hx, ly, hy, lz, hz, ref);
      const uint2 cw = P.counts4[node];  // leaves under each child, 16 bits each (saturated)
      const uint32_t c01 = cw.x, c23 = cw.y;
      const int ct[4] = {(int)(c01 & 0xffffu), (int)(c01 >> 16), (int)(c23 & 0xffffu), (int)(c23 >> 16)};
      float tn[4], tf[4];
      slab(r, lx.x, hx.x, ly.x, hy.x, lz.x, hz.x, tn[0], tf[0]);
      slab(r, lx.y, hx.y, ly.y, hy.y, lz.y, hz.y, tn[1], tf[1]);
      slab(r, lx.z, hx.z, ly.z, hy.z, lz.z, hz.z, tn[2], tf[2]);
      slab(r, lx.w, hx.w, ly.w, hy.w, lz.w, hz.w, tn[3], tf[3]);
      const int rf[4] = {ref.x, ref.y, ref.z, ref.w};
      for (int c = 0; c < 4; c++) {
        if (rf[c] == kEmptyChild || !box_ok(tn[c], tf[c], 0.0f, inf)) continue;
        if (rf[c] < 0 || ct[c] <= kProbeCut) {
          est += (uint32_t)ct[c];
        } else {
          const int slot = atomicAdd(s_count, 1);
          if (slot < kProbeFront) nxt[slot] = rf[c];
        }
      }
    }
    __syncthreads();
    n = min(*s_count, kProbeFront);
    __syncthreads();
    if (lane == 0) *s_count = 0;
    int *t = cur;
    cur = nxt;
    nxt = t;
    __syncthreads();
  }
  for (int off = 32; off > 0; off >>= 1) est += __shfl_xor(est, off);
  return (float)est;
}

// Tile of a block in a one-shot launch: workgroups go to the 8 XCDs round-robin,
// so block i is mapped to tile (i % 8) * (n/8) + i / 8 (with the remainder spread
// over the first XCDs): each XCD's tiles form one contiguous band of the image.
TT_D int xcd_tile(int n_tiles) {
  const int i = blockIdx.x;
  const int xcd = i & 7, j = i >> 3;
  const int per = n_tiles >> 3, rem = n_tiles & 7;
  return xcd * per + min(xcd, rem) + j;
}

// Pixel of this lane in tile `tile` of an image with P.tiles_x tiles per row.
template <int BLOCK>
TT_D bool tile_pixel(const TraceParams &P, int tile, int &x, int &y) {
  const int tx = tile % P.tiles_x, ty = tile / P.tiles_x;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  x = tx * TileShape<BLOCK>::W + (wave & 1) * 8 + (lane & 7);
  y = ty * TileShape<BLOCK>::H + (wave >> 1) * 8 + (lane >> 3);
  return x < P.W && y < P.H;
}

// forward.cu:34-74 computeScaling
TT_D void compute_scaling(const TraceParams &P, int idx, v3f org) {
  if (P.scaling[idx] > 0.0f) return;
  const int off = P.cumsum[idx];
  const int np = P.npts[idx];
  v3f cen = mk(0.0f, 0.0f, 0.0f);
  for (int i = 0; i < np; i++) cen = cen + ld3(P.verts + 3 * (size_t)(off + i));
  cen = cen * (1.0f / (float)np);
  v3f tri2org = cen - org;
  float mx = 0.0f;
  for (int i = 0; i < np; i++) {
    v3f a = ld3(P.verts + 3 * (size_t)(off + i)) - org;
    float len_ab = vlen(a) * vlen(tri2org);
    float ang = 0.0f;
    if (!(len_ab < 1e-12f)) {
      float c = fmaxf(-1.0f, fminf(1.0f, dot(a, tri2org) / len_ab));
      ang = acosf(c);
    }
    mx = fmaxf(ang, mx);
  }
  P.scaling[idx] = mx;
}

struct SegState {
  float C[3];
  float T;
  float N[3];
  int consumed, blended;
  // depth interval of the blended hits: first/last t, t of the terminating hit, first prim
  float t_first, t_last, t_term;
  uint32_t p_first;
  // bounce variant: sums over blended hits of w rho, w (1 - rho), w eta, w t (w = alpha T)
  float R, Q, eta, D;
};

TT_D void seg_init(SegState &st) {
  st.C[0] = st.C[1] = st.C[2] = 0.0f;
  st.N[0] = st.N[1] = st.N[2] = 0.0f;
  st.T = 1.0f;
  st.consumed = st.blended = 0;
  st.t_first = st.t_last = st.t_term = __builtin_huge_valf();
  st.p_first = kInvalid;
  st.R = st.Q = st.eta = st.D = 0.0f;
}

// The chunk loop of forward.cu:76-155 (FULL: with normal/statistics outputs) and
// of the variants (colour only).
// Depth interval (exact chunk mode): hits start at t >= tstart; the gathers stop at
// `cap` until the hits up to it are used up, and go on past it (t > cap) only while
// the ray is still open. The result is that of a trace starting at tstart for any cap.
// With P.ival_hard the trace ends at the cap instead, with two fallbacks: a ray whose
// first hit at or after tstart lies within ival_fb_start of it (or that has no hit
// up to the cap) is traced again from 0 without cap, before anything is blended;
// and a ray still open at the cap that has no blended hit yet, or whose last
// blended hit lies within ival_fb_end of the cap, goes on past it.
// IVAL bits select the interval features compiled in: kIvalIn (a (start, cap) interval per
// ray is read), kIvalOut (the blended hits' interval is recorded for ival_out), kIvalHard (the
// cap may be hard, with its fallbacks, when P.ival_hard is set). Without kIvalHard a cap is
// exact (the trace continues past it), with fewer live registers.
constexpr int kIvalIn = 1, kIvalOut = 2, kIvalHard = 4;
// WREC: each consumed hit reads its triangle's window record (P.tri_win, written by the
// forward's record pass) instead of forming the triangle's terms itself; same results.
template <int K, int BLOCK, bool FULL, bool COUNT, bool BOUNCE = false, int IVAL = 0, bool WREC = false>
TT_D void trace_segment(const TraceParams &P, v3f o, v3f d, SegState &st, int ray, v3f scaling_org, int *s_node,
                        float tstart = 0.0f, float cap = kInfDist, int hseg = -1) {
  RayPre r = prep_ray(o, d);
  KBuf<K> b;
  float last = 0.0f, tlo = tstart;
  long long plo = -1;
  constexpr bool IN = (IVAL & kIvalIn) != 0, OUT = (IVAL & kIvalOut) != 0, HARD = IN && (IVAL & kIvalHard) != 0;
  bool capped = IN && cap < P.far_t && cap >= tstart;  // a cap before the start is no cap
  bool hard = HARD && P.ival_hard && capped;
  bool probe = HARD && P.ival_hard && tstart > 0.0f;  // first gather of a hard interval with a start
  const bool logging = FULL && P.log_cap > 0;
  TravCount tc;
  if (IN && tstart > P.far_t) {  // an interval past the far plane: nothing to gather, no fallback
    if (COUNT) P.trav_stats[2 * (size_t)ray] = P.trav_stats[2 * (size_t)ray + 1] = 0;
    return;
  }
  while (true) {
    gather<K, BLOCK, COUNT>(P, r, tlo, plo, capped ? cap : P.far_t, b, s_node, tc);
    if (probe) {
      probe = false;
      if (b.p(0) == kInvalid || b.t(0) < tstart + P.ival_fb_start) {  // the proposal may have missed: full trace
        tlo = 0.0f;
        plo = -1;
        capped = hard = false;
        continue;
      }
    }
    if (COUNT) {
      P.trav_stats[2 * (size_t)ray] = tc.nodes;
      P.trav_stats[2 * (size_t)ray + 1] = tc.tris;
    }
    float lt = 0.0f;
    uint32_t lp = 0;
    bool spent = false;
    for (int i = 0; i < K; i++) {
      const float t = b.t(0);
      const uint32_t prim = b.p(0);
      kpop(b);
      if (prim == kInvalid) {
        spent = true;
        break;
      }
      lt = t;
      lp = prim;
      const int slot = st.consumed;
      st.consumed++;
      last = fmaxf((float)((double)last + kStepEps), t);
      int code;
      if (t < P.near_t) {
        code = 0;
      } else {
        Win w;
        v3f VA, VB, VC;
        if (WREC) {
          TriWin tw;
          float sg;
          window_record_load(P.tri_win + kWinRec * (size_t)prim, tw, VA, VB, VC, sg);
          window_hit(tw, VA, VB, VC, o, d, t, sg, w);
        } else {
          const float *vp = P.verts + 3 * (size_t)P.cumsum[prim];
          window32(ld3(vp), ld3(vp + 3), ld3(vp + 6), o, d, t, P.sigma[prim], w);
        }
        const float alpha = fminf(0.99f, P.opac[prim] * w.Cx);
        if (alpha < 1.0f / 255.0f) {
          code = 1;
        } else {
          const float testT = st.T * (1.0f - alpha);
          if (testT < P.cutoff) {
            if (logging && slot < P.log_cap) {
              P.log_prim[(size_t)ray * P.log_cap + slot] = (int)prim;
              P.log_t[(size_t)ray * P.log_cap + slot] = t;
              P.log_code[(size_t)ray * P.log_cap + slot] = 3;
            }
            if (OUT) st.t_term = t;
            return;
          }
          if (OUT && st.blended == 0) st.t_first = t, st.p_first = prim;
          if (OUT || HARD) st.t_last = t;
          const float bw = alpha * st.T;
          if (FULL) {
            P.density[prim] = 1.0f;
            compute_scaling(P, prim, scaling_org);
            atomicMax(reinterpret_cast<int *>(P.maxb) + prim, __float_as_int(bw));
            const v3f nrm = WREC ? face_normal(VA, VB, VC) : w.n;
            st.N[0] += nrm.x * bw;
            st.N[1] += nrm.y * bw;
            st.N[2] += nrm.z * bw;
          }
          bool c0, c1, c2;
          const v3f col = sh_color(P.shs + (size_t)prim * P.M * 3, P.deg, d, c0, c1, c2);
          st.C[0] += col.x * bw;
          st.C[1] += col.y * bw;
          st.C[2] += col.z * bw;
          if (BOUNCE) {
            const float rho = P.refl[prim];
            st.R += bw * rho;
            st.Q += bw * (1.0f - rho);
            st.eta += bw * P.ior[prim];
            st.D += bw * t;
          }
          st.T = testT;
          if (hseg >= 0 && P.hl && st.blended < P.hl_cap) {  // hit list of segment hseg
            const size_t e = ((size_t)hseg * P.hl_cap + st.blended) * P.hl_stride + ray;
            P.hl[e] = make_uint2(prim, __float_as_uint(t));
            if (P.hl_rgb) {  // the colour as blended; a clamped channel as -0.0f (+ 0.0f: no other -0)
              float *c = P.hl_rgb + 3 * e;
              c[0] = c0 ? -0.0f : col.x + 0.0f;
              c[1] = c1 ? -0.0f : col.y + 0.0f;
              c[2] = c2 ? -0.0f : col.z + 0.0f;
            }
          }
          st.blended++;
          code = 2;
        }
      }
      if (logging && slot < P.log_cap) {
        P.log_prim[(size_t)ray * P.log_cap + slot] = (int)prim;
        P.log_t[(size_t)ray * P.log_cap + slot] = t;
        P.log_code[(size_t)ray * P.log_cap + slot] = code;
      }
    }
    if (spent) {
      if (!capped || (hard && !(st.t_last > cap - P.ival_fb_end))) return;
      capped = false;  // everything up to cap is used: continue strictly past it
      tlo = cap;
      plo = (long long)kInvalid;
      continue;
    }
    if (P.chunk_mode == 1) {
      tlo = lt;
      plo = (long long)lp;
    } else if (last > tlo) {
      tlo = last;
      plo = -1;
    } else {
      plo = (long long)lp;
    }
  }
}

// Start of the bounce variant's secondary rays (this project's model): in exact chunk mode
// they begin at the key (near_t, -1) -- hits before near_t would only be consumed and
// skipped, so the blended result is the same and the gather does not collect them; in
// compat mode at 0 (its chunk boundaries depend on every consumed hit).
TT_D float bounce_start(const TraceParams &P) { return P.chunk_mode == 1 ? P.near_t : 0.0f; }

// Sort key of a bounce pixel's secondary rays: the pixel's 128 x 128 region (Morton order,
// 8 bits), then the reflected direction's octahedral bin (6 + 6 bits), then the pixel's
// Morton order inside the region at 2-pixel granularity (12 bits): a wave traces rays that
// start close together and point in similar directions.
TT_D uint32_t morton2(uint32_t x, uint32_t y, int bits) {
  uint32_t m = 0;
  for (int b = 0; b < bits; b++) m |= ((x >> b) & 1u) << (2 * b) | ((y >> b) & 1u) << (2 * b + 1);
  return m;
}
TT_D uint32_t ray_bin_key(v3f dir, int x, int y) {
  const float l1 = fabsf(dir.x) + fabsf(dir.y) + fabsf(dir.z);
  float u = dir.x / l1, v = dir.y / l1;
  if (dir.z < 0.0f) {
    const float uu = (1.0f - fabsf(v)) * (u >= 0.0f ? 1.0f : -1.0f);
    v = (1.0f - fabsf(u)) * (v >= 0.0f ? 1.0f : -1.0f);
    u = uu;
  }
  const uint32_t qu = (uint32_t)min(63, max(0, (int)((u * 0.5f + 0.5f) * 64.0f)));
  const uint32_t qv = (uint32_t)min(63, max(0, (int)((v * 0.5f + 0.5f) * 64.0f)));
  const uint32_t region = morton2((uint32_t)min(x >> 7, 15), (uint32_t)min(y >> 7, 15), 4);
  const uint32_t key =
      region << 24 | morton2(qu, qv, 6) << 12 | morton2((uint32_t)((x & 127) >> 1), (uint32_t)((y & 127) >> 1), 6);
  return min(key, 0xfffffffeu);  // 0xffffffff marks pixels without secondary rays
}

// One pixel of the forward programs (forward.cu:161-215 and the variants' raygens).
// Returns the pixel's blended hits over all its traced segments. VARIANT 5: the bounce
// variant's primary segment only -- the colour buffer then holds the primary's composite
// C and aux its transmittance (slot 5), R, Q and the two secondary rays, and ray_keys the
// pixel's sort key (of its reflected ray), for bounce_secondary.
template <int K, int BLOCK, int VARIANT, bool COUNT, int IVAL = 0, bool WREC = false>
TT_D int forward_pixel(const TraceParams &P, int pix, int *s_node, float *loss = nullptr) {
  v3f o, d;
  pixel_ray(P, (size_t)pix, o, d);
  SegState st;
  seg_init(st);
  if (VARIANT == 0) {
    float tstart = 0.0f, cap = kInfDist;
    if ((IVAL & kIvalIn) && P.ival_in) {
      const float2 iv = P.ival_in[pix];
      tstart = fmaxf(iv.x, 0.0f);  // a start before the origin is the full trace (hit keys need t >= 0)
      cap = iv.y;
    }
    trace_segment<K, BLOCK, true, COUNT, false, IVAL, WREC>(P, o, d, st, pix, o, s_node, tstart, cap, 0);
    if ((IVAL & kIvalOut) && P.ival_out)
      P.ival_out[pix] = make_float4(st.t_first, st.t_last, st.t_term, __uint_as_float(st.p_first));
    if (P.hl_count) P.hl_count[pix] = st.blended <= P.hl_cap ? st.blended : -1;
  } else if (VARIANT == 5) {
    trace_segment<K, BLOCK, true, COUNT, true, 0, WREC>(P, o, d, st, pix, o, s_node, 0.0f, kInfDist, 0);
    const Bounce g = bounce_rays(o, d, st.T, st.N, st.D, st.eta);
    if (P.hl_count) {
      P.hl_count[pix] = st.blended <= P.hl_cap ? st.blended : -1;
      P.hl_count[P.hl_stride + pix] = 0;
      P.hl_count[2 * P.hl_stride + pix] = 0;
    }
    const size_t R = (size_t)P.R;
    float *a = P.aux + (size_t)kBounceAux * pix;
    const v3f zero = mk(0.0f, 0.0f, 0.0f);
    const v3f orr = g.valid ? g.orr : zero, dr = g.valid ? g.dr : zero, ot = g.valid ? g.ot : zero,
              dt = g.valid ? g.dt : zero;
    const float rec[kBounceAux] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, st.T, st.R, st.Q, orr.x, orr.y,
                                   orr.z, dr.x, dr.y, dr.z, ot.x, ot.y, ot.z, dt.x, dt.y, dt.z};
    for (int i = 0; i < kBounceAux; i++) a[i] = rec[i];
    if (g.valid) {
      for (int c = 0; c < 3; c++) P.color[c * R + pix] = st.C[c];
      P.ray_keys[pix] = ray_bin_key(g.dr, pix % P.W, pix / P.W);
    } else {  // no secondary rays: the final colour now (Lr = Lt = 0), as VARIANT 4 forms it
      float E0[3];
      env_value(P.env, P.env_w, P.env_h, d, E0);
      for (int c = 0; c < 3; c++)
        P.color[c * R + pix] = ((st.C[c] + st.R * 0.0f) + st.Q * 0.0f) + st.T * E0[c];
      a[5] = 0.0f;
      P.ray_keys[pix] = 0xffffffffu;
    }
  } else if (VARIANT == 4) {
    trace_segment<K, BLOCK, true, COUNT, true, 0, WREC>(P, o, d, st, pix, o, s_node, 0.0f, kInfDist, 0);
    int nb[3] = {st.blended, 0, 0};  // blended hits of the primary, reflected and refracted segment
    const Bounce g = bounce_rays(o, d, st.T, st.N, st.D, st.eta);
    float Lr[3] = {0.0f, 0.0f, 0.0f}, Lt[3] = {0.0f, 0.0f, 0.0f};
    if (g.valid) {
      float E[3];
      SegState sr;
      seg_init(sr);
      trace_segment<K, BLOCK, false, false, false, 0, WREC>(P, g.orr, g.dr, sr, pix, g.orr, s_node, bounce_start(P), kInfDist, 1);
      env_value(P.env, P.env_w, P.env_h, g.dr, E);
      for (int c = 0; c < 3; c++) Lr[c] = sr.C[c] + sr.T * E[c];
      SegState stt;
      seg_init(stt);
      trace_segment<K, BLOCK, false, false, false, 0, WREC>(P, g.ot, g.dt, stt, pix, g.ot, s_node, bounce_start(P), kInfDist, 2);
      env_value(P.env, P.env_w, P.env_h, g.dt, E);
      for (int c = 0; c < 3; c++) Lt[c] = stt.C[c] + stt.T * E[c];
      st.consumed += sr.consumed + stt.consumed;
      st.blended += sr.blended + stt.blended;
      nb[1] = sr.blended;
      nb[2] = stt.blended;
    }
    if (P.hl_count)
      for (int sg = 0; sg < 3; sg++) P.hl_count[sg * P.hl_stride + pix] = nb[sg] <= P.hl_cap ? nb[sg] : -1;
    float E0[3];
    env_value(P.env, P.env_w, P.env_h, d, E0);
    const size_t R = (size_t)P.R;
    for (int c = 0; c < 3; c++) P.color[c * R + pix] = ((st.C[c] + st.R * Lr[c]) + st.Q * Lt[c]) + st.T * E0[c];
    float *a = P.aux + (size_t)kBounceAux * pix;
    const v3f zero = mk(0.0f, 0.0f, 0.0f);
    const v3f orr = g.valid ? g.orr : zero, dr = g.valid ? g.dr : zero, ot = g.valid ? g.ot : zero,
              dt = g.valid ? g.dt : zero;
    const float rec[kBounceAux] = {Lr[0], Lr[1], Lr[2], Lt[0], Lt[1], Lt[2], st.R, st.Q, orr.x, orr.y,
                                   orr.z, dr.x, dr.y, dr.z, ot.x, ot.y, ot.z, dt.x, dt.y, dt.z};
    for (int i = 0; i < kBounceAux; i++) a[i] = rec[i];
  } else {
    const float rad = P.extra_len > 0 ? P.extra[0] : 1.0f;
    const v3f cnt = P.extra_len > 3 ? mk(P.extra[1], P.extra[2], P.extra[3]) : mk(0.0f, 0.0f, 0.0f);
    const float eps = (float)kStepEps;
    if (VARIANT == 1) {  // reflect.cu:25-48
      float sth;
      if (hit_sphere(rad, cnt, o, d, sth)) {
        o = o + d * sth;
        const v3f nrm = normalize(o - cnt);
        o = o + nrm * eps;
        d = reflect3(normalize(d), nrm);
      }
      trace_segment<K, BLOCK, false, false, false, 0, WREC>(P, o, d, st, pix, o, s_node);
    } else if (VARIANT == 2) {  // refract.cu:189-213
      float sth;
      if (hit_sphere(rad, cnt, o, d, sth)) {
        const v3f hp = o + d * sth;
        const v3f nrm = normalize(hp - cnt);
        const v3f in = normalize(d);
        const float ior = 1.5f / 1.0003f;
        v3f nd = in;
        if (!refract_quirk(nd, in, nrm, ior)) {
          const v3f rn = dot(in, nrm) < 0.0f ? nrm : -nrm;
          nd = reflect3(in, rn);
        }
        o = hp + nrm * eps;
        d = nd;
      }
      trace_segment<K, BLOCK, false, false, false, 0, WREC>(P, o, d, st, pix, o, s_node);
    } else {  // envlight.cu:116-184
      const int samples = (int)(P.extra_len > 4 ? P.extra[4] : 0.0f);
      float sth;
      if (hit_sphere(rad, cnt, o, d, sth)) {
        o = o + d * sth;
        const v3f nrm0 = normalize(o - cnt);
        o = o + nrm0 * eps;
        float radc[3] = {0.0f, 0.0f, 0.0f};
        for (int i = 0; i < samples; i++) {
          const v3f nrm = normalize(o - cnt);
          uint32_t seed = (((uint32_t)pix + 1u) * 9781u) ^ (((uint32_t)i + 1u) * 6271u);
          const v3f dn = cos_hemi(seed, nrm);
          const v3f on = o + dn * eps;
          SegState ss;
          seg_init(ss);
          trace_segment<K, BLOCK, false, false, false, 0, WREC>(P, on, dn, ss, pix, on, s_node);
          for (int c = 0; c < 3; c++) radc[c] += ss.C[c] + ss.T * P.bg[c];
          st.consumed += ss.consumed;
          st.blended += ss.blended;
        }
        const float inv = 1.0f / (float)samples;
        for (int c = 0; c < 3; c++) st.C[c] += radc[c] * inv;
        st.T = 0.0f;
      } else {
        trace_segment<K, BLOCK, false, false, false, 0, WREC>(P, o, d, st, pix, o, s_node);
      }
    }
  }
  const size_t R = (size_t)P.R;
  if (VARIANT != 4 && VARIANT != 5)
    for (int c = 0; c < 3; c++) P.color[c * R + pix] = st.C[c] + st.T * P.bg[c];
  if (VARIANT == 0 && P.target) {  // fused loss: dL/dcolour for the backward, this pixel's loss
    float l = 0.0f;
    for (int c = 0; c < 3; c++) {
      const float diff = (st.C[c] + st.T * P.bg[c]) - P.target[c * R + pix];
      if (P.loss_kind == 1) {
        P.loss_grad[c * R + pix] = (2.0f * diff) * P.loss_norm;
        l += diff * diff;
      } else {
        P.loss_grad[c * R + pix] = diff > 0.0f ? P.loss_norm : (diff < 0.0f ? -P.loss_norm : 0.0f);
        l += fabsf(diff);
      }
    }
    if (loss) *loss = l;
  }
  if (VARIANT == 0 || VARIANT == 4 || VARIANT == 5) {
    P.others[0 * R + pix] = 0.0f;
    P.others[1 * R + pix] = 1.0f - st.T;
    P.others[2 * R + pix] = st.N[0];
    P.others[3 * R + pix] = st.N[1];
    P.others[4 * R + pix] = st.N[2];
    P.others[5 * R + pix] = 0.0f;
    P.others[6 * R + pix] = 0.0f;
    if (P.log_cap > 0) P.log_count[pix] = st.consumed;
  }
  if (P.stats) {
    P.stats[2 * (size_t)pix] = st.consumed;
    P.stats[2 * (size_t)pix + 1] = st.blended;
  }
  return st.blended;
}

// The secondary segments of bounce pixel pix after its primary (VARIANT 5): the reflected
// and refracted rays from aux, their radiance L = C' + T' E(d') into aux 0..5, the final
// colour ((C + R L_r) + Q L_t) + T E(d) (the operation order of VARIANT 4), the hit counts
// of both segments, and the statistics and the tile's blended hits added to.
template <int K, int BLOCK, bool WREC = false>
TT_D void bounce_secondary(const TraceParams &P, int pix, int *s_node) {
  float *a = P.aux + (size_t)kBounceAux * pix;
  const v3f orr = mk(a[8], a[9], a[10]), dr = mk(a[11], a[12], a[13]);
  const v3f ot = mk(a[14], a[15], a[16]), dt = mk(a[17], a[18], a[19]);
  float E[3], Lr[3], Lt[3];
  SegState sr;
  seg_init(sr);
  trace_segment<K, BLOCK, false, false, false, 0, WREC>(P, orr, dr, sr, pix, orr, s_node, bounce_start(P), kInfDist, 1);
  env_value(P.env, P.env_w, P.env_h, dr, E);
  for (int c = 0; c < 3; c++) Lr[c] = sr.C[c] + sr.T * E[c];
  SegState stt;
  seg_init(stt);
  trace_segment<K, BLOCK, false, false, false, 0, WREC>(P, ot, dt, stt, pix, ot, s_node, bounce_start(P), kInfDist, 2);
  env_value(P.env, P.env_w, P.env_h, dt, E);
  for (int c = 0; c < 3; c++) Lt[c] = stt.C[c] + stt.T * E[c];
  if (P.hl_count) {
    P.hl_count[P.hl_stride + pix] = sr.blended <= P.hl_cap ? sr.blended : -1;
    P.hl_count[2 * P.hl_stride + pix] = stt.blended <= P.hl_cap ? stt.blended : -1;
  }
  const size_t R = (size_t)P.R;
  const float T = a[5], Rw = a[6], Qw = a[7];
  float E0[3];
  v3f po, pd;
  pixel_ray(P, (size_t)pix, po, pd);
  env_value(P.env, P.env_w, P.env_h, pd, E0);
  for (int c = 0; c < 3; c++) {
    const float C = P.color[c * R + pix];
    P.color[c * R + pix] = ((C + Rw * Lr[c]) + Qw * Lt[c]) + T * E0[c];
  }
  for (int c = 0; c < 3; c++) a[c] = Lr[c], a[3 + c] = Lt[c];
  const int nb = sr.blended + stt.blended;
  if (P.stats) {
    P.stats[2 * (size_t)pix] += sr.consumed + stt.consumed;
    P.stats[2 * (size_t)pix + 1] += nb;
  }
  if (P.tile_cost && P.tile_hits && nb > 0) {
    const int x = pix % P.W, y = pix / P.W;
    atomicAdd(P.tile_hits + (y / TileShape<BLOCK>::H) * P.tiles_x + x / TileShape<BLOCK>::W, nb);
  }
}

// ------------------------------------------------------------------ backward
TT_D v3f vecmat(v3f v, const float A[9]) {
  return {A[0] * v.x + A[3] * v.y + A[6] * v.z, A[1] * v.x + A[4] * v.y + A[7] * v.z,
          A[2] * v.x + A[5] * v.y + A[8] * v.z};
}
TT_D void skew(v3f v, float S[9]) {
  S[0] = 0.0f, S[1] = -v.z, S[2] = v.y;
  S[3] = v.z, S[4] = 0.0f, S[5] = -v.x;
  S[6] = -v.y, S[7] = v.x, S[8] = 0.0f;
}
TT_D void outer(v3f u, v3f v, float M[9]) {
  M[0] = u.x * v.x, M[1] = u.x * v.y, M[2] = u.x * v.z;
  M[3] = u.y * v.x, M[4] = u.y * v.y, M[5] = u.y * v.z;
  M[6] = u.z * v.x, M[7] = u.z * v.y, M[8] = u.z * v.z;
}
TT_D void eye_scaled(float s, float M[9]) {
#pragma unroll
  for (int i = 0; i < 9; i++) M[i] = ((i % 4) == 0 ? 1.0f : 0.0f) * s;
}
// d(v/|v|)/dv = (I - n n^T)/|v| (vecmath.h:2798-2809); scale formed as 1/|v| in double.
TT_D void dnormvdv(v3f nv, v3f v, float out[9]) {
  float op[9];
  outer(nv, nv, op);
#pragma unroll
  for (int i = 0; i < 9; i++) out[i] = ((i % 4) == 0 ? 1.0f : 0.0f) - op[i];
  const float s = (float)(1.0 / (double)vlen(v));
#pragma unroll
  for (int i = 0; i < 9; i++) out[i] *= s;
}
// Jacobians of ne = ((p1-p3) x (p2-p3)) x (p1-p2) (vecmath.h:2811-2859).
TT_D void dne_dp1(v3f p1, v3f p2, v3f p3, float out[9]) {
  v3f b = p2 - p3, c = p1 - p2;
  v3f axb = cross(p1 - p3, b);
  float m[9], d[9], s[9];
  outer(b, c, m);
  eye_scaled(dot(b, c), d);
  skew(axb, s);
#pragma unroll
  for (int i = 0; i < 9; i++) out[i] = (m[i] - d[i]) + s[i];
}
TT_D void dne_dp2(v3f p1, v3f p2, v3f p3, float out[9]) {
  v3f a = p1 - p3, b = p2 - p3, c = p1 - p2;
  v3f axb = cross(p1 - p3, b);
  float d[9], m[9], s[9];
  eye_scaled(dot(a, c), d);
  outer(a, c, m);
  skew(axb, s);
#pragma unroll
  for (int i = 0; i < 9; i++) out[i] = (d[i] - m[i]) - s[i];
}
TT_D void dne_dp3(v3f p1, v3f p2, v3f p3, float out[9]) {
  v3f c = p1 - p2;
  float d[9], m[9];
  eye_scaled(dot(c, c), d);
  outer(c, c, m);
#pragma unroll
  for (int i = 0; i < 9; i++) out[i] = m[i] - d[i];
}

// ---- gradient fields: 9 vertex coordinates, opacity, sigma, 16 x 3 SH coefficients
constexpr int F_OPAC = 9, F_SIGMA = 10, F_SH = 11;

// Slot of `prim` in the block's table of SLOTS entries (Fibonacci hash, linear
// probing, 8 probes), or -1 when full.
template <int SLOTS>
TT_D int claim_slot(int *keys, uint32_t prim) {
  static_assert(SLOTS >= 64 && SLOTS <= 2048 && (SLOTS & (SLOTS - 1)) == 0, "table size");
  constexpr int bits = __builtin_ctz(SLOTS);
  uint32_t h = (prim * 2654435761u) >> (32 - bits);
  for (int probe = 0; probe < 8; probe++) {
    const int old = atomicCAS(&keys[h], -1, (int)prim);
    if (old == -1 || old == (int)prim) return (int)h;
    h = (h + 1) & (SLOTS - 1);
  }
  return -1;
}

// Register sinks: the 11 per-hit scalars (9 vertex coordinates, opacity, sigma) of a
// gradient record, and the 48 SH sums of the record reduce.
struct HitSink {
  float v[11];
  TT_D void add(int f, float x) { v[f] += x; }
};
struct ShSink {
  float v[48];
  TT_D void add(int f, float x) { v[f - F_SH] += x; }
};
// Straight to the global gradient buffers (records that found no room).
struct GlobalSink {
  float *g_verts, *g_opac, *g_sigma, *g_sh;
  TT_D void add(int f, float v) {
    if (f < 9)
      unsafeAtomicAdd(g_verts + f, v);
    else if (f == F_OPAC)
      unsafeAtomicAdd(g_opac, v);
    else if (f == F_SIGMA)
      unsafeAtomicAdd(g_sigma, v);
    else
      unsafeAtomicAdd(g_sh + (f - F_SH), v);
  }
};

// computeWindowFuncBwd (backward.cu:104-189): gradient of the max edge distance
// (|d_k - phi| <= 1e-6) and of the face normal w.r.t. the three vertices, hit
// point and incenter distance held constant. Per-vertex sums are formed in
// registers before they reach the sink.
template <class Sink>
TT_D void window_bwd(const v3f V[3], const Win &w, float dL_dphi, v3f dL_dn3, Sink &sink) {
  v3f g[3] = {mk(0.f, 0.f, 0.f), mk(0.f, 0.f, 0.f), mk(0.f, 0.f, 0.f)};
  // The edges at the window's maximum (one, or two or three at ties), in ascending order.
  // Each pass takes a lane's next such edge, with the vertices rotated so that it is edge
  // (p1, p2): one copy of the Jacobian code runs per pass, however the edges differ across
  // the wave's lanes (an unrolled loop over the three edges ran every edge's copy that any
  // lane needed). The sums into g keep the ascending-edge order.
  unsigned pend = 0;
  for (int k = 0; k < 3; k++) pend |= (fabsf(w.ed[k] - w.phi_max) > 1e-6f ? 0u : 1u) << k;
  const v3f V0 = V[0], V1 = V[1], V2 = V[2], R0 = w.rel[0], R1 = w.rel[1], R2 = w.rel[2];
  auto pick = [](v3f a0, v3f a1, v3f a2, int i) { return i == 0 ? a0 : (i == 1 ? a1 : a2); };
  while (pend) {
    const int k = __builtin_ctz(pend);
    pend &= pend - 1;
    const int k1 = k == 2 ? 0 : k + 1, k2 = k == 0 ? 2 : k - 1;
    const v3f p1 = pick(V0, V1, V2, k), p2 = pick(V0, V1, V2, k1), p3 = pick(V0, V1, V2, k2);
    const v3f n = cross(p1 - p3, p2 - p3);
    const v3f ne = cross(n, p1 - p2);
    const v3f Ne = normalize(ne);
    float J[9];
    dnormvdv(Ne, ne, J);
    const v3f dL_dNe = dL_dphi * pick(R0, R1, R2, k);
    const v3f dL_dne = vecmat(dL_dNe, J);
    float j1[9], j2[9], j3[9];
    dne_dp1(p1, p2, p3, j1);
    dne_dp2(p1, p2, p3, j2);
    dne_dp3(p1, p2, p3, j3);
    const v3f g1 = vecmat(dL_dne, j1) - dL_dphi * Ne;
    const v3f g2 = vecmat(dL_dne, j2);
    const v3f g3 = vecmat(dL_dne, j3);
    for (int j = 0; j < 3; j++) g[j] = g[j] + (j == k ? g1 : (j == k1 ? g2 : g3));
  }
  if (dL_dn3.x != 0.0f || dL_dn3.y != 0.0f || dL_dn3.z != 0.0f) {
    const v3f p1 = V[0], p2 = V[1], p3 = V[2];
    const v3f n = cross(p1 - p3, p2 - p3);
    const v3f Nn = normalize(n);
    float J[9];
    dnormvdv(Nn, n, J);
    const v3f dL_dn = vecmat(dL_dn3, J);
    float s1[9], s2[9], s3[9];
    skew(p3 - p2, s1);
    skew(p1 - p3, s2);
    skew(p2 - p1, s3);
    g[0] = g[0] + vecmat(dL_dn, s1);
    g[1] = g[1] + vecmat(dL_dn, s2);
    g[2] = g[2] + vecmat(dL_dn, s3);
  }
  for (int j = 0; j < 3; j++) {
    sink.add(3 * j, g[j].x);
    sink.add(3 * j + 1, g[j].y);
    sink.add(3 * j + 2, g[j].z);
  }
}

// computeColorFromSHBwd (backward.cu:33-102); clamped channels carry no gradient.
template <class Sink>
TT_D void sh_bwd(int deg, v3f dir0, bool c0, bool c1, bool c2, const float dc[3], Sink &sink) {
  const v3f dir = divs(dir0, vlen(dir0));
  v3f g = mk(dc[0], dc[1], dc[2]);
  g.x *= c0 ? 0.0f : 1.0f;
  g.y *= c1 ? 0.0f : 1.0f;
  g.z *= c2 ? 0.0f : 1.0f;
  const float x = dir.x, y = dir.y, z = dir.z;
  auto add = [&](int j, v3f v) {
    if (!c0) sink.add(F_SH + 3 * j, v.x);
    if (!c1) sink.add(F_SH + 3 * j + 1, v.y);
    if (!c2) sink.add(F_SH + 3 * j + 2, v.z);
  };
  add(0, SH_C0 * g);
  if (deg > 0) {
    add(1, (-SH_C1 * y) * g);
    add(2, (SH_C1 * z) * g);
    add(3, (-SH_C1 * x) * g);
    if (deg > 1) {
      const float xx = x * x, yy = y * y, zz = z * z;
      const float xy = x * y, yz = y * z, xz = x * z;
      add(4, (SH_C2_0 * xy) * g);
      add(5, (SH_C2_1 * yz) * g);
      add(6, (SH_C2_2 * (2.0f * zz - xx - yy)) * g);
      add(7, (SH_C2_3 * xz) * g);
      add(8, (SH_C2_4 * (xx - yy)) * g);
      if (deg > 2) {
        add(9, ((SH_C3_0 * y) * (3.0f * xx - yy)) * g);
        add(10, ((SH_C3_1 * xy) * z) * g);
        add(11, ((SH_C3_2 * y) * (4.0f * zz - xx - yy)) * g);
        add(12, ((SH_C3_3 * z) * (2.0f * zz - 3.0f * xx - 3.0f * yy)) * g);
        add(13, ((SH_C3_4 * x) * (4.0f * zz - xx - yy)) * g);
        add(14, ((SH_C3_5 * z) * (xx - yy)) * g);
        add(15, ((SH_C3_6 * x) * (xx - 3.0f * yy)) * g);
      }
    }
  }
}

// Gradient record of one blended hit. Records that the emitting block sums itself are
// 4 x float4 = 64 B, with the ray's direction kept once per lane in the block's LDS (rec_dirs):
//   (prim, clamp bits | lane << 3, dL/dc[0], dL/dc[1]) (dL/dc[2], vertex 0..2) (vertex 3..6)
//   (vertex 7, vertex 8, opacity, sigma)
// Records summed over the whole launch after a sort by triangle (the bounce secondaries)
// carry the direction, 5 x float4 = 80 B:
//   (prim, clamp bits, d.x, d.y) (d.z, dL/dc[3]) (vertex 0..3) (vertex 4..7) (vertex 8, opacity, sigma, 0)
// The SH gradient (48 values) is rebuilt from the direction, dL/dcolour and the clamp
// bits when the records are summed. A block writes its records into its own
// region of P.rec_rows, rows [s_rec[1], s_rec[1] + s_rec[2]); the row index comes from one
// LDS atomic per wave (counter s_rec[0]). A one-shot backward that follows a forward of the
// same image sizes each tile's region to that tile's blended-hit count in the forward
// (P.tile_hits, prefix sums in P.rec_off); otherwise a block has P.rec_lane_rows rows per
// lane. Past its region a hit's gradient goes to the global buffers with atomics instead.
constexpr int kRecLds = 4;  // s_rec: counter, region start, region rows, work items of the aggregation
template <int BLOCK>
constexpr int rec_batch() {  // records grouped per pass (two shorts each in the traversal LDS)
  return 16 * BLOCK;
}
template <int BLOCK>
constexpr int agg_slots() {  // triangles a block groups one batch of records into
  return 8 * BLOCK;
}
constexpr int kRowLocal = 4, kRowGlobal = 5;  // float4 per record: block-summed / launch-summed
template <int BLOCK>
constexpr int rec_lds_ints() {  // s_rec: header, the first batch's triangle ids, the lanes' ray directions
  return kRecLds + rec_batch<BLOCK>() + 3 * BLOCK;
}
template <int BLOCK>
TT_D float *rec_dirs(int *s_rec) {
  return reinterpret_cast<float *>(s_rec + kRecLds + rec_batch<BLOCK>());
}
constexpr int kStageSlots = 32;  // triangles whose sums are staged in LDS per coalesced flush
constexpr int kSegRecords = 16;  // records one thread sums at most (a triangle's longer runs are split)
template <int BLOCK>
constexpr int rec_tab_ints() {  // s_tab of backward_block_rec: grouping table + flush stage
  return 3 * agg_slots<BLOCK>() + 64 * kStageSlots;
}
template <int BLOCK, bool LOCAL>
TT_D void emit_record(const TraceParams &P, uint32_t prim, int off, v3f d, const float dc[3], bool c0, bool c1,
                      bool c2, const HitSink &h, int *s_rec) {
  const unsigned long long m = __ballot(1);
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((unsigned long long)m) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(&s_rec[0], __popcll(m));
  base = __shfl(base, leader);
  const int idx = base + __popcll(m & ((1ull << lane) - 1ull));
  if (idx < s_rec[2]) {
    if (idx < rec_batch<BLOCK>()) s_rec[kRecLds + idx] = (int)prim;  // the grouping's first batch reads it here
    if (P.rec_keys) P.rec_keys[(size_t)s_rec[1] + idx] = prim;
    const int cl = (c0 ? 1 : 0) | (c1 ? 2 : 0) | (c2 ? 4 : 0);
    if (LOCAL) {  // every record of this lane in this launch has this direction
      float *sd = rec_dirs<BLOCK>(s_rec) + 3 * threadIdx.x;
      sd[0] = d.x, sd[1] = d.y, sd[2] = d.z;
      float4 *row = P.rec_rows + kRowLocal * ((size_t)s_rec[1] + idx);
      row[0] = make_float4(__uint_as_float(prim), __int_as_float(cl | (int)threadIdx.x << 3), dc[0], dc[1]);
      row[1] = make_float4(dc[2], h.v[0], h.v[1], h.v[2]);
      row[2] = make_float4(h.v[3], h.v[4], h.v[5], h.v[6]);
      row[3] = make_float4(h.v[7], h.v[8], h.v[9], h.v[10]);
    } else {
      float4 *row = P.rec_rows + kRowGlobal * ((size_t)s_rec[1] + idx);
      row[0] = make_float4(__uint_as_float(prim), __int_as_float(cl), d.x, d.y);
      row[1] = make_float4(d.z, dc[0], dc[1], dc[2]);
      row[2] = make_float4(h.v[0], h.v[1], h.v[2], h.v[3]);
      row[3] = make_float4(h.v[4], h.v[5], h.v[6], h.v[7]);
      row[4] = make_float4(h.v[8], h.v[9], h.v[10], 0.0f);
    }
  } else {
    GlobalSink g{P.dverts + 3 * (size_t)off, P.dopac + prim, P.dsigma + prim, P.dsh + (size_t)prim * P.M * 3};
    for (int f = 0; f < 11; f++) g.add(f, h.v[f]);
    sh_bwd(P.deg, d, c0, c1, c2, dc, g);
  }
}

// Adds record r (an absolute record of P.rec_rows) to register sums; returns its triangle.
// LOCAL: a block-summed record, whose direction is in `dirs` (the block's rec_dirs).
template <bool LOCAL>
TT_D uint32_t add_record(const TraceParams &P, size_t r, float acc[11], ShSink &sh, const float *dirs = nullptr) {
  if (LOCAL) {
    const float4 *row = P.rec_rows + kRowLocal * r;
    const float4 r0 = row[0], r1 = row[1], r2 = row[2], r3 = row[3];
    const int cl = __float_as_int(r0.y);
    const float *sd = dirs + 3 * (cl >> 3);
    const float dc[3] = {r0.z, r0.w, r1.x};
    acc[0] += r1.y, acc[1] += r1.z, acc[2] += r1.w;
    acc[3] += r2.x, acc[4] += r2.y, acc[5] += r2.z, acc[6] += r2.w;
    acc[7] += r3.x, acc[8] += r3.y, acc[9] += r3.z, acc[10] += r3.w;
    sh_bwd(P.deg, mk(sd[0], sd[1], sd[2]), (cl & 1) != 0, (cl & 2) != 0, (cl & 4) != 0, dc, sh);
    return __float_as_uint(r0.x);
  }
  const float4 *row = P.rec_rows + kRowGlobal * r;
  const float4 r0 = row[0], r1 = row[1], r2 = row[2], r3 = row[3], r4 = row[4];
  const int cl = __float_as_int(r0.y);
  const float dc[3] = {r1.y, r1.z, r1.w};
  acc[0] += r2.x, acc[1] += r2.y, acc[2] += r2.z, acc[3] += r2.w;
  acc[4] += r3.x, acc[5] += r3.y, acc[6] += r3.z, acc[7] += r3.w;
  acc[8] += r4.x, acc[9] += r4.y, acc[10] += r4.z;
  sh_bwd(P.deg, mk(r0.z, r0.w, r1.x), (cl & 1) != 0, (cl & 2) != 0, (cl & 4) != 0, dc, sh);
  return __float_as_uint(r0.x);
}

// Stage row of one triangle's sums (64 floats): SH 0..47, vertices 48..56, opacity 57,
// sigma 58, triangle id 63 (-1: no triangle).
TT_D void stage_sums(float *row, int prim, int vrow, const float acc[11], const ShSink &sh) {
  for (int f = 0; f < 48; f++) row[f] = sh.v[f];
  for (int f = 0; f < 11; f++) row[48 + f] = acc[f];
  row[59] = __int_as_float(vrow);  // the triangle's first vertex row (cumsum)
  row[63] = __int_as_float(prim);
}

// Flush of kStageSlots staged rows by the whole block: lane = (row, field), so a wave's
// atomic instruction covers one triangle's consecutive SH / vertex fields.
template <int BLOCK>
TT_D void flush_stage(const TraceParams &P, const float *stage) {
  const int nsh3 = 3 * (P.deg + 1) * (P.deg + 1);
  for (int p = threadIdx.x; p < 64 * kStageSlots; p += BLOCK) {
    const int f = p & 63;
    const int prim = __float_as_int(stage[p | 63]);
    if (prim < 0 || f >= 59 || (f < 48 && f >= nsh3)) continue;
    const float v = stage[p];
    if (v == 0.0f) continue;
    // the field's address, picked per lane: one atomic instruction covers all of a row's fields
    float *dst = f < 48 ? P.dsh + (size_t)prim * P.M * 3 + f
                        : (f < 57 ? P.dverts + 3 * (size_t)__float_as_int(stage[(p & ~63) | 59]) + (f - 48)
                                  : (f == 57 ? P.dopac + prim : P.dsigma + prim));
    unsafeAtomicAdd(dst, v);
  }
}

// Flush of one triangle's sums to the global gradient buffers (FP32 atomics).
TT_D void flush_sums(const TraceParams &P, uint32_t prim, const float acc[11], const ShSink &sh) {
  float *gv = P.dverts + 3 * (size_t)P.cumsum[prim];
  for (int f = 0; f < 9; f++)
    if (acc[f] != 0.0f) unsafeAtomicAdd(gv + f, acc[f]);
  if (acc[F_OPAC] != 0.0f) unsafeAtomicAdd(P.dopac + prim, acc[F_OPAC]);
  if (acc[F_SIGMA] != 0.0f) unsafeAtomicAdd(P.dsigma + prim, acc[F_SIGMA]);
  float *gs = P.dsh + (size_t)prim * P.M * 3;
  const int nsh3 = 3 * (P.deg + 1) * (P.deg + 1);
  for (int f = 0; f < 48; f++)
    if (f < nsh3 && sh.v[f] != 0.0f) unsafeAtomicAdd(gs + f, sh.v[f]);
}

// backward.cu:191-350 for one traced segment: re-traverse it with the same chunk
// logic (or, given hl, read the forward's stored blended hits of the ray: hl[i * stride]
// for i < n_hl), recompute the forward state hit by hit and send each blended hit's
// gradient to the sink. fin = the segment's output colour (composite incl. background),
// g = the upstream gradient on it. MODE 0: a primary ray (depth / normal terms); 1: the
// bounce variant's primary (per-hit colours extended by rho L_r + (1 - rho) L_t, with the
// reflectivity gradient); 2: a bounce secondary ray (colour only). Returns the segment's
// final transmittance.
template <int K, int BLOCK, int MODE, bool WREC = false, bool LOCAL = true>
TT_D float backward_seg(const TraceParams &P, v3f o, v3f d, const float fin[3], const float g[3], float dL_ddepth,
                        v3f dL_dn2, float final_depth, v3f final_normal, const float *Lr, const float *Lt, float tlo,
                        long long plo, float far_t, int *s_node, int *s_rec,
                        const uint2 *hl = nullptr, int n_hl = 0, size_t stride = 0,
                        const float *hlc = nullptr) {
  float C[3] = {0.0f, 0.0f, 0.0f}, N[3] = {0.0f, 0.0f, 0.0f}, D = 0.0f, T = 1.0f;
  // one hit at or past the near plane; false: the ray terminates before blending it. rgb:
  // its colour from the forward's stored hits (clamped channel -0.0f), or null: evaluated here
  // With WREC the hit's inputs (window record, opacity, stored colour) come in `hi`, loaded
  // by the caller (fetch_hit) ahead of time; without, they are read here.
  struct HitIn {
    WinRec w;
    float opacity, rgb[3];
  };
  auto fetch_hit = [&](uint32_t prim, const float *rgb, HitIn &hi) {
    window_record_fetch(P.tri_win + kWinRec * (size_t)prim, hi.w);
    hi.opacity = P.opac[prim];
    if (rgb) hi.rgb[0] = rgb[0], hi.rgb[1] = rgb[1], hi.rgb[2] = rgb[2];
  };
  auto blend = [&](float t, uint32_t prim, const float *rgb, const HitIn &hi) -> bool {
    const int off = P.cumsum[prim];
    v3f V[3];
    Win w;
    float sigma, opacity;
    if (WREC) {
      TriWin tw;
      window_record_decode(hi.w, tw, V[0], V[1], V[2], sigma);
      window_hit(tw, V[0], V[1], V[2], o, d, t, sigma, w);
      if (MODE != 2) w.n = face_normal(V[0], V[1], V[2]);
      opacity = hi.opacity;
    } else {
      const float *vp = P.verts + 3 * (size_t)off;
      V[0] = ld3(vp), V[1] = ld3(vp + 3), V[2] = ld3(vp + 6);
      sigma = P.sigma[prim];
      window32(V[0], V[1], V[2], o, d, t, sigma, w);
      opacity = P.opac[prim];
    }
    const float alpha = fminf(0.99f, opacity * w.Cx);
    if (alpha < 1.0f / 255.0f) return true;
    const float testT = T * (1.0f - alpha);
    if (testT < P.cutoff) return false;
    const float bw = alpha * T;
    bool c0, c1, c2;
    v3f col;
    if (rgb) {
      const float r0 = WREC ? hi.rgb[0] : rgb[0], r1 = WREC ? hi.rgb[1] : rgb[1], r2 = WREC ? hi.rgb[2] : rgb[2];
      c0 = signbit(r0), c1 = signbit(r1), c2 = signbit(r2);
      col = mk(fabsf(r0), fabsf(r1), fabsf(r2));
    } else {
      col = sh_color(P.shs + (size_t)prim * P.M * 3, P.deg, d, c0, c1, c2);
    }
    float cc[3] = {col.x, col.y, col.z};
    if (MODE == 1) {
      const float rho = P.refl[prim];
      float dr = 0.0f;
      for (int c = 0; c < 3; c++) {
        dr += g[c] * (Lr[c] - Lt[c]);
        cc[c] += rho * Lr[c] + (1.0f - rho) * Lt[c];
      }
      unsafeAtomicAdd(P.drefl + prim, bw * dr);
    }
    C[0] += cc[0] * bw, C[1] += cc[1] * bw, C[2] += cc[2] * bw;
    if (MODE != 2) {
      N[0] += w.n.x * bw, N[1] += w.n.y * bw, N[2] += w.n.z * bw;
      D += bw * t;
    }
    float dL_dcolor[3];
    float dL_dalpha = 0.0f;
    const float inv1a = 1.0f / (1.0f - alpha);
    for (int c = 0; c < 3; c++) {
      dL_dcolor[c] = g[c] * bw;
      dL_dalpha += g[c] * (T * cc[c] - (fin[c] - C[c]) * inv1a);
    }
    if (MODE != 2) {
      dL_dalpha += dL_ddepth * (T * t - (final_depth - D) * inv1a);
      const v3f Nf = mk(N[0], N[1], N[2]);
      dL_dalpha += sumv(hmul(dL_dn2, T * w.n - (final_normal - Nf) * inv1a));
    }
    const float dL_dC = opacity * dL_dalpha;
    const float dL_dphi = ((sigma / w.phi_max) * dL_dC) * w.Cx;
    const v3f dL_dn3 = MODE != 2 ? dL_dn2 * bw : mk(0.0f, 0.0f, 0.0f);
    HitSink hs;
    for (int f = 0; f < 11; f++) hs.v[f] = 0.0f;
    hs.v[F_OPAC] = dL_dalpha * w.Cx;
    if (w.phi_final > 0.0f) hs.v[F_SIGMA] = (dL_dC * w.Cx) * log_tt(w.phi_final);
    window_bwd(V, w, dL_dphi, dL_dn3, hs);
    emit_record<BLOCK, LOCAL>(P, prim, off, d, dL_dcolor, c0, c1, c2, hs, s_rec);
    T = testT;
    return true;
  };
  if (hl) {
    if (WREC) {
      // software pipeline over the stored hits: while hit i is blended, the inputs of hit i + 1
      // and the entry of hit i + 2 are in flight (indices clamped to the last stored hit, so
      // every load reads a written entry; the surplus loads are never used)
      const int last = n_hl - 1;
      auto entry = [&](int j) { return hl[(size_t)j * stride]; };
      auto colours = [&](int j) { return hlc ? hlc + 3 * (size_t)j * stride : nullptr; };
      HitIn cur, nxt;
      uint2 hc = make_uint2(0u, 0u), hn = make_uint2(0u, 0u);
      if (n_hl > 0) {
        hc = entry(0);
        fetch_hit(hc.x, colours(0), cur);
        hn = entry(min(1, last));
      }
      for (int i = 0; i < n_hl; i++) {
        const int i1 = min(i + 1, last);
        fetch_hit(hn.x, colours(i1), nxt);
        const uint2 h2 = entry(min(i + 2, last));
        if (!blend(__uint_as_float(hc.y), hc.x, colours(i), cur)) break;
        cur = nxt;
        hc = hn;
        hn = h2;
      }
    } else {
      HitIn none;
      for (int i = 0; i < n_hl; i++) {
        const uint2 h = hl[(size_t)i * stride];
        if (!blend(__uint_as_float(h.y), h.x, hlc ? hlc + 3 * (size_t)i * stride : nullptr, none)) break;
      }
    }
    return T;
  }
  const RayPre r = prep_ray(o, d);
  KBuf<K> b;
  float last = 0.0f;
  TravCount tc;
  while (true) {
    gather<K, BLOCK, false>(P, r, tlo, plo, far_t, b, s_node, tc);
    float lt = 0.0f;
    uint32_t lp = 0;
    for (int i = 0; i < K; i++) {
      const float t = b.t(0);
      const uint32_t prim = b.p(0);
      kpop(b);
      if (prim == kInvalid) return T;
      lt = t;
      lp = prim;
      last = fmaxf((float)((double)last + kStepEps), t);
      if (t < P.near_t) continue;
      HitIn hi;  // fetched here: a fetch ahead of the blend spills at this kernel's register bound
      if (WREC) fetch_hit(prim, nullptr, hi);
      if (!blend(t, prim, nullptr, hi)) return T;
    }
    if (P.chunk_mode == 1) {
      tlo = lt;
      plo = (long long)lp;
    } else if (last > tlo) {
      tlo = last;
      plo = -1;
    } else {
      plo = (long long)lp;
    }
  }
}

// Environment-map gradient of one lookup: the upstream gradient g spread over the four
// bilinear taps of direction d.
TT_D void env_scatter(const TraceParams &P, v3f d, const float g[3]) {
  const EnvTaps e = env_taps(d, P.env_w, P.env_h);
  for (int j = 0; j < 4; j++)
    if (e.w[j] != 0.0f)
      for (int c = 0; c < 3; c++)
        if (g[c] != 0.0f) unsafeAtomicAdd(P.denv + 3 * e.idx[j] + c, e.w[j] * g[c]);
}

// The environment gradient of the bounce primaries of one wave (all 64 lanes must call it,
// `on` false for lanes without one): the four bilinear taps of each lane's direction d
// get w g, summed over the lanes sharing a texel before one atomic per texel and channel
// -- a wave of neighbouring primary rays looks up a handful of texels, so the per-lane
// form's atomics collided on the same addresses.
TT_D void env_scatter_wave(const TraceParams &P, v3f d, const float g[3], bool on) {
  EnvTaps e{};
  const bool any = on && (g[0] != 0.0f || g[1] != 0.0f || g[2] != 0.0f);
  if (any) e = env_taps(d, P.env_w, P.env_h);
  const int lane = threadIdx.x & 63;
  for (int j = 0; j < 4; j++) {
    const int idx = any && e.w[j] != 0.0f ? e.idx[j] : -1;
    const float v0 = any ? e.w[j] * g[0] : 0.0f, v1 = any ? e.w[j] * g[1] : 0.0f, v2 = any ? e.w[j] * g[2] : 0.0f;
    unsigned long long pending = __ballot(idx >= 0);
    while (pending) {  // one texel per round: its lanes' values summed across the wave
      const int leader = __ffsll((unsigned long long)pending) - 1;
      const int lidx = __shfl(idx, leader);
      const bool mine = idx == lidx;
      const unsigned long long peers = __ballot(mine);
      float s0 = mine ? v0 : 0.0f, s1 = mine ? v1 : 0.0f, s2 = mine ? v2 : 0.0f;
      for (int off = 32; off > 0; off >>= 1) {
        s0 += __shfl_xor(s0, off);
        s1 += __shfl_xor(s1, off);
        s2 += __shfl_xor(s2, off);
      }
      if (lane == leader) {
        if (s0 != 0.0f) unsafeAtomicAdd(P.denv + 3 * (size_t)lidx, s0);
        if (s1 != 0.0f) unsafeAtomicAdd(P.denv + 3 * (size_t)lidx + 1, s1);
        if (s2 != 0.0f) unsafeAtomicAdd(P.denv + 3 * (size_t)lidx + 2, s2);
      }
      pending &= ~peers;
    }
  }
}

// One pixel of the backward (backward.cu:191-391). VARIANT 4: the bounce variant's
// primary, its environment term and its two secondary rays (upstream R g and Q g);
// VARIANT 5, 6, 7: only its primary (with the environment term), only the reflected, only
// the refracted segment -- three launches whose sum is VARIANT 4, each with the register
// footprint of one segment; VARIANT 8: both secondary segments (records grouped globally).
template <int K, int BLOCK, int VARIANT, bool WREC = false>
TT_D void backward_ray(const TraceParams &P, int pix, int *s_node, int *s_rec, float *env_g = nullptr) {
  const size_t R = (size_t)P.R;
  v3f o, d;
  pixel_ray(P, (size_t)pix, o, d);
  const float final_depth = P.fothers[pix];
  const v3f final_normal = mk(P.fothers[2 * R + pix], P.fothers[3 * R + pix], P.fothers[4 * R + pix]);
  float final_color[3], dL_dpixel[3];
  for (int c = 0; c < 3; c++) {
    final_color[c] = P.fcolor[c * R + pix];
    dL_dpixel[c] = P.dpix[c * R + pix];
  }
  const float dL_ddepth = P.dothers[pix];
  const v3f dL_dn2 = mk(P.dothers[2 * R + pix], P.dothers[3 * R + pix], P.dothers[4 * R + pix]);
  if (VARIANT == 0) {
    float tlo = 0.0f, far_t = P.far_t;
    long long plo = -1;
    const uint2 *hl = nullptr;
    int n_hl = 0;
    if (P.hl_in) {  // the forward's stored blended hits (count -1: more than were stored)
      n_hl = P.hl_count_in[pix];
      if (n_hl >= 0) hl = P.hl_in + pix;
    }
    if (!hl && P.ival_bwd) {  // only the forward's blended hits: from the first one (inclusive) to the last
      const float4 iv = P.ival_bwd[pix];
      if (!(iv.x < kInfDist)) return;
      tlo = iv.x;
      plo = (long long)__float_as_uint(iv.w) - 1;
      far_t = iv.y;
    }
    const float *hlc = hl && P.hl_rgb_in ? P.hl_rgb_in + 3 * (size_t)pix : nullptr;
    backward_seg<K, BLOCK, 0, WREC>(P, o, d, final_color, dL_dpixel, dL_ddepth, dL_dn2, final_depth, final_normal, nullptr,
                              nullptr, tlo, plo, far_t, s_node, s_rec, hl, n_hl, P.hl_stride, hlc);
  } else {
    constexpr bool kPrim = VARIANT == 4 || VARIANT == 5;
    constexpr bool kRefl = VARIANT == 4 || VARIANT == 6 || VARIANT == 8;
    constexpr bool kRefr = VARIANT == 4 || VARIANT == 7 || VARIANT == 8;
    const float *a = P.aux + (size_t)kBounceAux * pix;
    float Lr[3], Lt[3];
    for (int c = 0; c < 3; c++) Lr[c] = a[c], Lt[c] = a[3 + c];
    const float Rw = a[6], Qw = a[7];
    const uint2 *hls[3] = {nullptr, nullptr, nullptr};  // the forward's hit lists of the three segments
    const float *hlcs[3] = {nullptr, nullptr, nullptr};  // and their colours
    int nhl[3] = {0, 0, 0};
    if (P.hl_in)
      for (int sg = 0; sg < 3; sg++) {
        if (!(sg == 0 ? kPrim : (sg == 1 ? kRefl : kRefr))) continue;
        nhl[sg] = P.hl_count_in[sg * P.hl_stride + pix];
        if (nhl[sg] >= 0) {
          const size_t e0 = (size_t)sg * P.hl_cap * P.hl_stride + pix;
          hls[sg] = P.hl_in + e0;
          if (P.hl_rgb_in) hlcs[sg] = P.hl_rgb_in + 3 * e0;
        }
      }
    if (kPrim) {
      const float Tp = backward_seg<K, BLOCK, 1, WREC>(P, o, d, final_color, dL_dpixel, dL_ddepth, dL_dn2, final_depth,
                                                 final_normal, Lr, Lt, 0.0f, -1, P.far_t, s_node, s_rec, hls[0],
                                                 nhl[0], P.hl_stride, hlcs[0]);
      const float ge[3] = {Tp * dL_dpixel[0], Tp * dL_dpixel[1], Tp * dL_dpixel[2]};
      if (env_g) {  // the caller aggregates it across the wave (env_scatter_wave)
        for (int c = 0; c < 3; c++) env_g[c] = ge[c];
      } else {
        env_scatter(P, d, ge);
      }
    }
    const v3f dr = mk(a[11], a[12], a[13]);
    if ((kRefl || kRefr) && (dr.x != 0.0f || dr.y != 0.0f || dr.z != 0.0f)) {
      const v3f zero = mk(0.0f, 0.0f, 0.0f);
      auto secondary = [&](auto which) {  // which: 0 reflected, 1 refracted (a compile-time constant)
        constexpr int s = decltype(which)::value;
        const float wgt = s == 0 ? Rw : Qw;
        const v3f so = s == 0 ? mk(a[8], a[9], a[10]) : mk(a[14], a[15], a[16]);
        const v3f sd = s == 0 ? dr : mk(a[17], a[18], a[19]);
        const float gs[3] = {wgt * dL_dpixel[0], wgt * dL_dpixel[1], wgt * dL_dpixel[2]};
        const float Ts = backward_seg<K, BLOCK, 2, WREC, VARIANT != 8>(P, so, sd, s == 0 ? Lr : Lt, gs, 0.0f, zero, 0.0f, zero, nullptr,
                                                   nullptr, bounce_start(P), -1, P.far_t, s_node, s_rec, hls[1 + s],
                                                   nhl[1 + s], P.hl_stride, hlcs[1 + s]);
        const float gse[3] = {Ts * gs[0], Ts * gs[1], Ts * gs[2]};
        env_scatter(P, sd, gse);
      };
      if (kRefl) secondary(std::integral_constant<int, 0>());
      if (kRefr) secondary(std::integral_constant<int, 1>());
    }
  }
}

// The record path of one block's backward (tile `tile`): the lanes trace their rays
// writing gradient records (no LDS atomics on the gradient fields); then, in batches of
// rec_batch records, the block groups its records by triangle in LDS (hash of the
// triangle id -> slot, counting sort by slot) and one thread per triangle sums its
// records -- rebuilding the SH gradient from each record's direction and dL/dcolour --
// and flushes the sums with one FP32 atomic per field. Every lane of the block must call
// this (block barriers; it ends with one, so the LDS can be reused by the caller). The
// sums of kStageSlots triangles at a time are staged in LDS and flushed by the whole block
// with one lane per (triangle, field), so that one atomic instruction covers a triangle's
// consecutive gradient fields. LDS: s_node (traversal stack and leaf queue, reused for the
// grouping), s_rec (kRecLds ints), s_tab (rec_tab_ints ints: table + stage).
template <int K, int BLOCK, int VARIANT, bool WREC = false>
TT_D void backward_block_rec(const TraceParams &P, bool valid, int pix, int tile, int *s_node, int *s_rec,
                             int *s_tab) {
  constexpr int kBatch = rec_batch<BLOCK>(), kAggSlots = agg_slots<BLOCK>();
  static_assert(kLdsInts * BLOCK * 4 >= 2 * kBatch * 2, "the grouping reuses the traversal LDS");
  static_assert(VARIANT != 4, "block-summed records hold one direction per lane: one segment kind per launch");
  // per slot: triangle id, record count (low 16 bits) and records listed so far (high 16 bits),
  // first list position and first work item (segment of <= kSegRecords records), 16 bits each
  int *t_keys = s_tab, *t_cf = s_tab + kAggSlots;
  short *t_off = reinterpret_cast<short *>(s_tab + 2 * kAggSlots), *t_item = t_off + kAggSlots;
  float *stage = reinterpret_cast<float *>(s_tab + 3 * kAggSlots);
  static_assert(kBatch < 32768 && kBatch / kSegRecords + kAggSlots < 32768, "16-bit list positions and items");
  if (threadIdx.x == 0) {
    const int lane_rows = P.rec_lane_rows > 0 ? P.rec_lane_rows : 16;
    int start = blockIdx.x * BLOCK * lane_rows, rows = BLOCK * lane_rows;
    if (P.rec_off) {
      start = P.rec_off[tile];
      rows = max(0, min(P.tile_hits[tile], P.rec_cap - start));
    }
    s_rec[0] = 0;
    s_rec[1] = start;
    s_rec[2] = rows;
  }
  for (int i = threadIdx.x; i < kAggSlots; i += BLOCK) t_keys[i] = -1, t_cf[i] = 0;
  __syncthreads();
  if (VARIANT == 5) {  // the bounce primaries: environment gradient aggregated per wave
    float eg[3] = {0.0f, 0.0f, 0.0f};
    if (valid) backward_ray<K, BLOCK, VARIANT, WREC>(P, pix, s_node, s_rec, eg);
    v3f po = mk(0.0f, 0.0f, 0.0f), pd = mk(0.0f, 0.0f, 1.0f);
    if (valid) pixel_ray(P, (size_t)pix, po, pd);
    env_scatter_wave(P, pd, eg, valid);
  } else if (valid) {
    backward_ray<K, BLOCK, VARIANT, WREC>(P, pix, s_node, s_rec);
  }
  __syncthreads();
  if (VARIANT == 8) return;  // records grouped globally (sorted by triangle) by the caller
  const int n_all = min(s_rec[0], s_rec[2]);
  const size_t start = (size_t)s_rec[1];
  short *slot_of = reinterpret_cast<short *>(s_node), *list = slot_of + kBatch;
  for (int b0 = 0; b0 < n_all; b0 += kBatch) {
    const int n = min(kBatch, n_all - b0);
    const size_t first = start + b0;
    if (b0 > 0) {
      for (int i = threadIdx.x; i < kAggSlots; i += BLOCK) t_keys[i] = -1, t_cf[i] = 0;
      __syncthreads();
    }
    for (int i = threadIdx.x; i < n; i += BLOCK) {
      const uint32_t prim =
          b0 == 0 ? (uint32_t)s_rec[kRecLds + i] : __float_as_uint(P.rec_rows[kRowLocal * (first + i)].x);
      const int sl = claim_slot<kAggSlots>(t_keys, prim);
      slot_of[i] = (short)sl;
      if (sl >= 0) atomicAdd(&t_cf[sl], 1);
    }
    __syncthreads();
    if (threadIdx.x < 64) {  // exclusive scan of the slot counts: kAggSlots / 64 per lane + a wave scan
      constexpr int per = kAggSlots / 64;
      const int l = threadIdx.x;
      int c[per], sum = 0, items = 0;
      for (int j = 0; j < per; j++) {
        c[j] = t_cf[per * l + j] & 0xffff;
        sum += c[j];
        items += (c[j] + kSegRecords - 1) / kSegRecords;
      }
      int inc = sum, inc_items = items;
      for (int off = 1; off < 64; off <<= 1) {
        const int v = __shfl_up(inc, off), w = __shfl_up(inc_items, off);
        if (l >= off) inc += v, inc_items += w;
      }
      int run = inc - sum, run_items = inc_items - items;
      for (int j = 0; j < per; j++) {
        t_off[per * l + j] = (short)run;
        t_item[per * l + j] = (short)run_items;
        run += c[j];
        run_items += (c[j] + kSegRecords - 1) / kSegRecords;
      }
      if (l == 63) s_rec[3] = inc_items;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += BLOCK) {
      const int sl = slot_of[i];
      if (sl >= 0) {
        list[t_off[sl] + (atomicAdd(&t_cf[sl], 0x10000) >> 16)] = (short)i;
      } else {  // no slot left: this record goes out on its own
        float acc[11] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        ShSink sh;
        for (int f = 0; f < 48; f++) sh.v[f] = 0.0f;
        const uint32_t prim = add_record<true>(P, first + i, acc, sh, rec_dirs<BLOCK>(s_rec));
        flush_sums(P, prim, acc, sh);
      }
    }
    __syncthreads();
    const int n_items = s_rec[3];
    for (int i0 = 0; i0 < n_items; i0 += BLOCK) {  // work item = (slot, segment of <= kSegRecords records)
      const int it = i0 + threadIdx.x;
      int prim = -1, vrow = 0;
      float acc[11] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
      ShSink sh;
      for (int f = 0; f < 48; f++) sh.v[f] = 0.0f;
      if (it < n_items) {
        int lo = 0, hi = kAggSlots;  // the last slot whose first item is <= it
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (t_item[mid] <= it) lo = mid;
          else hi = mid;
        }
        prim = t_keys[lo];
        vrow = P.cumsum[prim];
        const int b = t_off[lo] + (it - t_item[lo]) * kSegRecords,
                  e = min(t_off[lo] + (t_cf[lo] & 0xffff), b + kSegRecords);
#pragma unroll 2
        for (int j = b; j < e; j++) add_record<true>(P, first + list[j], acc, sh, rec_dirs<BLOCK>(s_rec));
      }
      const int live = min(BLOCK, n_items - i0);  // work items of this pass (the rest of the rows are empty)
      for (int q = 0; q * kStageSlots < live; q++) {
        if (threadIdx.x / kStageSlots == q) stage_sums(stage + 64 * (threadIdx.x % kStageSlots), prim, vrow, acc, sh);
        __syncthreads();
        flush_stage<BLOCK>(P, stage);
        __syncthreads();
      }
    }
  }
  __syncthreads();
}

// Sums of records sorted by triangle (keys; rows[j] is the record row of sorted entry j;
// entries with key 0xffffffff, rows never written, sort last). One wave per window of 64
// sorted entries (a grid-stride loop over the windows that ends at the first window of
// unwritten rows): each lane rebuilds one record's 59 gradient values into LDS, then for
// each run of one triangle in the window the lanes sum one field each down the run and
// flush it with one FP32 atomic instruction (lane = field), so a triangle costs one atomic
// per field per window it appears in.
constexpr int kRedRow = 61;  // floats per staged record (59 used), odd: rows start in distinct banks
__global__ __launch_bounds__(64) void k_rec_reduce_sorted(TraceParams P, const uint32_t *keys, const uint32_t *rows,
                                                          int n) {
  __shared__ float s_v[64 * kRedRow];
  const int lane = threadIdx.x;
  const int nsh3 = 3 * (P.deg + 1) * (P.deg + 1);
  const long long n_win = ((long long)n + 63) / 64;
  for (long long w = blockIdx.x; w < n_win; w += gridDim.x) {
    const long long j = w * 64 + lane;
    const uint32_t key = j < n ? keys[j] : 0xffffffffu;
    if (__shfl(key, 0) == 0xffffffffu) break;  // sorted: no written rows from here on
    float acc[11];
    ShSink sh;
    for (int f = 0; f < 11; f++) acc[f] = 0.0f;
    for (int f = 0; f < 48; f++) sh.v[f] = 0.0f;
    if (key != 0xffffffffu) add_record<false>(P, rows[j], acc, sh);
    float *mine = s_v + lane * kRedRow;
    for (int f = 0; f < 48; f++) mine[f] = sh.v[f];
    for (int f = 0; f < 11; f++) mine[48 + f] = acc[f];
    const uint32_t prev = __shfl_up(key, 1);
    unsigned long long heads = __ballot(lane == 0 || prev != key);
    __syncthreads();
    while (heads) {  // runs [s, e) of one key, in order
      const int s = __ffsll(heads) - 1;
      heads &= heads - 1;
      const int e = heads ? __ffsll(heads) - 1 : 64;
      const uint32_t prim = __shfl(key, s);
      if (prim == 0xffffffffu || lane >= 59) continue;
      float sum = 0.0f;
      for (int r = s; r < e; r++) sum += s_v[r * kRedRow + lane];
      if (sum == 0.0f) continue;
      if (lane < 48) {
        if (lane < nsh3) unsafeAtomicAdd(P.dsh + (size_t)prim * P.M * 3 + lane, sum);
      } else if (lane < 57) {
        unsafeAtomicAdd(P.dverts + 3 * (size_t)P.cumsum[prim] + (lane - 48), sum);
      } else if (lane == 57) {
        unsafeAtomicAdd(P.dopac + prim, sum);
      } else {
        unsafeAtomicAdd(P.dsigma + prim, sum);
      }
    }
    __syncthreads();
  }
}

__global__ void k_iota(uint32_t *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = (uint32_t)i;
}

__global__ void k_fill(float *p, size_t n, float v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_int(int *p, int n, int v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace
}  // namespace tt
