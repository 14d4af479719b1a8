// Internal declarations shared by the C-ABI layer and the kernel translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

namespace tt {

// LBVH in HBM.
//   nodes  : 4 x 16 B per internal node
//            [0] child0 box (lo.x, hi.x, lo.y, hi.y)   [1] child1 box (lo.x, hi.x, lo.y, hi.y)
//            [2] (c0.lo.z, c0.hi.z, c1.lo.z, c1.hi.z)   [3] int (child0, child1, parent, 0)
//            child >= 0: internal node index; child < 0: leaf ~s (sorted position s)
//   leaves : 3 x 16 B per leaf in Morton order: (A.xyz, prim bits) (B.xyz, 0) (C.xyz, 0)
struct Bvh {
  int P = -1;           // triangles in the structure (-1: none built)
  int n_leaves = 0;     // P, or 2 when P == 1 (second leaf is an unhittable dummy)
  int n_internal = 0;   // n_leaves - 1
  size_t cap_leaves = 0;
  float4 *nodes = nullptr;
  float4 *leaves = nullptr;
  int *leaf_parent = nullptr;
  int *flags = nullptr;
  uint32_t *keys[2] = {nullptr, nullptr};
  uint32_t *vals[2] = {nullptr, nullptr};
  uint32_t *hist = nullptr;     // radix digit counts [256][blocks]
  size_t cap_hist = 0;
  uint32_t *bounds = nullptr;   // ordered-uint centroid bounds [6]
  float4 *centroids = nullptr;  // [P]
  const int *cumsum_seen = nullptr;  // cumsum pointer of the last build (refit checks P only)
};

void bvh_free(Bvh &b);
// Builds (mode 0) or refits (mode 1) on `stream`. Returns an empty string or an error message.
std::string bvh_build(Bvh &b, const float *verts, int num_vertices, const int *cumsum, int P, int mode,
                      hipStream_t stream);

struct TraceParams {
  int W, H, R;
  int tiles_x, n_tiles;
  const float *ray_o, *ray_d;
  const float *verts;
  const int *npts, *cumsum;
  int P;
  const float *shs;
  int M, deg;
  const float *opac, *sigma;
  const float4 *nodes, *leaves;
  int n_leaves;
  int chunk_mode;
  float cutoff, near_t, far_t;
  // forward
  const float *bg;
  const float *extra;
  int extra_len;
  float *color, *others, *maxb, *density, *scaling;
  int *stats;
  int log_cap;
  int *log_prim;
  float *log_t;
  int *log_code;
  int *log_count;
  // backward
  const float *fcolor, *fothers, *dpix, *dothers;
  float *dverts, *dsigma, *dopac, *dsh;
};

// Launch the forward kernel for (k, variant). Returns an error message or "".
std::string launch_forward(const TraceParams &p, int k, int variant, hipStream_t stream);
std::string launch_backward(const TraceParams &p, int k, hipStream_t stream);
std::string launch_fill(float *ptr, size_t n, float v, hipStream_t stream);

}  // namespace tt
