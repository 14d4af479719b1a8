// Internal declarations shared by the C-ABI layer and the kernel translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace tt {

constexpr int kMaxLevels = 256;
constexpr int kEmptyChild = 0x7fffffff;  // bound on LBVH height bookkeeping (a Karras tree has at most 64)

// LBVH in HBM.
//   nodes  : 4 x 16 B per internal node
//            [0] child0 box (lo.x, hi.x, lo.y, hi.y)   [1] child1 box (lo.x, hi.x, lo.y, hi.y)
//            [2] (c0.lo.z, c0.hi.z, c1.lo.z, c1.hi.z)   [3] int (child0, child1, parent, 0)
//            child >= 0: internal node index; child < 0: leaf ~s (sorted position s)
//   leaves : 3 x 16 B per leaf in Morton order: (A.xyz, prim bits) (B.xyz, 0) (C.xyz, 0)
//   nodes4 : 5 x 16 B per binary node n, the traversal's 4-wide node: n's children, with
//            each internal child replaced by its own two children (2 to 4 entries), the
//            child boxes quantized to 256 steps over a per-axis origin and power-of-two
//            scale, each step count stored as an f16 integer
//            [0] origin xyz + exponents [1] x planes [2] y planes [3] z planes
//            [4] int child[4] (>= 0 binary node, < 0 leaf ~s, kEmptyChild unused)
//   counts4: per binary node n, the leaves under each of nodes4[n]'s children (16 bits each)
struct Bvh {
  int P = -1;           // triangles in the structure (-1: none built)
  int n_leaves = 0;     // P, or 2 when P == 1 (second leaf is an unhittable dummy)
  int n_internal = 0;   // n_leaves - 1
  size_t cap_leaves = 0;
  float4 *nodes = nullptr;
  float4 *nodes4 = nullptr;   // 4-wide view, 5 x 16 B per binary node n: the up-to-4 grandchildren of n,
                              // their boxes quantized to 256 steps (layout at k_collapse4)
  uint2 *counts4 = nullptr;   // leaves under each child of nodes4[n] (the tile probe's estimate)
  float4 *leaves = nullptr;
  int *leaf_parent = nullptr;
  unsigned long long *hword = nullptr;  // per node: the two children's reported heights
  int *height = nullptr;                // per internal node: 1 + max child height (leaf = 0)
  int *level_nodes = nullptr;           // internal nodes ordered by height
  uint32_t *level_meta = nullptr;       // device: histogram / offsets / cursors of the levels
  std::vector<int> level_off;           // host: level h occupies level_nodes[level_off[h], level_off[h+1])
  int height_max = 0;                   // root height = bound on the traversal stack depth
  uint32_t *keys[2] = {nullptr, nullptr};
  uint32_t *vals[2] = {nullptr, nullptr};
  uint32_t *hist = nullptr;     // radix digit counts [256][blocks]
  size_t cap_hist = 0;
  uint32_t *bounds = nullptr;   // ordered-uint centroid bounds [6]
  float4 *centroids = nullptr;  // [P]
  const int *cumsum_seen = nullptr;  // cumsum pointer of the last build (refit checks P only)
  double *sah_part = nullptr;        // per-block partial sums of the SAH cost
  size_t cap_sah = 0;
};

void bvh_free(Bvh &b);
// SAH cost of the current nodes (unit traversal and intersection costs): 1 + the sum over
// every non-root node of SA(node) / SA(root), written to out[0] on `stream`.
std::string bvh_cost(Bvh &b, float *out, hipStream_t stream);
// Builds (mode 0) or refits (mode 1) on `stream`. Returns an empty string or an error message.
std::string bvh_build(Bvh &b, const float *verts, int num_vertices, const int *cumsum, int P, int mode,
                      hipStream_t stream);

constexpr int kWinRec = 8;  // float4 per window record (128 B per triangle)

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
  // per-triangle window records (kWinRec float4 each) of this launch, or null; tri_win_ready:
  // they already hold this launch's inputs (the host skips the record pass)
  float4 *tri_win;
  int tri_win_ready;
  const float4 *nodes, *leaves, *nodes4;
  const uint2 *counts4;
  int n_leaves;
  int *ov_stack;  // traversal stack spill: ov_depth entries per launched lane
  int ov_depth;
  int chunk_mode;
  float cutoff, near_t, far_t;
  // forward
  const float *bg;
  const float *extra;
  int extra_len;
  float *color, *others, *maxb, *density, *scaling;
  int *stats;
  int *trav_stats;
  int log_cap;
  int *log_prim;
  float *log_t;
  int *log_code;
  int *log_count;
  // backward
  const float *fcolor, *fothers, *dpix, *dothers;
  float *dverts, *dsigma, *dopac, *dsh;
  // depth intervals: forward input (start, cap) and output (first, last, term, prim_first);
  // backward input = the forward's output
  const float2 *ival_in;
  void *ival_ready;  // host side: a hipEvent_t the traversal launch waits for (ival_in produced elsewhere)
  int ival_hard;  // 1: hits past the cap are dropped (the proposal is trusted); 0: exact continuation
  float ival_fb_start, ival_fb_end;  // hard mode fallbacks (distances), see trace_segment
  float4 *ival_out;
  const float4 *ival_bwd;
  // bounce variant: per-triangle reflectivity / IOR, environment map [env_h][env_w][3],
  // per-pixel record for the backward (kBounceAux floats), and their gradients
  const float *refl, *ior, *env;
  int env_w, env_h;
  float *aux;
  float *drefl, *denv;
  // tile schedule: the one-shot forward records each tile's wall time into tile_cost and
  // then orders the tiles by it (descending) into tile_order; a backward given tile_order
  // takes its tiles in that order (any permutation gives the same per-pixel results)
  float *tile_cost;
  int *tile_order;
  const int *tile_sched;
  // the forward's own schedule: tiles ordered by a BVH probe of each tile's centre ray
  // (estimated cost, descending); null: tiles in XCD band order
  int *probe_order;
  // backward gradient records (80-byte rows, one per blended hit), aggregated per block;
  // null: the per-block LDS gradient table instead. rec_cap = rows allocated.
  float4 *rec_rows;
  int rec_cap;
  // per-tile record regions: the forward counts each tile's blended hits into tile_hits
  // (when non-null) and the exclusive prefix sums go to rec_off; a backward given rec_off
  // gives tile t the rows [rec_off[t], rec_off[t] + tile_hits[t])
  int *tile_hits;
  const int *rec_off;
  int *rec_off_out;
  // the backward tiles 16x16 (two forward tiles): the forward folds its tiles' blended hits
  // and wall times into these (its rec_off_out and tile_order are then over backward tiles)
  int *tile_hits_b;
  float *tile_cost_b;
  // stored blended hits of the primary rays (variant 0, one-shot launches): the forward
  // writes hit i of pixel p as (prim, t bits) to hl[i * R + p] for i < hl_cap and the
  // count to hl_count[p] (-1: more than hl_cap); a backward given hl_in reads them
  // instead of re-traversing (pixels with count -1 are re-traversed)
  uint2 *hl;
  int *hl_count;
  int hl_cap;
  const uint2 *hl_in;
  float *hl_rgb;          // optional: colours of the stored hits, [entry][3] (clamped channel: -0.0f)
  const float *hl_rgb_in;
  const int *hl_count_in;
  size_t hl_stride;  // entries between hit i and hit i + 1 of a pixel (the launch's total rays)
  int rec_lane_rows;  // rows per launched lane of a fixed record region (no per-tile regions)
  // fused ray generation: this view's pinhole camera (14 doubles in device memory: eye,
  // forward, right, up, tan_x, tan_y), or null (rays read from ray_o / ray_d); cams = the
  // camera table [n][14] of a multi-camera launch
  const double *cam, *cams;
  // fused loss of a forward: target image [3][R] of this view; dL/dcolour is written to
  // loss_grad [3][R] and each tile's loss sum to loss_part[tile] (global tile index)
  const float *target;
  float *loss_grad, *loss_part;
  int loss_kind;    // 1: mean of (c - t)^2, 2: mean of |c - t| over the batch's channels and pixels
  float loss_norm;  // 1 / (3 x the batch's rays)
  // global grouping of records (the bounce variant's secondary segments, whose rays are too
  // incoherent for per-block grouping): the triangle of every record row (rows never
  // written hold 0xffffffff) and the radix-sort scratch (rec_cap entries each, hist:
  // 256 per 4096 rows + one per 4096 of those + 1); null: per-block grouping only
  uint32_t *rec_keys;
  uint32_t *sort_keys2, *sort_vals, *sort_vals2, *sort_hist;
  // the bounce forward's secondary rays binned for coherence: one key per pixel (pixel
  // region, reflected-direction bin, pixel; 0xffffffff: no secondary rays), sorted with the
  // sort_* scratch (R entries each); null: all three segments in one launch
  uint32_t *ray_keys;
};
// Per-pixel record of the bounce variant: L_r[3] L_t[3] R Q o_r[3] d_r[3] o_t[3] d_t[3]
// (d_r = 0: no secondary rays).
constexpr int kBounceAux = 20;

// Cameras of one persistent multi-camera launch (passed by value as a kernel argument).
constexpr int kMaxCams = 32;
struct MultiCams {
  int n, total_tiles;
  int W[kMaxCams], H[kMaxCams], tiles_x[kMaxCams];
  int tile_begin[kMaxCams + 1];  // camera c owns global tiles [tile_begin[c], tile_begin[c+1])
  long long ray_off[kMaxCams];   // first ray of camera c in the concatenated ray / pixel buffers
  int q_begin[9];                // 8 work queues: queue q serves global tiles [q_begin[q], q_begin[q+1])
};
// Tile shape of a persistent launch with `block` lanes per block (64 -> 8x8, 128 -> 16x8, 256 -> 16x16).
void multi_tile_shape(int block, int &tw, int &th);
// Blocks that fill the device for the persistent kernel (k, block, backward?), 0 if unsupported.
int multi_grid(int k, int block, bool backward, bool wrec = true);
// Persistent launch over all cameras; `heads` = 8 device ints (reset inside, graph-safe).
// A forward with a fused loss (p.loss_part) writes the batch's loss to *loss_out.
std::string launch_multi(const TraceParams &p, const MultiCams &mc, int *heads, int k, int block, int tiles_per_fetch,
                         int grid, bool backward, hipStream_t stream, float *loss_out = nullptr);
std::string launch_fill_int(int *ptr, int n, int v, hipStream_t stream);

// Message returned by tt_last_error() on this thread (shared by every entry point).
void set_last_error(const std::string &msg);

// Tiles (blocks) of a one-shot trace launch for an image of W x H, and of the backward
// (whose tiles are two forward tiles stacked).
int trace_tiles(int W, int H);
int backward_tiles(int W, int H);
// Tiles ordered by cost, most expensive first (one block), into order[n].
std::string launch_tile_order(const float *cost, int n, int *order, hipStream_t stream);
// The window records of p.tri_win from p's vertices and sigma (one per triangle).
std::string launch_tri_win(const TraceParams &p, hipStream_t stream);
// Lanes a trace launch uses for an image of W x H (the stack spill needs one slot each).
size_t trace_lanes(int W, int H);
constexpr int kLdsStack = 16;  // traversal stack entries per lane kept in LDS; deeper ones spill

// Optional HIP events recorded on the launch stream right before and after the traversal
// kernel(s) of a launch (not the schedule kernels around them).
struct KernelEvents {
  hipEvent_t before = nullptr, after = nullptr;
};

// Launch the forward kernel for (k, variant). Returns an error message or "".
std::string launch_forward(const TraceParams &p, int k, int variant, hipStream_t stream,
                           const KernelEvents *ev = nullptr);
std::string launch_backward(const TraceParams &p, int k, int variant, hipStream_t stream,
                            const KernelEvents *ev = nullptr);
std::string launch_fill(float *ptr, size_t n, float v, hipStream_t stream);

}  // namespace tt
