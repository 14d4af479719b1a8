// C ABI of libtritracer_hip.so (declared in include/tritracer_hip.h).
//
// Thin layer: validates arguments, owns the LBVH buffers, launches the kernels
// on the caller's stream, and reports failures as status codes plus a
// thread-local message (the reference throws Exception from CUDA_CHECK /
// OPTIX_CHECK, tritracer/exception.h:32-51; its C++ caller does the same from
// these codes).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/tritracer_hip.h"
#include "tt_internal.h"

struct tt_tracer {
  int device = 0;
  tt::Bvh bvh;
  int *ov_stack = nullptr;  // traversal stack spill (entries below the LDS part)
  size_t ov_cap = 0;
  float4 *tri_win = nullptr;  // window records of the traversal launches (128 B per triangle)
  size_t tri_win_cap = 0;
  // the inputs the last record pass was issued for (a backward told the scene is unchanged
  // reuses the records when its inputs are these)
  const void *win_verts = nullptr, *win_cumsum = nullptr, *win_sigma = nullptr;
  int win_P = -1;
  int *heads = nullptr;                    // 8 work-queue heads of the persistent launches
  std::map<long long, int> grids;          // persistent grid size per (k, block, backward)
  hipStream_t capture_stream = nullptr;    // private stream used only to record graphs
  hipGraphExec_t graph = nullptr;          // the recorded multi-camera step
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool build_timed = false;
  // tile schedule measured by the last one-shot forward (per-tile wall time -> order),
  // used by a backward over an image of the same size
  float *tile_cost = nullptr;
  int *tile_order = nullptr;
  int *probe_order = nullptr;
  int *tile_hits = nullptr;  // blended hits per tile in the last forward
  int *tile_hits_b = nullptr;    // the same folded into the backward's (16x16) tiles
  float *tile_cost_b = nullptr;  // wall times folded likewise (the backward's order)
  int *rec_off = nullptr;    // their exclusive prefix sums: the backward's record regions
  int tile_cap = 0;
  int sched_W = -1, sched_H = -1;
  // the last multi-camera forward's tile count, block and rays (its schedule is valid for a
  // backward of the same batch)
  int multi_tiles = -1, multi_block = -1;
  long long multi_rays = -1;
  // backward gradient records: 2048 rows of 80 B per block of the one-shot backward
  float4 *rec_rows = nullptr;
  size_t rec_blocks = 0;  // rows allocated
  // blended hits of the last one-shot forwards (their record-region total), copied to pinned
  // host memory after each forward; read only once its event has completed, so the pool
  // grows to the largest total seen without making a backward wait for its forward
  int *hits_total_host = nullptr;
  hipEvent_t hits_total_ev = nullptr;
  bool hits_total_pending = false;
  size_t hits_high_water = 0;
  // global grouping of the bounce variant's secondary-segment records: per record row a key
  // and the radix-sort scratch (keys, values x 2, digit counts)
  uint32_t *rec_keys = nullptr, *sort_keys2 = nullptr, *sort_vals = nullptr, *sort_vals2 = nullptr;
  uint32_t *sort_hist = nullptr;
  size_t sort_cap = 0;
  // State generations. A captured graph holds the pointers, stack depth and tile schedule of
  // the moment it was recorded: mem_gen changes with every TT_BUILD (new nodes, possibly a
  // deeper tree) and every (re)allocation of the tracer's scratch; sched_gen with every eager
  // launch that rewrites the tile schedule for a different tile layout. A replay checks both
  // (the schedule only for a backward-only graph, whose order comes from outside the graph).
  unsigned long long mem_gen = 1, sched_gen = 1;
  unsigned long long graph_mem_gen = 0, graph_sched_gen = 0;
  int graph_what = 0;
  // LBVH quality (tt_build TT_REFIT_AUTO): SAH costs on the device ([0] right after the last
  // TT_BUILD, [1] the last one evaluated) with a pinned host copy, the policy and counters
  float *sah_dev = nullptr, *sah_host = nullptr;
  float rebuild_ratio = 1.25f;
  int n_builds = 0, n_refits = 0, n_auto_rebuilds = 0;
  float *loss_part = nullptr;  // per-tile sums of the fused loss of a multi-camera forward
  int loss_cap = 0;
  int noalloc_grid[3][3] = {};  // persistent forward grid per (k, block) of tt_multi_forward_noalloc (0: unset)
  // per-kernel timing (tt_set_kernel_timing): a ring of event pairs per direction (0 forward,
  // 1 backward) around the traversal kernels, and the number of timed launches
  bool timing = false;
  hipEvent_t tev[2][TT_TIMING_RING][2] = {};
  unsigned long long tcount[2] = {0, 0};
};

// Scratch of tt_multi_forward_noalloc: one static device array that every GPU the library
// runs on holds (its own copy, at its own address) until the process ends. In host code the
// name `scratch` is a host-side stand-in, not a device address: kernels and fills get the
// address hipGetSymbolAddress returned for the current device, kept in g_scratch_dev.
// Layout (ints): [0, 64) work-queue heads, then TT_NOALLOC_MAX_TILES tile costs (float),
// TT_NOALLOC_MAX_TILES tile order entries, TT_NOALLOC_MAX_TILES per-tile loss sums (float, the
// fused loss), then the traversal stack spill.
static __device__ __align__(256) int scratch[TT_NOALLOC_SCRATCH_BYTES / sizeof(int)];

namespace {
constexpr int kScratchDevices = 64;
constexpr size_t kScratchInts = TT_NOALLOC_SCRATCH_BYTES / sizeof(int);
constexpr size_t kScratchCostOff = 64, kScratchOrderOff = kScratchCostOff + TT_NOALLOC_MAX_TILES;
constexpr size_t kScratchLossOff = kScratchOrderOff + TT_NOALLOC_MAX_TILES;
constexpr size_t kScratchSpillOff = kScratchLossOff + TT_NOALLOC_MAX_TILES;
static_assert(kScratchSpillOff % 64 == 0, "spill area must stay 256-byte aligned");
std::mutex g_scratch_mu;
int *g_scratch_dev[kScratchDevices];  // device address of `scratch` on device d (null: not set up)

// Once per device, with that device current (called from tt_create).
hipError_t setup_scratch(int dev) {
  if (dev < 0 || dev >= kScratchDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_dev[dev]) return hipSuccess;
  void *ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch_dev[dev] = static_cast<int *>(ptr);
  return hipSuccess;
}
}  // namespace

namespace {
thread_local std::string g_err;
}  // namespace

void tt::set_last_error(const std::string &msg) {
  g_err = msg;
}

namespace {
int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}
int ok() {
  g_err.clear();
  return TT_OK;
}

#define TT_ARG(cond, msg) \
  do {                    \
    if (!(cond)) return fail(TT_ERR_INVALID_ARG, msg); \
  } while (0)

int check_common(const tt_tracer *tr, int W, int H, const float *ro, const float *rd, const float *verts,
                 const int *npts, const int *cumsum, int P, const float *shs, int M, int deg, const float *opac,
                 const float *sigma, int k, int chunk_mode) {
  TT_ARG(tr, "tracer is NULL");
  TT_ARG(W >= 0 && H >= 0, "image size must be non-negative");
  TT_ARG((long long)W * H < (1ll << 31), "too many rays for one launch");
  TT_ARG(P >= 0, "number of triangles must be non-negative");
  TT_ARG(k == 8 || k == 16 || k == 32, "k (hits per chunk) must be 8, 16 or 32");
  TT_ARG(chunk_mode == TT_CHUNK_COMPAT || chunk_mode == TT_CHUNK_EXACT, "unknown chunk mode");
  if ((long long)W * H > 0) TT_ARG(ro && rd, "ray buffers are NULL");
  if (P > 0) {
    TT_ARG(verts && npts && cumsum && opac && sigma, "scene buffers are NULL");
    TT_ARG(deg >= 0 && deg <= 3, "sh_degree must be in [0, 3]");
    TT_ARG(shs != nullptr && M >= (deg + 1) * (deg + 1),
           "SH coefficients are required (precomputed colours are not supported by the tracer, as in the "
           "reference where params.cols is never set): shs must be [P, M, 3] with M >= (sh_degree+1)^2");
    if (tr->bvh.P < 0) return fail(TT_ERR_STATE, "no acceleration structure: build it first (rebuild_gas = 0)");
    if (tr->bvh.P != P)
      return fail(TT_ERR_STATE, "the acceleration structure holds " + std::to_string(tr->bvh.P) +
                                    " triangles but " + std::to_string(P) + " were passed; rebuild it");
  }
  return TT_OK;
}

// Stack spill for a launch of `lanes` lanes: a 4-wide walk defers at most 3 children at
// each of at most `height` 4-wide nodes on its path (each 4-wide step descends at least one
// binary level); entries beyond the LDS part spill.
int prepare_spill_lanes(tt_tracer *tr, size_t lanes, tt::TraceParams &p) {
  const int depth = 3 * tr->bvh.height_max + 1 - tt::kLdsStack;
  p.ov_depth = depth > 0 ? depth : 0;
  p.ov_stack = tr->ov_stack;
  if (p.ov_depth == 0 || tr->bvh.n_leaves == 0) return TT_OK;
  const size_t need = lanes * (size_t)p.ov_depth;
  if (need > tr->ov_cap) {
    if (tr->ov_stack) (void)hipFree(tr->ov_stack);
    tr->ov_stack = nullptr;
    tr->ov_cap = 0;
    hipError_t e = hipMalloc(&tr->ov_stack, need * sizeof(int));
    if (e != hipSuccess) return fail(TT_ERR_HIP, std::string("traversal stack spill allocation failed: ") +
                                                     hipGetErrorString(e));
    tr->ov_cap = need;
    tr->mem_gen++;
  }
  p.ov_stack = tr->ov_stack;
  return TT_OK;
}
int prepare_spill(tt_tracer *tr, int W, int H, tt::TraceParams &p) {
  return prepare_spill_lanes(tr, tt::trace_lanes(W, H), p);
}

// Window records of a traversal launch: kWinRec float4 per triangle, formed by a record pass
// before the launch from the vertices and sigma it is given -- or, with `reuse`, taken from the
// last pass when that pass was issued for the same three input pointers and P (the caller
// vouches that their contents did not change).
int prepare_tri_win(tt_tracer *tr, tt::TraceParams &p, bool reuse = false) {
  const size_t n = (size_t)(p.P > 0 ? p.P : 0);
  if (n > tr->tri_win_cap) {
    if (tr->tri_win) (void)hipFree(tr->tri_win);
    tr->tri_win = nullptr;
    tr->tri_win_cap = 0;
    tr->win_P = -1;
    hipError_t e = hipMalloc(&tr->tri_win, n * tt::kWinRec * sizeof(float4));
    if (e != hipSuccess) return fail(TT_ERR_HIP, std::string("window record allocation failed: ") +
                                                     hipGetErrorString(e));
    tr->tri_win_cap = n;
    tr->mem_gen++;
  }
  p.tri_win = tr->tri_win;
  p.tri_win_ready = reuse && tr->win_P == p.P && tr->win_verts == p.verts && tr->win_cumsum == p.cumsum &&
                    tr->win_sigma == p.sigma;
  tr->win_verts = p.verts;
  tr->win_cumsum = p.cumsum;
  tr->win_sigma = p.sigma;
  tr->win_P = p.P;
  return TT_OK;
}

void free_rec_sort(tt_tracer *tr) {
  for (uint32_t **q : {&tr->rec_keys, &tr->sort_keys2, &tr->sort_vals, &tr->sort_vals2, &tr->sort_hist}) {
    if (*q) (void)hipFree(*q);
    *q = nullptr;
  }
  tr->sort_cap = 0;
}

// Key and sort buffers for n entries: the bounce backward's global grouping of its p.rec_cap
// record rows, the bounce forward's binning of its R secondary-ray pixels (shared: the two
// never run at once on a tracer's stream).
int prepare_rec_sort(tt_tracer *tr, tt::TraceParams &p, size_t n) {
  if (n > tr->sort_cap) {
    free_rec_sort(tr);
    const size_t counts = 256 * ((n + 4095) / 4096), hist = counts + (counts + 4095) / 4096 + 1;
    if (hipMalloc(&tr->rec_keys, n * 4) != hipSuccess || hipMalloc(&tr->sort_keys2, n * 4) != hipSuccess ||
        hipMalloc(&tr->sort_vals, n * 4) != hipSuccess || hipMalloc(&tr->sort_vals2, n * 4) != hipSuccess ||
        hipMalloc(&tr->sort_hist, hist * 4) != hipSuccess) {
      free_rec_sort(tr);
      return fail(TT_ERR_HIP, "hipMalloc of the record sort buffers failed");
    }
    tr->sort_cap = n;
    tr->mem_gen++;
  }
  p.rec_keys = tr->rec_keys;
  p.sort_keys2 = tr->sort_keys2;
  p.sort_vals = tr->sort_vals;
  p.sort_vals2 = tr->sort_vals2;
  p.sort_hist = tr->sort_hist;
  return TT_OK;
}

void free_records(tt_tracer *tr) {
  if (tr->rec_rows) (void)hipFree(tr->rec_rows);
  tr->rec_rows = nullptr;
  tr->rec_blocks = 0;
}

// Gradient-record rows for a backward: 16 rows of 80 B per launched lane (2.6 GB for a
// 1920x1080 one-shot backward).
int prepare_records(tt_tracer *tr, size_t lanes, tt::TraceParams &p, int lane_rows = 16) {
  const size_t rows = lanes * (size_t)lane_rows;
  if (rows > (size_t)0x7fffffff) return fail(TT_ERR_INVALID_ARG, "too many rays for one backward launch");
  if (rows > tr->rec_blocks) {
    free_records(tr);
    if (hipMalloc(&tr->rec_rows, rows * 5 * sizeof(float4)) != hipSuccess) {
      tr->rec_rows = nullptr;
      return fail(TT_ERR_HIP, "hipMalloc of the gradient records failed");
    }
    tr->rec_blocks = rows;
    tr->mem_gen++;
  }
  p.rec_rows = tr->rec_rows;
  p.rec_cap = (int)rows;
  p.rec_lane_rows = lane_rows;
  return TT_OK;
}

// Tile schedule buffers (cost, orders, blended hits, record offsets) for nt tiles.
int ensure_tiles(tt_tracer *tr, int nt) {
  if (nt <= tr->tile_cap) return TT_OK;
  tr->mem_gen++;
  if (tr->tile_cost) (void)hipFree(tr->tile_cost);
  if (tr->tile_order) (void)hipFree(tr->tile_order);
  if (tr->probe_order) (void)hipFree(tr->probe_order);
  if (tr->tile_hits) (void)hipFree(tr->tile_hits);
  if (tr->rec_off) (void)hipFree(tr->rec_off);
  if (tr->tile_hits_b) (void)hipFree(tr->tile_hits_b);
  if (tr->tile_cost_b) (void)hipFree(tr->tile_cost_b);
  tr->tile_cost = nullptr;
  tr->tile_order = nullptr;
  tr->probe_order = nullptr;
  tr->tile_hits = nullptr;
  tr->rec_off = nullptr;
  tr->tile_hits_b = nullptr;
  tr->tile_cost_b = nullptr;
  tr->tile_cap = 0;
  tr->sched_W = tr->sched_H = -1;
  tr->multi_tiles = -1;
  if (hipMalloc(&tr->tile_cost, (size_t)nt * sizeof(float)) != hipSuccess ||
      hipMalloc(&tr->tile_order, (size_t)nt * sizeof(int)) != hipSuccess ||
      hipMalloc(&tr->probe_order, (size_t)nt * sizeof(int)) != hipSuccess ||
      hipMalloc(&tr->tile_hits, (size_t)nt * sizeof(int)) != hipSuccess ||
      hipMalloc(&tr->rec_off, ((size_t)nt + 1) * sizeof(int)) != hipSuccess ||
      hipMalloc(&tr->tile_hits_b, (size_t)nt * sizeof(int)) != hipSuccess ||
      hipMalloc(&tr->tile_cost_b, (size_t)nt * sizeof(float)) != hipSuccess)
    return fail(TT_ERR_HIP, "hipMalloc of the tile schedule failed");
  tr->tile_cap = nt;
  return TT_OK;
}

// The record-region total of an earlier forward, if its copy has landed (never waits).
void poll_hits_total(tt_tracer *tr) {
  if (!tr->hits_total_pending || hipEventQuery(tr->hits_total_ev) != hipSuccess) return;
  tr->hits_total_pending = false;
  if (*tr->hits_total_host > 0) tr->hits_high_water = std::max(tr->hits_high_water, (size_t)*tr->hits_total_host);
}

void fill_common(tt::TraceParams &p, const tt_tracer *tr, int W, int H, const float *ro, const float *rd,
                 const float *verts, const int *npts, const int *cumsum, int P, const float *shs, int M, int deg,
                 const float *opac, const float *sigma, int chunk_mode, float cutoff, float near_t, float far_t) {
  p = tt::TraceParams{};
  p.W = W;
  p.H = H;
  p.R = W * H;
  p.ray_o = ro;
  p.ray_d = rd;
  p.verts = verts;
  p.npts = npts;
  p.cumsum = cumsum;
  p.P = P;
  p.shs = shs;
  p.M = M;
  p.deg = deg;
  p.opac = opac;
  p.sigma = sigma;
  p.nodes = tr->bvh.nodes;
  p.nodes4 = tr->bvh.nodes4;
  p.counts4 = tr->bvh.counts4;
  p.leaves = tr->bvh.leaves;
  p.n_leaves = P > 0 ? tr->bvh.n_leaves : 0;
  p.chunk_mode = chunk_mode;
  p.cutoff = cutoff;
  p.near_t = near_t;
  p.far_t = far_t;
}
}  // namespace

extern "C" {

const char *tt_last_error(void) {
  return g_err.c_str();
}

const char *tt_version(void) {
  return "tritracer_hip 0.1 (gfx950)";
}

int tt_create(tt_tracer **out) {
  if (!out) return fail(TT_ERR_INVALID_ARG, "out is NULL");
  tt_tracer *tr = new tt_tracer();
  hipError_t e = hipGetDevice(&tr->device);
  if (e != hipSuccess) {
    delete tr;
    return fail(TT_ERR_HIP, std::string("hipGetDevice failed: ") + hipGetErrorString(e));
  }
  e = setup_scratch(tr->device);
  if (e != hipSuccess && e != hipErrorInvalidDevice) {  // beyond kScratchDevices: the no-alloc path reports it
    delete tr;
    return fail(TT_ERR_HIP, std::string("hipGetSymbolAddress of the static scratch failed: ") + hipGetErrorString(e));
  }
  *out = tr;
  return ok();
}

int tt_destroy(tt_tracer *tr) {
  if (!tr) return ok();
  int dev = 0;
  (void)hipGetDevice(&dev);
  (void)hipSetDevice(tr->device);
  tt::bvh_free(tr->bvh);
  if (tr->ov_stack) (void)hipFree(tr->ov_stack);
  if (tr->tri_win) (void)hipFree(tr->tri_win);
  if (tr->heads) (void)hipFree(tr->heads);
  if (tr->graph) (void)hipGraphExecDestroy(tr->graph);
  if (tr->capture_stream) (void)hipStreamDestroy(tr->capture_stream);
  if (tr->ev0) (void)hipEventDestroy(tr->ev0);
  if (tr->ev1) (void)hipEventDestroy(tr->ev1);
  for (auto &dir : tr->tev)
    for (auto &pair : dir)
      for (hipEvent_t &ev : pair)
        if (ev) (void)hipEventDestroy(ev);
  if (tr->tile_cost) (void)hipFree(tr->tile_cost);
  if (tr->tile_order) (void)hipFree(tr->tile_order);
  if (tr->probe_order) (void)hipFree(tr->probe_order);
  if (tr->tile_hits) (void)hipFree(tr->tile_hits);
  if (tr->rec_off) (void)hipFree(tr->rec_off);
  if (tr->tile_hits_b) (void)hipFree(tr->tile_hits_b);
  if (tr->tile_cost_b) (void)hipFree(tr->tile_cost_b);
  free_records(tr);
  free_rec_sort(tr);
  if (tr->loss_part) (void)hipFree(tr->loss_part);
  if (tr->sah_dev) (void)hipFree(tr->sah_dev);
  if (tr->sah_host) (void)hipHostFree(tr->sah_host);
  if (tr->hits_total_host) (void)hipHostFree(tr->hits_total_host);
  if (tr->hits_total_ev) (void)hipEventDestroy(tr->hits_total_ev);
  (void)hipSetDevice(dev);
  delete tr;
  return ok();
}

int tt_build(tt_tracer *tr, const float *vertices_dev, int num_vertices, const int *cumsum_dev, int P, int mode,
             void *stream) {
  TT_ARG(tr, "tracer is NULL");
  TT_ARG(mode == TT_BUILD || mode == TT_REFIT || mode == TT_REUSE || mode == TT_REFIT_AUTO,
         "mode must be 0 (build), 1 (refit), 2 (reuse) or 3 (refit, rebuild when degraded)");
  TT_ARG(P >= 0 && num_vertices >= 0, "sizes must be non-negative");
  if (mode == TT_REUSE) {
    if (tr->bvh.P < 0) return fail(TT_ERR_STATE, "rebuild_gas = 2 (reuse) but no acceleration structure exists");
    return ok();
  }
  if (P > 0) TT_ARG(vertices_dev && cumsum_dev, "vertex / cumsum buffers are NULL");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (!tr->ev0) {
    if (hipEventCreate(&tr->ev0) != hipSuccess || hipEventCreate(&tr->ev1) != hipSuccess)
      return fail(TT_ERR_HIP, "hipEventCreate failed");
  }
  if (!tr->sah_dev) {
    if (hipMalloc(&tr->sah_dev, 2 * sizeof(float)) != hipSuccess ||
        hipHostMalloc((void **)&tr->sah_host, 2 * sizeof(float)) != hipSuccess)
      return fail(TT_ERR_HIP, "allocation of the SAH cost buffers failed");
  }
  (void)hipEventRecord(tr->ev0, s);
  const bool automatic = mode == TT_REFIT_AUTO;
  if (mode == TT_BUILD) tr->mem_gen++;  // new nodes and tree height: captured graphs are stale
  std::string err = tt::bvh_build(tr->bvh, vertices_dev, num_vertices, cumsum_dev, P, automatic ? TT_REFIT : mode, s);
  if (!err.empty()) {
    tr->build_timed = false;
    return fail(mode != TT_BUILD && err.find("refit") != std::string::npos ? TT_ERR_STATE : TT_ERR_HIP, err);
  }
  if (mode == TT_BUILD) {
    tr->n_builds++;
    err = tt::bvh_cost(tr->bvh, tr->sah_dev, s);
    if (err.empty() && hipMemcpyAsync(tr->sah_dev + 1, tr->sah_dev, sizeof(float), hipMemcpyDeviceToDevice, s) != hipSuccess)
      err = "SAH cost copy failed";
  } else {
    tr->n_refits++;
  }
  if (automatic && err.empty()) {  // the refit tree's cost against the cost right after the last build
    err = tt::bvh_cost(tr->bvh, tr->sah_dev + 1, s);
    if (err.empty() && (hipMemcpyAsync(tr->sah_host, tr->sah_dev, 2 * sizeof(float), hipMemcpyDeviceToHost, s) !=
                            hipSuccess || hipStreamSynchronize(s) != hipSuccess))
      err = "SAH cost read-back failed";
    if (err.empty() && tr->rebuild_ratio > 0.0f && tr->sah_host[1] > tr->rebuild_ratio * tr->sah_host[0]) {
      tr->mem_gen++;
      tr->n_auto_rebuilds++;
      tr->n_builds++;
      err = tt::bvh_build(tr->bvh, vertices_dev, num_vertices, cumsum_dev, P, TT_BUILD, s);
      if (err.empty()) err = tt::bvh_cost(tr->bvh, tr->sah_dev, s);
      if (err.empty() && hipMemcpyAsync(tr->sah_dev + 1, tr->sah_dev, sizeof(float), hipMemcpyDeviceToDevice, s) !=
                             hipSuccess)
        err = "SAH cost copy failed";
    }
  }
  if (!err.empty()) {
    tr->build_timed = false;
    return fail(TT_ERR_HIP, err);
  }
  (void)hipEventRecord(tr->ev1, s);
  tr->build_timed = true;
  return ok();
}

int tt_bvh_cost(tt_tracer *tr, float *cost_dev, void *stream) {
  TT_ARG(tr && cost_dev, "NULL argument");
  if (tr->bvh.P < 0) return fail(TT_ERR_STATE, "no acceleration structure: build it first");
  std::string err = tt::bvh_cost(tr->bvh, cost_dev, static_cast<hipStream_t>(stream));
  return err.empty() ? ok() : fail(TT_ERR_HIP, err);
}

int tt_set_rebuild_policy(tt_tracer *tr, float max_ratio) {
  TT_ARG(tr, "tracer is NULL");
  TT_ARG(max_ratio == max_ratio, "max_ratio is NaN");
  tr->rebuild_ratio = max_ratio;
  return ok();
}

int tt_build_stats(tt_tracer *tr, int counts[3], float costs[2]) {
  TT_ARG(tr && counts && costs, "NULL argument");
  counts[0] = tr->n_builds;
  counts[1] = tr->n_refits;
  counts[2] = tr->n_auto_rebuilds;
  costs[0] = costs[1] = -1.0f;
  if (tr->sah_dev && tr->n_builds > 0) {
    if (hipDeviceSynchronize() != hipSuccess || hipMemcpy(costs, tr->sah_dev, 2 * sizeof(float), hipMemcpyDeviceToHost) !=
                                                    hipSuccess)
      return fail(TT_ERR_HIP, "SAH cost read-back failed");
  }
  return ok();
}

int tt_set_kernel_timing(tt_tracer *tr, int enable) {
  TT_ARG(tr, "tracer is NULL");
  if (enable && !tr->tev[0][0][0]) {
    for (auto &dir : tr->tev)
      for (auto &pair : dir)
        for (hipEvent_t &ev : pair) {
          hipError_t e = hipEventCreate(&ev);
          if (e != hipSuccess) return fail(TT_ERR_HIP, std::string("hipEventCreate failed: ") + hipGetErrorString(e));
        }
  }
  tr->timing = enable != 0;
  tr->tcount[0] = tr->tcount[1] = 0;
  return ok();
}

int tt_kernel_times(tt_tracer *tr, int which, float *ms_out, int max_n, int *count_out) {
  TT_ARG(tr && ms_out && count_out, "NULL argument");
  TT_ARG(which == 0 || which == 1, "which must be 0 (forward) or 1 (backward)");
  TT_ARG(max_n >= 0, "max_n must be non-negative");
  const unsigned long long have = tr->tcount[which];
  unsigned long long n = std::min<unsigned long long>({have, (unsigned long long)max_n, (unsigned long long)TT_TIMING_RING});
  for (unsigned long long i = 0; i < n; i++) {
    const unsigned slot = (unsigned)((have - n + i) % TT_TIMING_RING);
    hipEvent_t b = tr->tev[which][slot][0], e = tr->tev[which][slot][1];
    hipError_t r = hipEventSynchronize(e);
    float ms = 0.0f;
    if (r == hipSuccess) r = hipEventElapsedTime(&ms, b, e);
    if (r != hipSuccess) return fail(TT_ERR_HIP, std::string("kernel timing events: ") + hipGetErrorString(r));
    ms_out[i] = ms;
  }
  *count_out = (int)n;
  return ok();
}

int tt_num_primitives(const tt_tracer *tr) {
  return tr ? tr->bvh.P : -1;
}

float tt_last_build_ms(const tt_tracer *tr) {
  if (!tr || !tr->build_timed) return -1.0f;
  if (hipEventSynchronize(tr->ev1) != hipSuccess) return -1.0f;
  float ms = -1.0f;
  if (hipEventElapsedTime(&ms, tr->ev0, tr->ev1) != hipSuccess) return -1.0f;
  return ms;
}

int tt_trace_fwd(tt_tracer *tr, const tt_fwd_args *a, void *stream) {
  TT_ARG(a, "args is NULL");
  int rc = check_common(tr, a->width, a->height, a->ray_o_dev, a->ray_d_dev, a->vertices_dev, a->npts_dev,
                        a->cumsum_dev, a->P, a->shs_dev, a->M, a->sh_degree, a->opacities_dev, a->sigma_dev, a->k,
                        a->chunk_mode);
  if (rc != TT_OK) return rc;
  TT_ARG(a->variant >= TT_VARIANT_FORWARD && a->variant <= TT_VARIANT_BOUNCE, "unknown variant");
  const long long R = (long long)a->width * a->height;
  if (R > 0) {
    TT_ARG(a->color_dev && a->bg_dev, "colour / background buffers are NULL");
    if (a->variant == TT_VARIANT_BOUNCE) {
      TT_ARG(a->envmap_dev && a->env_width > 0 && a->env_height > 0, "bounce variant needs an environment map");
      TT_ARG((long long)a->env_width * a->env_height < (1ll << 29), "environment map too large");
      TT_ARG(a->bounce_aux_dev, "bounce variant needs bounce_aux");
      if (a->P > 0) TT_ARG(a->reflectivity_dev && a->ior_dev, "bounce variant needs reflectivity and ior");
    }
    if (a->variant == TT_VARIANT_FORWARD || a->variant == TT_VARIANT_BOUNCE) {
      TT_ARG(a->others_dev, "others buffer is NULL");
      if (a->P > 0) TT_ARG(a->max_blending_dev && a->density_dev && a->scaling_dev, "statistics buffers are NULL");
    }
    if (a->variant == TT_VARIANT_REFLECT || a->variant == TT_VARIANT_REFRACT)
      TT_ARG(a->extra_dev && a->extra_len >= 4, "reflect/refract need extra_params = (r, cx, cy, cz)");
    if (a->variant == TT_VARIANT_ENVLIGHT)
      TT_ARG(a->extra_dev && a->extra_len >= 5, "envlight needs extra_params = (r, cx, cy, cz, samples)");
    if (a->hit_log_cap > 0)
      TT_ARG(a->hit_log_prim_dev && a->hit_log_t_dev && a->hit_log_code_dev && a->hit_log_count_dev,
             "hit log buffers are NULL");
    if (a->interval_in_dev || a->interval_out_dev)
      TT_ARG(a->variant == TT_VARIANT_FORWARD, "depth intervals are supported by the forward variant only");
    if (a->interval_in_dev)
      TT_ARG(a->chunk_mode == TT_CHUNK_EXACT, "interval_in requires TT_CHUNK_EXACT");
    if (a->hitlist_dev || a->hitlist_count_dev) {
      TT_ARG(a->variant == TT_VARIANT_FORWARD || a->variant == TT_VARIANT_BOUNCE,
             "the hit list is supported by TT_VARIANT_FORWARD and TT_VARIANT_BOUNCE");
      TT_ARG(a->hitlist_dev && a->hitlist_count_dev && a->hitlist_cap > 0,
             "the hit list needs hitlist, hitlist_count and hitlist_cap > 0");
    }
  }
  if (R == 0) return ok();
  tt::TraceParams p;
  fill_common(p, tr, a->width, a->height, a->ray_o_dev, a->ray_d_dev, a->vertices_dev, a->npts_dev, a->cumsum_dev,
              a->P, a->shs_dev, a->M, a->sh_degree, a->opacities_dev, a->sigma_dev, a->chunk_mode, a->cutoff,
              a->near_t, a->far_t);
  p.bg = a->bg_dev;
  p.extra = a->extra_dev;
  p.extra_len = a->extra_len;
  p.color = a->color_dev;
  p.others = a->others_dev;
  p.maxb = a->max_blending_dev;
  p.density = a->density_dev;
  p.scaling = a->scaling_dev;
  p.stats = a->stats_dev;
  p.trav_stats = a->trav_stats_dev;
  p.log_cap = a->hit_log_cap > 0 ? a->hit_log_cap : 0;
  p.log_prim = a->hit_log_prim_dev;
  p.log_t = a->hit_log_t_dev;
  p.log_code = a->hit_log_code_dev;
  p.log_count = a->hit_log_count_dev;
  p.ival_in = reinterpret_cast<const float2 *>(a->interval_in_dev);
  p.ival_ready = a->interval_ready_event;
  p.ival_out = reinterpret_cast<float4 *>(a->interval_out_dev);
  p.ival_hard = a->interval_in_dev && a->interval_hard_cap ? 1 : 0;
  p.ival_fb_start = a->interval_fallback_start;
  p.ival_fb_end = a->interval_fallback_end;
  if (a->hitlist_dev) {
    p.hl = reinterpret_cast<uint2 *>(a->hitlist_dev);
    p.hl_rgb = a->hitlist_color_dev;
    p.hl_count = a->hitlist_count_dev;
    p.hl_cap = a->hitlist_cap;
    p.hl_stride = (size_t)a->width * a->height;
  }
  if (a->variant == TT_VARIANT_BOUNCE) {
    p.refl = a->reflectivity_dev;
    p.ior = a->ior_dev;
    p.env = a->envmap_dev;
    p.env_w = a->env_width;
    p.env_h = a->env_height;
    p.aux = a->bounce_aux_dev;
  }
  if ((rc = prepare_spill(tr, a->width, a->height, p)) != TT_OK) return rc;
  if ((rc = prepare_tri_win(tr, p)) != TT_OK) return rc;
  if (a->variant == TT_VARIANT_BOUNCE && (long long)a->width * a->height > 0) {  // secondary-ray binning
    if ((rc = prepare_rec_sort(tr, p, (size_t)a->width * a->height)) != TT_OK) return rc;
    p.ray_keys = p.rec_keys;
    p.rec_keys = nullptr;
  }
  const int nt = tt::trace_tiles(a->width, a->height);
  if ((rc = ensure_tiles(tr, nt)) != TT_OK) return rc;
  p.tile_cost = tr->tile_cost;
  p.tile_order = tr->tile_order;
  p.probe_order = tr->probe_order;
  p.tile_hits = tr->tile_hits;
  p.rec_off_out = tr->rec_off;
  p.tile_hits_b = tr->tile_hits_b;
  p.tile_cost_b = tr->tile_cost_b;
  tt::KernelEvents kev;
  const unsigned slot_f = (unsigned)(tr->tcount[0] % TT_TIMING_RING);
  if (tr->timing) kev.before = tr->tev[0][slot_f][0], kev.after = tr->tev[0][slot_f][1];
  std::string err = tt::launch_forward(p, a->k, a->variant, static_cast<hipStream_t>(stream),
                                       tr->timing ? &kev : nullptr);
  if (!err.empty()) return fail(TT_ERR_HIP, err);
  if (tr->timing) tr->tcount[0]++;
  poll_hits_total(tr);
  if (!tr->hits_total_pending && a->variant == TT_VARIANT_FORWARD) {  // publish this frame's record total
    if (!tr->hits_total_host && hipHostMalloc((void **)&tr->hits_total_host, sizeof(int)) != hipSuccess)
      tr->hits_total_host = nullptr;
    if (tr->hits_total_host && !tr->hits_total_ev && hipEventCreateWithFlags(&tr->hits_total_ev, hipEventDisableTiming) != hipSuccess)
      tr->hits_total_ev = nullptr;
    if (tr->hits_total_host && tr->hits_total_ev) {
      const int nb = tt::backward_tiles(a->width, a->height);
      hipStream_t s = static_cast<hipStream_t>(stream);
      if (hipMemcpyAsync(tr->hits_total_host, tr->rec_off + nb, sizeof(int), hipMemcpyDeviceToHost, s) == hipSuccess &&
          hipEventRecord(tr->hits_total_ev, s) == hipSuccess)
        tr->hits_total_pending = true;
    }
  }
  tr->sched_W = a->width;
  tr->sched_H = a->height;
  tr->multi_tiles = -1;
  tr->sched_gen++;  // tile_order now holds the one-shot tiles of this image
  return ok();
}

int tt_trace_bwd(tt_tracer *tr, const tt_bwd_args *a, void *stream) {
  TT_ARG(a, "args is NULL");
  int rc = check_common(tr, a->width, a->height, a->ray_o_dev, a->ray_d_dev, a->vertices_dev, a->npts_dev,
                        a->cumsum_dev, a->P, a->shs_dev, a->M, a->sh_degree, a->opacities_dev, a->sigma_dev, a->k,
                        a->chunk_mode);
  if (rc != TT_OK) return rc;
  const long long R = (long long)a->width * a->height;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (a->P > 0) {
    TT_ARG(a->dL_dvertices_dev && a->dL_dsigma_dev && a->dL_dopacity_dev && a->dL_dsh_dev,
           "gradient buffers are NULL");
    TT_ARG(a->dL_dvertices_rows >= 0, "dL_dvertices_rows must be non-negative");
    if (a->zero_grads) {
      std::string e;
      e = tt::launch_fill(a->dL_dvertices_dev, (size_t)a->dL_dvertices_rows * 3, 0.0f, s);
      if (e.empty()) e = tt::launch_fill(a->dL_dsigma_dev, (size_t)a->P, 0.0f, s);
      if (e.empty()) e = tt::launch_fill(a->dL_dopacity_dev, (size_t)a->P, 0.0f, s);
      if (e.empty()) e = tt::launch_fill(a->dL_dsh_dev, (size_t)a->P * a->M * 3, 0.0f, s);
      if (e.empty() && a->variant == TT_VARIANT_BOUNCE && a->dL_dreflectivity_dev)
        e = tt::launch_fill(a->dL_dreflectivity_dev, (size_t)a->P, 0.0f, s);
      if (!e.empty()) return fail(TT_ERR_HIP, e);
    }
  }
  TT_ARG(a->variant == TT_VARIANT_FORWARD || a->variant == TT_VARIANT_BOUNCE,
         "the backward supports TT_VARIANT_FORWARD and TT_VARIANT_BOUNCE");
  if (a->variant == TT_VARIANT_BOUNCE) {
    TT_ARG(a->envmap_dev && a->env_width > 0 && a->env_height > 0 && a->dL_denvmap_dev,
           "bounce backward needs the environment map and its gradient buffer");
    TT_ARG((long long)a->env_width * a->env_height < (1ll << 29), "environment map too large");
    TT_ARG(a->interval_dev == nullptr, "interval is not supported by the bounce backward");
    if (a->zero_grads) {
      std::string e = tt::launch_fill(a->dL_denvmap_dev, (size_t)a->env_width * a->env_height * 3, 0.0f, s);
      if (!e.empty()) return fail(TT_ERR_HIP, e);
    }
    if (R > 0 && a->P > 0)
      TT_ARG(a->bounce_aux_dev && a->reflectivity_dev && a->dL_dreflectivity_dev,
             "bounce backward needs bounce_aux, reflectivity and its gradient buffer");
  }
  if (R == 0 || a->P == 0) return ok();
  TT_ARG(a->color_dev && a->others_dev && a->dL_dpix_dev && a->dL_dothers_dev,
         "forward outputs / upstream gradients are NULL");
  if (a->interval_dev) TT_ARG(a->chunk_mode == TT_CHUNK_EXACT, "interval requires TT_CHUNK_EXACT");
  if (a->hitlist_dev || a->hitlist_count_dev) {
    TT_ARG(a->hitlist_dev && a->hitlist_count_dev && a->hitlist_cap > 0,
           "the hit list needs hitlist, hitlist_count and hitlist_cap > 0");
  }
  tt::TraceParams p;
  fill_common(p, tr, a->width, a->height, a->ray_o_dev, a->ray_d_dev, a->vertices_dev, a->npts_dev, a->cumsum_dev,
              a->P, a->shs_dev, a->M, a->sh_degree, a->opacities_dev, a->sigma_dev, a->chunk_mode, a->cutoff,
              a->near_t, a->far_t);
  p.fcolor = a->color_dev;
  p.fothers = a->others_dev;
  p.dpix = a->dL_dpix_dev;
  p.dothers = a->dL_dothers_dev;
  p.dverts = a->dL_dvertices_dev;
  p.dsigma = a->dL_dsigma_dev;
  p.dopac = a->dL_dopacity_dev;
  p.dsh = a->dL_dsh_dev;
  p.ival_bwd = reinterpret_cast<const float4 *>(a->interval_dev);
  if (a->hitlist_dev) {
    p.hl_in = reinterpret_cast<const uint2 *>(a->hitlist_dev);
    p.hl_rgb_in = a->hitlist_color_dev;
    p.hl_count_in = a->hitlist_count_dev;
    p.hl_cap = a->hitlist_cap;
    p.hl_stride = (size_t)a->width * a->height;
  }
  if (a->variant == TT_VARIANT_BOUNCE) {
    p.refl = a->reflectivity_dev;
    p.env = a->envmap_dev;
    p.env_w = a->env_width;
    p.env_h = a->env_height;
    p.aux = const_cast<float *>(a->bounce_aux_dev);
    p.drefl = a->dL_dreflectivity_dev;
    p.denv = a->dL_denvmap_dev;
  }
  if ((rc = prepare_spill(tr, a->width, a->height, p)) != TT_OK) return rc;
  if ((rc = prepare_tri_win(tr, p, a->same_scene_as_forward != 0)) != TT_OK) return rc;
  {  // 16 rows per lane, or the largest record total a forward has needed so far
    poll_hits_total(tr);
    const size_t lanes = tt::trace_lanes(a->width, a->height);
    const size_t want = std::min<size_t>(tr->hits_high_water, (size_t)1 << 28);  // at most 20 GiB of rows
    const int lane_rows = (int)std::max<size_t>(16, (want + lanes - 1) / lanes);
    if ((rc = prepare_records(tr, lanes, p, lane_rows)) != TT_OK) return rc;
  }
  if (a->variant == TT_VARIANT_BOUNCE && (rc = prepare_rec_sort(tr, p, (size_t)p.rec_cap)) != TT_OK) return rc;
  if (tr->sched_W == a->width && tr->sched_H == a->height) {  // the last forward's schedule and record regions
    p.tile_sched = tr->tile_order;
    p.tile_hits = tr->tile_hits_b;  // the backward's tiles
    p.rec_off = tr->rec_off;
  }
  tt::KernelEvents kev;
  const unsigned slot_b = (unsigned)(tr->tcount[1] % TT_TIMING_RING);
  if (tr->timing) kev.before = tr->tev[1][slot_b][0], kev.after = tr->tev[1][slot_b][1];
  std::string err = tt::launch_backward(p, a->k, a->variant, s, tr->timing ? &kev : nullptr);
  if (!err.empty()) return fail(TT_ERR_HIP, err);
  if (tr->timing) tr->tcount[1]++;
  return ok();
}

}  // extern "C"

// ------------------------------------------------------------- multi-camera
namespace {

// Validation, camera table and launch parameters of a multi-camera launch (no allocation).
int multi_params(tt_tracer *tr, const tt_multi_args *a, bool backward, tt::MultiCams &mc, tt::TraceParams &p,
                 long long &rays_out) {
  TT_ARG(a, "args is NULL");
  TT_ARG(a->n_cameras >= 1 && a->n_cameras <= TT_MAX_CAMERAS, "n_cameras must be in [1, 32]");
  TT_ARG(a->width && a->height, "camera size arrays are NULL");
  TT_ARG(a->block == 64 || a->block == 128 || a->block == 256, "block must be 64, 128 or 256");
  TT_ARG(a->tiles_per_fetch >= 1, "tiles_per_fetch must be >= 1");
  long long rays = 0;
  int tw, th;
  tt::multi_tile_shape(a->block, tw, th);
  mc = tt::MultiCams{};
  mc.n = a->n_cameras;
  int tiles = 0;
  for (int c = 0; c < a->n_cameras; c++) {
    TT_ARG(a->width[c] > 0 && a->height[c] > 0, "camera sizes must be positive");
    mc.W[c] = a->width[c];
    mc.H[c] = a->height[c];
    mc.tiles_x[c] = (a->width[c] + tw - 1) / tw;
    mc.tile_begin[c] = tiles;
    mc.ray_off[c] = rays;
    tiles += mc.tiles_x[c] * ((a->height[c] + th - 1) / th);
    rays += (long long)a->width[c] * a->height[c];
    TT_ARG(tiles < (1 << 30), "too many tiles");
  }
  mc.tile_begin[a->n_cameras] = tiles;
  mc.total_tiles = tiles;
  for (int q = 0; q <= 8; q++) mc.q_begin[q] = (int)((long long)tiles * q / 8);
  const int with_rays = a->cameras_dev ? 0 : 1;  // fused ray generation reads no ray buffers
  int rc = check_common(tr, with_rays, with_rays, a->ray_o_dev, a->ray_d_dev, a->vertices_dev, a->npts_dev,
                        a->cumsum_dev, a->P, a->shs_dev, a->M, a->sh_degree, a->opacities_dev, a->sigma_dev, a->k,
                        a->chunk_mode);
  if (rc != TT_OK) return rc;
  fill_common(p, tr, 1, 1, a->ray_o_dev, a->ray_d_dev, a->vertices_dev, a->npts_dev, a->cumsum_dev, a->P, a->shs_dev,
              a->M, a->sh_degree, a->opacities_dev, a->sigma_dev, a->chunk_mode, a->cutoff, a->near_t, a->far_t);
  p.bg = a->bg_dev;
  p.cams = a->cameras_dev;
  if (!backward && a->target_dev) {
    TT_ARG(a->loss_kind == TT_LOSS_L2 || a->loss_kind == TT_LOSS_L1, "loss_kind must be TT_LOSS_L2 or TT_LOSS_L1");
    TT_ARG(a->loss_grad_dev && a->loss_dev, "the fused loss needs loss_grad_dev and loss_dev");
    p.target = a->target_dev;
    p.loss_grad = a->loss_grad_dev;
    p.loss_kind = a->loss_kind;
    p.loss_norm = (float)(1.0 / (3.0 * (double)rays));
  }
  if (!backward) {
    TT_ARG(a->color_dev && a->others_dev && a->bg_dev, "colour / others / background buffers are NULL");
    if (a->P > 0) TT_ARG(a->max_blending_dev && a->density_dev && a->scaling_dev, "statistics buffers are NULL");
    p.color = a->color_dev;
    p.others = a->others_dev;
    p.maxb = a->max_blending_dev;
    p.density = a->density_dev;
    p.scaling = a->scaling_dev;
    p.stats = a->stats_dev;
  } else {
    TT_ARG(a->color_dev && a->others_dev && a->dL_dpix_dev && a->dL_dothers_dev,
           "forward outputs / upstream gradients are NULL");
    if (a->P > 0)
      TT_ARG(a->dL_dvertices_dev && a->dL_dsigma_dev && a->dL_dopacity_dev && a->dL_dsh_dev, "gradient buffers are NULL");
    p.fcolor = a->color_dev;
    p.fothers = a->others_dev;
    p.dpix = a->dL_dpix_dev;
    p.dothers = a->dL_dothers_dev;
    p.dverts = a->dL_dvertices_dev;
    p.dsigma = a->dL_dsigma_dev;
    p.dopac = a->dL_dopacity_dev;
    p.dsh = a->dL_dsh_dev;
  }
  if (a->hitlist_dev || a->hitlist_count_dev) {
    TT_ARG(a->hitlist_dev && a->hitlist_count_dev && a->hitlist_cap > 0,
           "the hit list needs hitlist, hitlist_count and hitlist_cap > 0");
    p.hl_stride = (size_t)rays;
    if (!backward) {
      p.hl = reinterpret_cast<uint2 *>(a->hitlist_dev);
      p.hl_rgb = a->hitlist_color_dev;
      p.hl_count = a->hitlist_count_dev;
      p.hl_cap = a->hitlist_cap;
    } else {
      p.hl_in = reinterpret_cast<const uint2 *>(a->hitlist_dev);
      p.hl_rgb_in = a->hitlist_color_dev;
      p.hl_count_in = a->hitlist_count_dev;
      p.hl_cap = a->hitlist_cap;
    }
  }
  rays_out = rays;
  return TT_OK;
}

// multi_params plus the tracer's persistent grid, work-queue heads, tile schedule, record
// rows and stack spill (allocated or grown here).
int multi_setup(tt_tracer *tr, const tt_multi_args *a, bool backward, tt::MultiCams &mc, tt::TraceParams &p,
                int &grid, bool reuse_records = false) {
  long long rays = 0;
  int rc = multi_params(tr, a, backward, mc, p, rays);
  if (rc != TT_OK) return rc;
  const int tiles = mc.total_tiles;
  const long long key = ((long long)a->k << 20) | ((long long)a->block << 1) | (backward ? 1 : 0);
  auto it = tr->grids.find(key);
  if (it == tr->grids.end()) {
    const int g = tt::multi_grid(a->k, a->block, backward);
    if (g <= 0) return fail(TT_ERR_HIP, "could not size the persistent grid");
    it = tr->grids.emplace(key, g).first;
  }
  grid = it->second;
  if (!tr->heads) {
    hipError_t e = hipMalloc(&tr->heads, 8 * sizeof(int));
    if (e != hipSuccess) return fail(TT_ERR_HIP, std::string("queue allocation failed: ") + hipGetErrorString(e));
    tr->mem_gen++;
  }
  if ((rc = ensure_tiles(tr, tiles)) != TT_OK) return rc;
  if (p.target) {
    if (tiles > tr->loss_cap) {
      if (tr->loss_part) (void)hipFree(tr->loss_part);
      tr->loss_part = nullptr;
      tr->loss_cap = 0;
      if (hipMalloc(&tr->loss_part, (size_t)tiles * sizeof(float)) != hipSuccess)
        return fail(TT_ERR_HIP, "hipMalloc of the loss partial sums failed");
      tr->loss_cap = tiles;
      tr->mem_gen++;
    }
    p.loss_part = tr->loss_part;
  }
  if (!backward) {  // schedule: probe order for the forward, measured tile times for the backward
    p.tile_cost = tr->tile_cost;
    p.tile_order = tr->tile_order;
    p.probe_order = tr->probe_order;
    tr->sched_W = tr->sched_H = -1;
    if (tr->multi_tiles != tiles || tr->multi_block != a->block || tr->multi_rays != rays) tr->sched_gen++;
    tr->multi_tiles = tiles;
    tr->multi_block = a->block;
    tr->multi_rays = rays;
  } else if (tr->multi_tiles == tiles && tr->multi_block == a->block && tr->multi_rays == rays) {
    p.tile_sched = tr->tile_order;
  }
  if (backward) {  // a fixed region per resident block, reused tile after tile: 64 rows per lane
    const int rc2 = prepare_records(tr, (size_t)grid * a->block, p, 64);
    if (rc2 != TT_OK) return rc2;
  }
  if ((rc = prepare_tri_win(tr, p, reuse_records)) != TT_OK) return rc;
  return prepare_spill_lanes(tr, (size_t)grid * a->block, p);
}

int zero_grads(const tt_multi_args *a, hipStream_t s) {
  if (!a->zero_grads || a->P == 0) return TT_OK;
  std::string e = tt::launch_fill(a->dL_dvertices_dev, (size_t)a->dL_dvertices_rows * 3, 0.0f, s);
  if (e.empty()) e = tt::launch_fill(a->dL_dsigma_dev, (size_t)a->P, 0.0f, s);
  if (e.empty()) e = tt::launch_fill(a->dL_dopacity_dev, (size_t)a->P, 0.0f, s);
  if (e.empty()) e = tt::launch_fill(a->dL_dsh_dev, (size_t)a->P * a->M * 3, 0.0f, s);
  if (!e.empty()) return fail(TT_ERR_HIP, e);
  return TT_OK;
}

int multi_issue(tt_tracer *tr, const tt_multi_args *a, int what, hipStream_t s) {
  tt::MultiCams mc;
  tt::TraceParams p;
  int grid = 0, rc;
  if (what & 1) {
    if ((rc = multi_setup(tr, a, false, mc, p, grid)) != TT_OK) return rc;
    std::string e = tt::launch_multi(p, mc, tr->heads, a->k, a->block, a->tiles_per_fetch, grid, false, s,
                                     a->loss_dev);
    if (!e.empty()) return fail(TT_ERR_HIP, e);
  }
  if (what & 2) {  // after this call's forward, the backward reads the records it formed
    if ((rc = multi_setup(tr, a, true, mc, p, grid, (what & 1) != 0)) != TT_OK) return rc;
    if ((rc = zero_grads(a, s)) != TT_OK) return rc;
    if (a->P > 0) {
      std::string e = tt::launch_multi(p, mc, tr->heads, a->k, a->block, a->tiles_per_fetch, grid, true, s);
      if (!e.empty()) return fail(TT_ERR_HIP, e);
    }
  }
  return TT_OK;
}

}  // namespace

extern "C" {

int tt_multi_forward(tt_tracer *tr, const tt_multi_args *a, void *stream) {
  int rc = multi_issue(tr, a, 1, static_cast<hipStream_t>(stream));
  return rc == TT_OK ? ok() : rc;
}

int tt_multi_backward(tt_tracer *tr, const tt_multi_args *a, void *stream) {
  int rc = multi_issue(tr, a, 2, static_cast<hipStream_t>(stream));
  return rc == TT_OK ? ok() : rc;
}

int tt_multi_forward_noalloc(tt_tracer *tr, const tt_multi_args *a, void *stream) {
  TT_ARG(tr, "tracer is NULL");
  tt::MultiCams mc;
  tt::TraceParams p;
  long long rays = 0;
  int rc = multi_params(tr, a, false, mc, p, rays);
  if (rc != TT_OK) return rc;
  if (mc.total_tiles > TT_NOALLOC_MAX_TILES)
    return fail(TT_ERR_INVALID_ARG, "the batch has " + std::to_string(mc.total_tiles) + " tiles; the no-alloc forward "
                                    "takes at most TT_NOALLOC_MAX_TILES (" + std::to_string(TT_NOALLOC_MAX_TILES) +
                                    "); use tt_multi_forward or a larger block");
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev != tr->device)
    return fail(TT_ERR_STATE, "the current device is not the tracer's device");
  int *base = dev >= 0 && dev < kScratchDevices ? g_scratch_dev[dev] : nullptr;
  if (!base) return fail(TT_ERR_STATE, "the static scratch is not set up on this device (device index >= 64)");
  const int ki = a->k == 8 ? 0 : (a->k == 16 ? 1 : 2), bi = a->block == 64 ? 0 : (a->block == 128 ? 1 : 2);
  int &g = tr->noalloc_grid[ki][bi];
  if (g == 0) g = tt::multi_grid(a->k, a->block, false, false);
  if (g <= 0) return fail(TT_ERR_HIP, "could not size the persistent grid");
  int grid = g;
  const int depth = 3 * tr->bvh.height_max + 1 - tt::kLdsStack;
  p.ov_depth = depth > 0 ? depth : 0;
  p.ov_stack = base + kScratchSpillOff;
  if (p.ov_depth > 0 && tr->bvh.n_leaves > 0) {  // the grid whose stack spill fits the array
    const size_t fit = (kScratchInts - kScratchSpillOff) / ((size_t)a->block * p.ov_depth);
    if (fit < 1)
      return fail(TT_ERR_UNSUPPORTED, "the acceleration structure is too deep for the static scratch's stack spill");
    if ((size_t)grid > fit) grid = (int)fit;
  }
  p.tile_cost = reinterpret_cast<float *>(base + kScratchCostOff);
  p.probe_order = base + kScratchOrderOff;
  if (p.target) p.loss_part = reinterpret_cast<float *>(base + kScratchLossOff);
  p.tile_order = nullptr;  // no schedule for a backward: the tracer's own is left alone
  tr->sched_W = tr->sched_H = -1;
  tr->multi_tiles = -1;    // a following tt_multi_backward takes its tiles in queue order
  tr->sched_gen++;
  std::string e = tt::launch_multi(p, mc, base, a->k, a->block, a->tiles_per_fetch, grid, false,
                                   static_cast<hipStream_t>(stream), a->loss_dev);
  if (!e.empty()) return fail(TT_ERR_HIP, e);
  return ok();
}

int tt_multi_capture(tt_tracer *tr, const tt_multi_args *a, int what, void *stream) {
  TT_ARG(tr && a, "tracer / args is NULL");
  TT_ARG(what >= 1 && what <= 3, "what must be 1 (forward), 2 (backward) or 3 (both)");
  (void)stream;
  // all allocation happens here, outside the capture
  tt::MultiCams mc;
  tt::TraceParams p;
  int grid, rc;
  if ((what & 1) && (rc = multi_setup(tr, a, false, mc, p, grid)) != TT_OK) return rc;
  if ((what & 2) && (rc = multi_setup(tr, a, true, mc, p, grid)) != TT_OK) return rc;
  if (!tr->capture_stream) {
    hipError_t e = hipStreamCreateWithFlags(&tr->capture_stream, hipStreamNonBlocking);
    if (e != hipSuccess) return fail(TT_ERR_HIP, std::string("hipStreamCreate failed: ") + hipGetErrorString(e));
  }
  hipError_t e = hipStreamBeginCapture(tr->capture_stream, hipStreamCaptureModeRelaxed);
  if (e != hipSuccess) return fail(TT_ERR_HIP, std::string("hipStreamBeginCapture failed: ") + hipGetErrorString(e));
  rc = multi_issue(tr, a, what, tr->capture_stream);
  hipGraph_t g = nullptr;
  e = hipStreamEndCapture(tr->capture_stream, &g);
  if (rc != TT_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(TT_ERR_HIP, std::string("hipStreamEndCapture failed: ") + hipGetErrorString(e));
  if (tr->graph) (void)hipGraphExecDestroy(tr->graph);
  tr->graph = nullptr;
  e = hipGraphInstantiate(&tr->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    tr->graph = nullptr;
    return fail(TT_ERR_HIP, std::string("hipGraphInstantiate failed: ") + hipGetErrorString(e));
  }
  tr->graph_mem_gen = tr->mem_gen;
  tr->graph_sched_gen = tr->sched_gen;
  tr->graph_what = what;
  return ok();
}

int tt_multi_replay(tt_tracer *tr, void *stream) {
  TT_ARG(tr, "tracer is NULL");
  if (!tr->graph) return fail(TT_ERR_STATE, "no graph has been captured");
  if (tr->graph_mem_gen != tr->mem_gen)
    return fail(TT_ERR_STATE, "the captured graph is stale: the acceleration structure was rebuilt or the tracer's "
                              "scratch buffers were reallocated since tt_multi_capture; capture again");
  if (!(tr->graph_what & 1) && tr->graph_sched_gen != tr->sched_gen)
    return fail(TT_ERR_STATE, "the captured backward-only graph is stale: a forward of another tile layout rewrote "
                              "the tile schedule since tt_multi_capture; capture again");
  hipError_t e = hipGraphLaunch(tr->graph, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return fail(TT_ERR_HIP, std::string("hipGraphLaunch failed: ") + hipGetErrorString(e));
  return ok();
}

int tt_multi_autotune(tt_tracer *tr, tt_multi_args *a, const int *ks, int nk, const int *blocks, int nb,
                      const int *tpf, int nt, int reps, float *results_ms, void *stream) {
  TT_ARG(tr && a && ks && blocks && tpf && results_ms, "NULL argument");
  TT_ARG(nk > 0 && nb > 0 && nt > 0 && reps > 0, "empty sweep");
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return fail(TT_ERR_HIP, "hipEventCreate failed");
  const tt_multi_args base = *a;
  float best = -1.0f;
  tt_multi_args best_a = base;
  int rc = TT_OK;
  for (int i = 0; i < nk && rc == TT_OK; i++)
    for (int j = 0; j < nb && rc == TT_OK; j++)
      for (int l = 0; l < nt && rc == TT_OK; l++) {
        float &out = results_ms[(i * nb + j) * nt + l];
        out = -1.0f;
        if (base.chunk_mode != TT_CHUNK_EXACT && ks[i] != base.k) continue;  // k changes results in compat mode
        tt_multi_args c = base;
        c.k = ks[i];
        c.block = blocks[j];
        c.tiles_per_fetch = tpf[l];
        if ((rc = multi_issue(tr, &c, 3, s)) != TT_OK) break;  // warm-up (and allocations)
        std::vector<float> ms;
        for (int r = 0; r < reps; r++) {
          (void)hipEventRecord(e0, s);
          if ((rc = multi_issue(tr, &c, 3, s)) != TT_OK) break;
          (void)hipEventRecord(e1, s);
          (void)hipEventSynchronize(e1);
          float t = 0.0f;
          (void)hipEventElapsedTime(&t, e0, e1);
          ms.push_back(t);
        }
        if (rc != TT_OK) break;
        std::sort(ms.begin(), ms.end());
        out = ms[ms.size() / 2];
        if (best < 0.0f || out < best) best = out, best_a = c;
      }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != TT_OK) return rc;
  a->k = best_a.k;
  a->block = best_a.block;
  a->tiles_per_fetch = best_a.tiles_per_fetch;
  return ok();
}

}  // extern "C"
