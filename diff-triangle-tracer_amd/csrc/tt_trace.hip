// One-shot traversal kernels (one block per 16x8 tile of one image) and their
// launchers. The device code is in tt_trace_dev.h.
#include <hip/hip_runtime.h>

#include <string>

#include "tt_sort.h"
#include "tt_trace_dev.h"

namespace tt {
namespace {

constexpr int kBlock = 128;  // forward lanes per block: two 8x8 wave tiles -> 16x8 pixels
constexpr int kTileW = TileShape<kBlock>::W, kTileH = TileShape<kBlock>::H;
// The backward takes 16x16 tiles (two forward tiles stacked): a block then groups the
// records of twice the pixels, whose triangles overlap more, so fewer sums reach the
// gradient atomics (measured 3.46 -> 3.21 ms; the forward itself is faster on 16x8 tiles).
constexpr int kBlockB = 256;
static_assert(TileShape<kBlockB>::W == kTileW && TileShape<kBlockB>::H == 2 * kTileH, "backward tile = 2 forward tiles");

template <int K, int VARIANT, bool COUNT = false, int IVAL = 0>
__global__ __launch_bounds__(kBlock, 3) void k_forward(TraceParams P) {
  __shared__ int s_node[kLdsInts * kBlock];
  const long long c0 = P.tile_cost ? wall_clock64() : 0;
  const int tile = P.tile_sched ? P.tile_sched[blockIdx.x] : xcd_tile(P.n_tiles);
  int x, y, blended = 0;
  if (tile_pixel<kBlock>(P, tile, x, y))
    blended = forward_pixel<K, kBlock, VARIANT, COUNT, IVAL, true>(P, y * P.W + x, s_node);
  if (P.tile_cost) {  // the tile's wall time (100 MHz ticks, the backward's schedule key) and blended hits
    for (int off = 32; off > 0; off >>= 1) blended += __shfl_xor(blended, off);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) s_node[threadIdx.x >> 6] = blended;
    __syncthreads();
    if (threadIdx.x == 0) {
      P.tile_cost[tile] = (float)(wall_clock64() - c0);
      if (P.tile_hits) {
        int sum = 0;
        for (int w = 0; w < kBlock / 64; w++) sum += s_node[w];
        P.tile_hits[tile] = sum;
      }
    }
  }
}

// The bounce variant's secondary segments, one lane per pixel in the order of the sorted
// ray keys (pixels without secondary rays sort last and are skipped). Memory latency bounds
// it more than issue: four waves per SIMD (three: 19.7 ms per bounce forward, five: 20.8).
template <int K>
__global__ __launch_bounds__(kBlock, 4) void k_bounce_secondary(TraceParams P, const uint32_t *keys,
                                                                const uint32_t *order) {
  __shared__ int s_node[kLdsInts * kBlock];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < P.R && keys[i] != 0xffffffffu) bounce_secondary<K, kBlock, true>(P, (int)order[i], s_node);
}

// The forward's window records: the terms of each triangle's window function that do not
// depend on the ray (FP64 edge-plane normals, inverse lengths, incenter distance) with its
// vertices and sigma, one 128-byte record per triangle, read by every consumed hit.
__global__ __launch_bounds__(256) void k_tri_win(TraceParams P) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= P.P) return;
  const float *vp = P.verts + 3 * (size_t)P.cumsum[i];
  window_record_store(P.tri_win + kWinRec * (size_t)i, ld3(vp), ld3(vp + 3), ld3(vp + 6), P.sigma[i]);
}

// Exclusive prefix sums of the tiles' blended-hit counts (one block): the backward's
// per-tile record regions; off[n] = the total (capped at 2^31 - 1).
__global__ __launch_bounds__(1024) void k_tile_offsets(const int *hits, int n, int *off) {
  __shared__ long long s_part[1024];
  const int t = threadIdx.x;
  const int per = (n + 1023) / 1024, b = min(n, t * per), e = min(n, b + per);
  long long sum = 0;
  for (int i = b; i < e; i++) sum += hits[i];
  s_part[t] = sum;
  __syncthreads();
  for (int d = 1; d < 1024; d <<= 1) {  // Hillis-Steele inclusive scan
    const long long v = t >= d ? s_part[t - d] : 0;
    __syncthreads();
    s_part[t] += v;
    __syncthreads();
  }
  long long run = s_part[t] - sum;
  for (int i = b; i < e; i++) {
    off[i] = (int)min(run, (long long)0x7fffffff);
    run += hits[i];
  }
  if (t == 1023) off[n] = (int)min(s_part[1023], (long long)0x7fffffff);
}

// Cost estimate of every tile before the forward (probe_tile), one wave per tile.
__global__ __launch_bounds__(64) void k_tile_probe(TraceParams P, float *cost) {
  __shared__ int s_front[2 * kProbeFront];
  __shared__ int s_count;
  const float est = probe_tile<kBlock>(P, blockIdx.x, s_front, &s_count);
  if (threadIdx.x == 0) cost[blockIdx.x] = est;
}

// Tiles ordered by cost, most expensive first (counting sort over 1024 buckets of
// sqrt(cost) in one block). The backward then starts its heaviest tiles first, so the
// end of the launch is not left to a few long blocks.
__global__ __launch_bounds__(1024) void k_tile_order(const float *cost, int n, int *order) {
  __shared__ unsigned int hist[1024];
  __shared__ float s_max[16];
  const int t = threadIdx.x;
  float m = 0.0f;
  for (int i = t; i < n; i += 1024) m = fmaxf(m, cost[i]);
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
  if ((t & 63) == 0) s_max[t >> 6] = m;
  hist[t] = 0u;
  __syncthreads();
  float mx = 0.0f;
  for (int w = 0; w < 16; w++) mx = fmaxf(mx, s_max[w]);
  const float scale = mx > 0.0f ? 1023.0f / sqrtf(mx) : 0.0f;
  auto bucket = [&](int i) { return 1023 - min(1023, (int)(sqrtf(fmaxf(cost[i], 0.0f)) * scale)); };
  for (int i = t; i < n; i += 1024) atomicAdd(&hist[bucket(i)], 1u);
  __syncthreads();
  // exclusive scan of the 1024 bucket counts (Hillis-Steele, inclusive, then shift)
  unsigned int v = hist[t];
  for (int off = 1; off < 1024; off <<= 1) {
    __syncthreads();
    const unsigned int add = t >= off ? hist[t - off] : 0u;
    __syncthreads();
    hist[t] += add;
  }
  __syncthreads();
  hist[t] -= v;
  __syncthreads();
  for (int i = t; i < n; i += 1024) order[atomicAdd(&hist[bucket(i)], 1u)] = i;
}

// Backward through gradient records aggregated per block (no LDS atomics on the
// gradient fields; 16 KB of LDS per block instead of 42 KB).
template <int K, int VARIANT>
__global__ __launch_bounds__(kBlockB, 2) void k_backward_rec(TraceParams P) {
  __shared__ int s_node[kLdsInts * kBlockB];
  __shared__ int s_rec[rec_lds_ints<kBlockB>()];
  __shared__ int s_tab[rec_tab_ints<kBlockB>()];
  int x, y;
  const int tile = P.tile_sched ? P.tile_sched[blockIdx.x] : xcd_tile(P.n_tiles);
  const bool valid = tile_pixel<kBlockB>(P, tile, x, y);
  backward_block_rec<K, kBlockB, VARIANT, true>(P, valid, y * P.W + x, tile, s_node, s_rec, s_tab);
}

// Statistics of the forward's 16x8 tiles folded into the backward's 16x16 tiles: blended
// hits (the record regions) and wall time (the backward's order).
__global__ __launch_bounds__(256) void k_fold_tiles(const int *hits, const float *cost, int tiles_x, int n_f,
                                                    int n_b, int *hits_b, float *cost_b) {
  const int b = blockIdx.x * 256 + threadIdx.x;
  if (b >= n_b) return;
  const int bx = b % tiles_x, by = b / tiles_x;
  const int f0 = 2 * by * tiles_x + bx, f1 = f0 + tiles_x;
  int h = 0;
  float c = 0.0f;
  if (f0 < n_f) h += hits ? hits[f0] : 0, c += cost[f0];
  if (f1 < n_f) h += hits ? hits[f1] : 0, c += cost[f1];
  if (hits_b) hits_b[b] = h;
  cost_b[b] = c;
}

template <int K>
hipError_t fwd_k(const TraceParams &p, int variant, hipStream_t s) {
  dim3 g(p.n_tiles), bl(kBlock);
  switch (variant) {
    case 0:
      if (p.trav_stats)
        hipLaunchKernelGGL((k_forward<K, 0, true, kIvalIn | kIvalOut | kIvalHard>), g, bl, 0, s, p);
      else if (p.ival_hard && p.ival_out)
        hipLaunchKernelGGL((k_forward<K, 0, false, kIvalIn | kIvalOut | kIvalHard>), g, bl, 0, s, p);
      else if (p.ival_hard)
        hipLaunchKernelGGL((k_forward<K, 0, false, kIvalIn | kIvalHard>), g, bl, 0, s, p);
      else if (p.ival_out)
        hipLaunchKernelGGL((k_forward<K, 0, false, kIvalIn | kIvalOut>), g, bl, 0, s, p);
      else if (p.ival_in)
        hipLaunchKernelGGL((k_forward<K, 0, false, kIvalIn>), g, bl, 0, s, p);
      else
        hipLaunchKernelGGL((k_forward<K, 0>), g, bl, 0, s, p);
      break;
    case 1:
      hipLaunchKernelGGL((k_forward<K, 1>), g, bl, 0, s, p);
      break;
    case 2:
      hipLaunchKernelGGL((k_forward<K, 2>), g, bl, 0, s, p);
      break;
    case 3:
      hipLaunchKernelGGL((k_forward<K, 3>), g, bl, 0, s, p);
      break;
    default:
      if (p.ray_keys && p.sort_keys2 && p.sort_vals && p.sort_vals2 && p.sort_hist && p.log_cap == 0) {
        // primary segments, then the secondary rays binned by direction and pixel locality
        hipLaunchKernelGGL((k_forward<K, 5>), g, bl, 0, s, p);
        const int n = p.R;
        const int sblk = (n + kSortTile - 1) / kSortTile;
        hipLaunchKernelGGL(k_iota, dim3((n + 255) / 256), dim3(256), 0, s, p.sort_vals, n);
        uint32_t *kb[2] = {p.ray_keys, p.sort_keys2}, *vb[2] = {p.sort_vals, p.sort_vals2};
        int cur = 0;
        for (int shift = 0; shift < 32; shift += 8) {
          hipLaunchKernelGGL(k_hist, dim3(sblk), dim3(kSortThreads), 0, s, kb[cur], n, nullptr, shift, p.sort_hist,
                             sblk);
          const int nh = 256 * sblk, nc = (nh + kScanChunk - 1) / kScanChunk;
          uint32_t *totals = p.sort_hist + nh;
          hipLaunchKernelGGL(k_scan_local, dim3(nc), dim3(256), 0, s, p.sort_hist, nh, totals);
          hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, totals, nc);
          hipLaunchKernelGGL(k_scan_add, dim3(nc), dim3(256), 0, s, p.sort_hist, nh, totals);
          hipLaunchKernelGGL(k_scatter, dim3(sblk), dim3(kSortThreads), 0, s, kb[cur], vb[cur], kb[cur ^ 1],
                             vb[cur ^ 1], n, nullptr, shift, p.sort_hist, sblk);
          cur ^= 1;
        }
        hipLaunchKernelGGL((k_bounce_secondary<K>), dim3((n + kBlock - 1) / kBlock), bl, 0, s, p, kb[cur], vb[cur]);
      } else {
        hipLaunchKernelGGL((k_forward<K, 4>), g, bl, 0, s, p);
      }
      break;
  }
  return hipGetLastError();
}

}  // namespace

size_t trace_lanes(int W, int H) {  // of the forward's or the backward's tiling, the larger
  const size_t tx = (size_t)(W + kTileW - 1) / kTileW, ty = (size_t)(H + kTileH - 1) / kTileH;
  const size_t tyb = (size_t)(H + 2 * kTileH - 1) / (2 * kTileH);
  return tx * ty * kBlock > tx * tyb * kBlockB ? tx * ty * kBlock : tx * tyb * kBlockB;
}

static void tiles(TraceParams &p, int th = kTileH) {
  p.tiles_x = (p.W + kTileW - 1) / kTileW;
  p.n_tiles = p.tiles_x * ((p.H + th - 1) / th);
}

std::string launch_forward(const TraceParams &p0, int k, int variant, hipStream_t stream, const KernelEvents *ev) {
  TraceParams p = p0;
  tiles(p);
  if (p.n_tiles == 0) return "";
  hipError_t e;
  if (k != 8 && k != 16 && k != 32) return "k must be 8, 16 or 32";
  if (p.P > 0 && !p.tri_win) return "the forward needs its window-record buffer";
  if (p.P > 0) hipLaunchKernelGGL(k_tri_win, dim3((p.P + 255) / 256), dim3(256), 0, stream, p);
  p.tile_sched = nullptr;
  if (p.tile_cost && p.probe_order) {  // the forward's schedule: estimated tile costs
    hipLaunchKernelGGL(k_tile_probe, dim3(p.n_tiles), dim3(64), 0, stream,
                       p, p.tile_cost);
    hipLaunchKernelGGL(k_tile_order, dim3(1), dim3(1024), 0, stream, p.tile_cost, p.n_tiles, p.probe_order);
    e = hipGetLastError();
    if (e != hipSuccess) return std::string("tile probe launch failed: ") + hipGetErrorString(e);
    p.tile_sched = p.probe_order;
  }
  if (p.ival_ready) {  // the proposed intervals come from another stream: wait only now
    e = hipStreamWaitEvent(stream, static_cast<hipEvent_t>(p.ival_ready), 0);
    if (e != hipSuccess) return std::string("hipStreamWaitEvent on interval_ready_event failed: ") + hipGetErrorString(e);
  }
  if (ev) (void)hipEventRecord(ev->before, stream);
  if (k == 8)
    e = fwd_k<8>(p, variant, stream);
  else if (k == 16)
    e = fwd_k<16>(p, variant, stream);
  else if (k == 32)
    e = fwd_k<32>(p, variant, stream);
  else
    return "k must be 8, 16 or 32";
  if (ev) (void)hipEventRecord(ev->after, stream);
  if (e != hipSuccess) return std::string("forward launch failed: ") + hipGetErrorString(e);
  if (p.tile_cost && p.tile_order && p.tile_cost_b) {  // the backward's schedule and record regions
    TraceParams pb = p;
    tiles(pb, 2 * kTileH);
    hipLaunchKernelGGL(k_fold_tiles, dim3((pb.n_tiles + 255) / 256), dim3(256), 0, stream, p.tile_hits, p.tile_cost,
                       p.tiles_x, p.n_tiles, pb.n_tiles, p.tile_hits ? p.tile_hits_b : nullptr, p.tile_cost_b);
    if (p.tile_hits && p.tile_hits_b && p.rec_off_out)
      hipLaunchKernelGGL(k_tile_offsets, dim3(1), dim3(1024), 0, stream, p.tile_hits_b, pb.n_tiles, p.rec_off_out);
    hipLaunchKernelGGL(k_tile_order, dim3(1), dim3(1024), 0, stream, p.tile_cost_b, pb.n_tiles, p.tile_order);
    e = hipGetLastError();
    if (e != hipSuccess) return std::string("tile order launch failed: ") + hipGetErrorString(e);
  }
  return "";
}

std::string launch_tri_win(const TraceParams &p, hipStream_t stream) {
  if (p.P <= 0) return "";
  if (!p.tri_win) return "no window-record buffer";
  hipLaunchKernelGGL(k_tri_win, dim3((p.P + 255) / 256), dim3(256), 0, stream, p);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return std::string("window record launch failed: ") + hipGetErrorString(e);
  return "";
}

std::string launch_tile_order(const float *cost, int n, int *order, hipStream_t stream) {
  if (n == 0) return "";
  hipLaunchKernelGGL(k_tile_order, dim3(1), dim3(1024), 0, stream, cost, n, order);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return std::string("tile order launch failed: ") + hipGetErrorString(e);
  return "";
}

int trace_tiles(int W, int H) {  // the forward's tiles (at least as many as the backward's)
  return ((W + kTileW - 1) / kTileW) * ((H + kTileH - 1) / kTileH);
}

int backward_tiles(int W, int H) {
  return ((W + kTileW - 1) / kTileW) * ((H + 2 * kTileH - 1) / (2 * kTileH));
}

std::string launch_backward(const TraceParams &p0, int k, int variant, hipStream_t stream, const KernelEvents *ev) {
  TraceParams p = p0;
  tiles(p, 2 * kTileH);
  if (p.n_tiles == 0) return "";
  dim3 g(p.n_tiles), bl(kBlockB);
  if (k != 8 && k != 16 && k != 32) return "k must be 8, 16 or 32";
  // the window records of the vertices this backward is given, unless the host knows the
  // last forward's records hold them
  if (p.P > 0 && !p.tri_win) return "the backward needs its window-record buffer";
  if (p.P > 0 && !p.tri_win_ready) hipLaunchKernelGGL(k_tri_win, dim3((p.P + 255) / 256), dim3(256), 0, stream, p);
  if (ev) (void)hipEventRecord(ev->before, stream);
  if (p.rec_rows) {
    if (variant == 4) {
      // the primary segment (coherent rays) with per-block grouping of its records
      TraceParams p1 = p;
      p1.rec_keys = nullptr;
      if (k == 8) hipLaunchKernelGGL((k_backward_rec<8, 5>), g, bl, 0, stream, p1);
      if (k == 16) hipLaunchKernelGGL((k_backward_rec<16, 5>), g, bl, 0, stream, p1);
      if (k == 32) hipLaunchKernelGGL((k_backward_rec<32, 5>), g, bl, 0, stream, p1);
      if (p.rec_keys && p.sort_keys2 && p.sort_vals && p.sort_vals2 && p.sort_hist) {
        // the secondary segments: their records (keys = triangle ids) are sorted by triangle
        // over the whole launch and summed run by run
        const int n = p.rec_cap;
        hipLaunchKernelGGL(k_fill_int, dim3((n + 255) / 256), dim3(256), 0, stream,
                           reinterpret_cast<int *>(p.rec_keys), n, -1);
        if (k == 8) hipLaunchKernelGGL((k_backward_rec<8, 8>), g, bl, 0, stream, p);
        if (k == 16) hipLaunchKernelGGL((k_backward_rec<16, 8>), g, bl, 0, stream, p);
        if (k == 32) hipLaunchKernelGGL((k_backward_rec<32, 8>), g, bl, 0, stream, p);
        hipLaunchKernelGGL(k_iota, dim3((n + 255) / 256), dim3(256), 0, stream, p.sort_vals, n);
        const int sblk = (n + kSortTile - 1) / kSortTile;
        int bits = 1;  // digits covering P, so that 0xffffffff (masked) sorts after every triangle id
        while (bits < 32 && (1ll << bits) <= (long long)p.P) bits++;
        uint32_t *kb[2] = {p.rec_keys, p.sort_keys2}, *vb[2] = {p.sort_vals, p.sort_vals2};
        int cur = 0;
        for (int shift = 0; shift < bits; shift += 8) {
          hipLaunchKernelGGL(k_hist, dim3(sblk), dim3(kSortThreads), 0, stream, kb[cur], n, nullptr, shift,
                             p.sort_hist, sblk);
          const int nh = 256 * sblk, nc = (nh + kScanChunk - 1) / kScanChunk;
          uint32_t *totals = p.sort_hist + nh;
          hipLaunchKernelGGL(k_scan_local, dim3(nc), dim3(256), 0, stream, p.sort_hist, nh, totals);
          hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, stream, totals, nc);
          hipLaunchKernelGGL(k_scan_add, dim3(nc), dim3(256), 0, stream, p.sort_hist, nh, totals);
          hipLaunchKernelGGL(k_scatter, dim3(sblk), dim3(kSortThreads), 0, stream, kb[cur], vb[cur], kb[cur ^ 1],
                             vb[cur ^ 1], n, nullptr, shift, p.sort_hist, sblk);
          cur ^= 1;
        }
        const int windows = (n + 63) / 64;
        hipLaunchKernelGGL(k_rec_reduce_sorted, dim3(windows < 16384 ? windows : 16384), dim3(64), 0, stream, p,
                           kb[cur], vb[cur], n);
      } else {  // one launch per secondary segment with per-block grouping
        if (k == 8) {
          hipLaunchKernelGGL((k_backward_rec<8, 6>), g, bl, 0, stream, p1);
          hipLaunchKernelGGL((k_backward_rec<8, 7>), g, bl, 0, stream, p1);
        }
        if (k == 16) {
          hipLaunchKernelGGL((k_backward_rec<16, 6>), g, bl, 0, stream, p1);
          hipLaunchKernelGGL((k_backward_rec<16, 7>), g, bl, 0, stream, p1);
        }
        if (k == 32) {
          hipLaunchKernelGGL((k_backward_rec<32, 6>), g, bl, 0, stream, p1);
          hipLaunchKernelGGL((k_backward_rec<32, 7>), g, bl, 0, stream, p1);
        }
      }
    } else {
      if (k == 8) hipLaunchKernelGGL((k_backward_rec<8, 0>), g, bl, 0, stream, p);
      if (k == 16) hipLaunchKernelGGL((k_backward_rec<16, 0>), g, bl, 0, stream, p);
      if (k == 32) hipLaunchKernelGGL((k_backward_rec<32, 0>), g, bl, 0, stream, p);
    }
  } else {
    return "the backward needs its gradient-record rows";
  }
  if (ev) (void)hipEventRecord(ev->after, stream);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return std::string("backward launch failed: ") + hipGetErrorString(e);
  return "";
}

std::string launch_fill(float *ptr, size_t n, float v, hipStream_t stream) {
  if (n == 0) return "";
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, ptr, n, v);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return std::string("fill launch failed: ") + hipGetErrorString(e);
  return "";
}

}  // namespace tt
