// Persistent multi-camera traversal: every camera of a batch (uneven resolutions)
// is rendered forward or backward by ONE launch that fills the device with
// resident blocks; blocks pull tiles from 8 work queues. Queue q serves one
// contiguous band of the camera-major tile order and is the home queue of the
// blocks dispatched to XCD q (workgroups go to XCDs round-robin), so a block
// stays on pixels that share its XCD's L2 until its band is done, then steals
// from the other queues. The number of tiles per fetch is a tuning knob.
//
// A batched multi-view training step (several cameras per optimizer step) has
// no counterpart in the reference, which traces one camera per call
// (tracer.cpp:16-142); per-pixel work is the same device code as the one-shot
// kernels (tt_trace_dev.h), so the results equal per-camera launches.
#include <hip/hip_runtime.h>

#include <string>

#include "tt_trace_dev.h"

namespace tt {
namespace {

// Thread 0 of a block: claims `G` tiles, home queue first. Returns the first tile
// (end in `end`) or -1 when every queue is drained.
TT_D int fetch_tiles(int *heads, const MultiCams &mc, int G, int &end) {
  const int home = blockIdx.x & 7;
  for (int i = 0; i < 8; i++) {
    const int q = (home + i) & 7;
    const int qb = mc.q_begin[q], qe = mc.q_begin[q + 1];
    if (qb >= qe) continue;
    if (__hip_atomic_load(heads + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= qe - qb) continue;
    const int old = atomicAdd(heads + q, G);
    if (qb + old < qe) {
      end = min(qb + old + G, qe);
      return qb + old;
    }
  }
  return -1;
}

// The launch parameters seen from camera c = the camera owning global tile `tile`.
TT_D TraceParams camera_view(const TraceParams &P, const MultiCams &mc, int tile, int &local_tile) {
  int c = 0;
  while (c + 1 < mc.n && mc.tile_begin[c + 1] <= tile) c++;
  TraceParams Q = P;
  Q.W = mc.W[c];
  Q.H = mc.H[c];
  Q.R = mc.W[c] * mc.H[c];
  Q.tiles_x = mc.tiles_x[c];
  const long long off = mc.ray_off[c];
  Q.ray_o = P.ray_o + 3 * off;
  Q.ray_d = P.ray_d + 3 * off;
  if (P.color) Q.color = P.color + 3 * off;
  if (P.others) Q.others = P.others + 7 * off;
  if (P.stats) Q.stats = P.stats + 2 * off;
  if (P.fcolor) Q.fcolor = P.fcolor + 3 * off;
  if (P.fothers) Q.fothers = P.fothers + 7 * off;
  if (P.dpix) Q.dpix = P.dpix + 3 * off;
  if (P.dothers) Q.dothers = P.dothers + 7 * off;
  if (P.cams) Q.cam = P.cams + kCamDoubles * c;
  if (P.target) Q.target = P.target + 3 * off;
  if (P.loss_grad) Q.loss_grad = P.loss_grad + 3 * off;
  if (P.hl) Q.hl = P.hl + off;  // hit lists keep the launch-wide stride P.hl_stride
  if (P.hl_count) Q.hl_count = P.hl_count + off;
  if (P.hl_in) Q.hl_in = P.hl_in + off;
  if (P.hl_rgb) Q.hl_rgb = P.hl_rgb + 3 * off;
  if (P.hl_rgb_in) Q.hl_rgb_in = P.hl_rgb_in + 3 * off;
  if (P.hl_count_in) Q.hl_count_in = P.hl_count_in + off;
  local_tile = tile - mc.tile_begin[c];
  return Q;
}

// Cost estimate of every global tile (probe_tile on its camera), one wave per tile.
template <int BLOCK>
__global__ __launch_bounds__(64) void k_probe_multi(TraceParams P, MultiCams mc, float *cost) {
  __shared__ int s_front[2 * kProbeFront];
  __shared__ int s_count;
  const int t = blockIdx.x;
  int lt;
  const TraceParams Q = camera_view(P, mc, t, lt);
  const float est = probe_tile<BLOCK>(Q, lt, s_front, &s_count);
  if (threadIdx.x == 0) cost[t] = est;
}

// WREC: the consumed hits read the launch's window records (P.tri_win); the allocation-free
// forward has none and forms each triangle's terms per hit.
template <int K, int BLOCK, bool WREC>
__global__ __launch_bounds__(BLOCK, 3) void k_forward_multi(TraceParams P, MultiCams mc, int *heads, int G) {
  __shared__ int s_node[kLdsInts * BLOCK];
  __shared__ int s_range[2];
  __shared__ float s_loss[BLOCK / 64];
  while (true) {
    if (threadIdx.x == 0) {
      int e = 0;
      s_range[0] = fetch_tiles(heads, mc, G, e);
      s_range[1] = e;
    }
    __syncthreads();
    const int b = s_range[0], e = s_range[1];
    __syncthreads();
    if (b < 0) break;
    for (int pos = b; pos < e; pos++) {
      const int t = P.tile_sched ? P.tile_sched[pos] : pos;
      const long long c0 = P.tile_cost ? wall_clock64() : 0;
      int lt, x, y;
      const TraceParams Q = camera_view(P, mc, t, lt);
      float loss = 0.0f;
      if (tile_pixel<BLOCK>(Q, lt, x, y)) forward_pixel<K, BLOCK, 0, false, 0, WREC>(Q, y * Q.W + x, s_node, &loss);
      if (P.loss_part) {  // the tile's loss: lanes summed in a fixed order (wave tree, then the waves)
        for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off);
        __syncthreads();
        if ((threadIdx.x & 63) == 0) s_loss[threadIdx.x >> 6] = loss;
        __syncthreads();
        if (threadIdx.x == 0) {
          float sum = 0.0f;
          for (int w = 0; w < BLOCK / 64; w++) sum += s_loss[w];
          P.loss_part[t] = sum;
        }
      }
      if (P.tile_cost) {  // the tile's wall time: the backward's schedule key
        __syncthreads();
        if (threadIdx.x == 0) P.tile_cost[t] = (float)(wall_clock64() - c0);
      }
    }
  }
}

template <int K, int BLOCK>
__global__ __launch_bounds__(BLOCK, 2) void k_backward_multi(TraceParams P, MultiCams mc, int *heads, int G) {
  __shared__ int s_node[kLdsInts * BLOCK];
  __shared__ int s_rec[rec_lds_ints<BLOCK>()];
  __shared__ int s_tab[rec_tab_ints<BLOCK>()];
  __shared__ int s_range[2];
  while (true) {
    if (threadIdx.x == 0) {
      int e = 0;
      s_range[0] = fetch_tiles(heads, mc, G, e);
      s_range[1] = e;
    }
    __syncthreads();
    const int b = s_range[0], e = s_range[1];
    __syncthreads();
    if (b < 0) break;
    for (int pos = b; pos < e; pos++) {
      const int t = P.tile_sched ? P.tile_sched[pos] : pos;
      int lt, x, y;
      const TraceParams Q = camera_view(P, mc, t, lt);
      const bool valid = tile_pixel<BLOCK>(Q, lt, x, y);
      backward_block_rec<K, BLOCK, 0, true>(Q, valid, y * Q.W + x, lt, s_node, s_rec, s_tab);
    }
  }
}

// The batch's loss from the per-tile sums (one block, a fixed order: deterministic), scaled
// to the mean over all channels and pixels.
__global__ __launch_bounds__(1024) void k_loss_total(const float *part, int n, float norm, float *loss) {
  __shared__ double s[1024];
  const int t = threadIdx.x;
  const int per = (n + 1023) / 1024, b = min(n, t * per), e = min(n, b + per);
  double sum = 0.0;
  for (int i = b; i < e; i++) sum += (double)part[i];
  s[t] = sum;
  __syncthreads();
  for (int d = 512; d > 0; d >>= 1) {
    if (t < d) s[t] += s[t + d];
    __syncthreads();
  }
  if (t == 0) *loss = (float)(s[0] * (double)norm);
}

template <int K, int BLOCK>
const void *kernel_of(bool backward, bool wrec) {
  if (backward) return reinterpret_cast<const void *>(&k_backward_multi<K, BLOCK>);
  return wrec ? reinterpret_cast<const void *>(&k_forward_multi<K, BLOCK, true>)
              : reinterpret_cast<const void *>(&k_forward_multi<K, BLOCK, false>);
}

template <int K>
const void *kernel_k(int block, bool backward, bool wrec) {
  if (block == 64) return kernel_of<K, 64>(backward, wrec);
  if (block == 128) return kernel_of<K, 128>(backward, wrec);
  if (block == 256) return kernel_of<K, 256>(backward, wrec);
  return nullptr;
}

const void *multi_kernel(int k, int block, bool backward, bool wrec) {
  if (k == 8) return kernel_k<8>(block, backward, wrec);
  if (k == 16) return kernel_k<16>(block, backward, wrec);
  if (k == 32) return kernel_k<32>(block, backward, wrec);
  return nullptr;
}

}  // namespace

void multi_tile_shape(int block, int &tw, int &th) {
  tw = block == 64 ? 8 : 16;
  th = block == 256 ? 16 : 8;
}

int multi_grid(int k, int block, bool backward, bool wrec) {
  const void *fn = multi_kernel(k, block, backward, wrec);
  if (!fn) return 0;
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, block, 0) != hipSuccess) return 0;
  return cus * (per_cu > 0 ? per_cu : 1);
}

std::string launch_multi(const TraceParams &p, const MultiCams &mc, int *heads, int k, int block, int G, int grid,
                         bool backward, hipStream_t stream, float *loss_out) {
  // the backward always reads window records; a forward does when it is given them
  if (backward && p.P > 0 && !p.tri_win) return "the backward needs its window-record buffer";
  const void *fn = multi_kernel(k, block, backward, p.tri_win != nullptr);
  if (!fn) return "unsupported (k, block): k in {8, 16, 32}, block in {64, 128, 256}";
  if (mc.total_tiles == 0) return "";
  if (p.tri_win && p.P > 0 && !p.tri_win_ready) {
    std::string e = launch_tri_win(p, stream);
    if (!e.empty()) return e;
  }
  TraceParams pp = p;
  MultiCams m = mc;
  const int n = mc.total_tiles;
  if (!backward && p.tile_cost && p.probe_order) {  // heaviest tiles first, by the probe's estimate
    const dim3 g(n), b(64);
    if (block == 64) hipLaunchKernelGGL(k_probe_multi<64>, g, b, 0, stream, p, mc, p.tile_cost);
    if (block == 128) hipLaunchKernelGGL(k_probe_multi<128>, g, b, 0, stream, p, mc, p.tile_cost);
    if (block == 256) hipLaunchKernelGGL(k_probe_multi<256>, g, b, 0, stream, p, mc, p.tile_cost);
    std::string e = launch_tile_order(p.tile_cost, n, p.probe_order, stream);
    if (!e.empty()) return e;
    pp.tile_sched = p.probe_order;
  }
  if (pp.tile_sched) {  // one queue in schedule order
    m.q_begin[0] = 0;
    for (int q = 1; q <= 8; q++) m.q_begin[q] = n;
  }
  std::string e = launch_fill_int(heads, 8, 0, stream);
  if (!e.empty()) return e;
  void *args[] = {&pp, &m, &heads, &G};
  hipError_t err = hipLaunchKernel(fn, dim3(grid), dim3(block), args, 0, stream);
  if (err != hipSuccess) return std::string("persistent launch failed: ") + hipGetErrorString(err);
  if (!backward && p.loss_part && loss_out) {
    hipLaunchKernelGGL(k_loss_total, dim3(1), dim3(1024), 0, stream, p.loss_part, n, p.loss_norm, loss_out);
    err = hipGetLastError();
    if (err != hipSuccess) return std::string("loss reduction launch failed: ") + hipGetErrorString(err);
  }
  if (!backward && p.tile_cost && p.tile_order) return launch_tile_order(p.tile_cost, n, p.tile_order, stream);
  return "";
}

std::string launch_fill_int(int *ptr, int n, int v, hipStream_t stream) {
  hipLaunchKernelGGL(k_fill_int, dim3((n + 255) / 256), dim3(256), 0, stream, ptr, n, v);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return std::string("fill launch failed: ") + hipGetErrorString(e);
  return "";
}

}  // namespace tt
