// LSD radix sort of (uint32 key, uint32 value) pairs, 8-bit digits: per-block LDS
// digit histograms, a one-block exclusive scan of the [digit][block] counts, and a
// stable scatter ranked within each wave by ballot peer masks (wave64). Used by the
// LBVH build (Morton codes) and the backward's gradient reduce (triangle ids). The
// element count may live in device memory (n_dev) so that a launch sized for a
// capacity sorts what a previous kernel produced without a host round trip.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tt {
namespace {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortThreads * kSortItems;

__device__ __forceinline__ int sort_count(int n, const int *n_dev) {
  return n_dev ? min(*n_dev, n) : n;
}

__global__ __launch_bounds__(kSortThreads) void k_hist(const uint32_t *__restrict__ keys, int n, const int *n_dev,
                                                        int shift, uint32_t *__restrict__ hist, int nblocks) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  n = sort_count(n, n_dev);
  int base = blockIdx.x * kSortTile;
  int end = min(base + kSortTile, n);
  for (int i = base + threadIdx.x; i < end; i += kSortThreads) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
  __syncthreads();
  hist[(size_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of `n` counts in place, one block of 1024 threads.
__global__ __launch_bounds__(1024) void k_scan(uint32_t *data, int n) {
  __shared__ uint32_t s[1024];
  int per = (n + 1023) / 1024;
  int b = threadIdx.x * per;
  int e = min(b + per, n);
  uint32_t sum = 0;
  for (int i = b; i < e; i++) sum += data[i];
  s[threadIdx.x] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    uint32_t v = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0u;
    __syncthreads();
    s[threadIdx.x] += v;
    __syncthreads();
  }
  uint32_t run = s[threadIdx.x] - sum;
  for (int i = b; i < e; i++) {
    uint32_t v = data[i];
    data[i] = run;
    run += v;
  }
}

// Exclusive scan of n counts in place over many blocks (the digit counts of a sort of tens
// of millions of keys): kScanChunk counts per block scanned locally with the chunk total to
// totals[block], the totals scanned by k_scan, then added back.
constexpr int kScanChunk = 256 * 16;
__global__ __launch_bounds__(256) void k_scan_local(uint32_t *data, int n, uint32_t *totals) {
  __shared__ uint32_t s[256];
  const int base = blockIdx.x * kScanChunk + threadIdx.x * 16;
  uint32_t v[16], sum = 0;
  for (int k = 0; k < 16; k++) {
    v[k] = base + k < n ? data[base + k] : 0u;
    sum += v[k];
  }
  s[threadIdx.x] = sum;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const uint32_t x = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0u;
    __syncthreads();
    s[threadIdx.x] += x;
    __syncthreads();
  }
  uint32_t run = s[threadIdx.x] - sum;
  for (int k = 0; k < 16; k++) {
    if (base + k < n) data[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 255) totals[blockIdx.x] = s[255];
}
__global__ __launch_bounds__(256) void k_scan_add(uint32_t *data, int n, const uint32_t *offs) {
  const uint32_t o = offs[blockIdx.x];
  const int base = blockIdx.x * kScanChunk;
  for (int k = threadIdx.x; k < kScanChunk; k += 256)
    if (base + k < n) data[base + k] += o;
}

__global__ __launch_bounds__(kSortThreads) void k_scatter(const uint32_t *__restrict__ kin,
                                                           const uint32_t *__restrict__ vin, uint32_t *kout,
                                                           uint32_t *vout, int n, const int *n_dev, int shift,
                                                           const uint32_t *__restrict__ offs, int nblocks) {
  constexpr int W = kSortThreads / 64;
  __shared__ uint32_t base[256];
  __shared__ uint32_t cnt[256];
  __shared__ uint32_t wcnt[W][256];
  n = sort_count(n, n_dev);
  if (blockIdx.x * kSortTile >= n) return;  // uniform: past the count
  int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  base[tid] = offs[(size_t)tid * nblocks + blockIdx.x];
  cnt[tid] = 0;
  for (int w = 0; w < W; w++) wcnt[w][tid] = 0;
  __syncthreads();
  const unsigned long long lt_mask = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  for (int r = 0; r < kSortItems; r++) {
    int i = blockIdx.x * kSortTile + r * kSortThreads + tid;
    bool valid = i < n;
    uint32_t key = valid ? kin[i] : 0u;
    uint32_t val = valid ? vin[i] : 0u;
    uint32_t dg = (key >> shift) & 255u;
    unsigned long long peers = __ballot(valid);
    for (int bit = 0; bit < 8; bit++) {
      unsigned long long bb = __ballot(valid && ((dg >> bit) & 1u));
      peers &= ((dg >> bit) & 1u) ? bb : ~bb;
    }
    int rank = __popcll(peers & lt_mask);
    if (valid && rank == 0) wcnt[wave][dg] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t pos = base[dg] + cnt[dg] + (uint32_t)rank;
      for (int w = 0; w < wave; w++) pos += wcnt[w][dg];
      kout[pos] = key;
      vout[pos] = val;
    }
    __syncthreads();
    uint32_t add = 0;
    for (int w = 0; w < W; w++) add += wcnt[w][tid], wcnt[w][tid] = 0;
    cnt[tid] += add;
    __syncthreads();
  }
}

}  // namespace
}  // namespace tt
