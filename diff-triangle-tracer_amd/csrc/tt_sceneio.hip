// Scene I/O on the device: the element bodies of a binary little-endian PLY scene file
// (vertex rows and face rows with per-face splat properties) to and from the tracer's
// layout. The reference has no scene I/O (its callers keep scenes as torch tensors); the
// file layout is this build's specification, declared in include/tritracer_hip.h.
//
// Face rows are byte-packed (213 bytes for 16 SH coefficients), so no field is aligned.
// Each block stages the byte range of its faces through LDS with aligned dword global
// vector accesses (the LDS image keeps the global 16-byte granules: byte g sits at LDS
// byte g - (first & ~15)), reads / writes the fields there, and moves the outputs
// with coalesced stores. Pure data movement: bit-exact by construction.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <string>

#include "../../include/tritracer_hip.h"
#include "tt_internal.h"

namespace {

constexpr int kIoThreads = 256;
constexpr int kIoMaxFaces = 128;        // faces per block (fewer for very wide rows)
constexpr int kIoLdsBytes = 32768;      // staged face bytes per block
constexpr int kShFields = TT_SCENE_MAX_COEFFS * 3;

// Little-endian float at LDS byte `off` of a word array: two aligned dword reads joined by
// a byte-granular funnel shift (the word after the last staged one is allocated and only
// contributes bytes that are not used).
__device__ __forceinline__ float lds_f32(const unsigned *w, int off) {
  const int i = off >> 2;
  return __uint_as_float(__builtin_amdgcn_alignbyte(w[i + 1], w[i], (unsigned)(off & 3)));
}
__device__ __forceinline__ float global_f32_bytes(const unsigned char *p) {
  const unsigned b = (unsigned)p[0] | (unsigned)p[1] << 8 | (unsigned)p[2] << 16 | (unsigned)p[3] << 24;
  return __uint_as_float(b);
}

struct UnpackIn {
  tt_scene_unpack_args a;
  int faces_per_block;
  int vert_aligned;  // vertex rows and xyz offsets are 4-byte aligned: dword gathers
};

__global__ __launch_bounds__(kIoThreads) void k_scene_unpack(UnpackIn U) {
  __shared__ __align__(16) unsigned s_words[kIoLdsBytes / 4 + 8];
  __shared__ int s_off[kShFields];
  __shared__ int s_idx[3 * kIoMaxFaces];
  const tt_scene_unpack_args &a = U.a;
  const int t = threadIdx.x;
  const int f0 = blockIdx.x * U.faces_per_block;
  const int nf = min(U.faces_per_block, a.P - f0);
  const long long stride = a.face_stride;
  const long long b0 = f0 * stride, b1 = b0 + nf * stride, end_all = (long long)a.P * stride;
  const long long w0 = b0 & ~15ll;  // 16-byte granules (the buffer is 16-byte aligned)
  const int shift = (int)(b0 - w0);
  const int n_quads = (int)((b1 - w0 + 15) >> 4);
  for (int q = t; q < n_quads; q += kIoThreads) {
    const long long g = w0 + 16ll * q;
    uint4 v;
    if (g + 16 <= end_all) {
      v = *reinterpret_cast<const uint4 *>(a.face_dev + g);
    } else {
      unsigned w[4] = {0u, 0u, 0u, 0u};
      for (int k = 0; k < 16; k++)
        if (g + k < end_all) w[k >> 2] |= (unsigned)a.face_dev[g + k] << (8 * (k & 3));
      v = make_uint4(w[0], w[1], w[2], w[3]);
    }
    reinterpret_cast<uint4 *>(s_words)[q] = v;
  }
  for (int k = t; k < kShFields; k += kIoThreads) s_off[k] = k < 3 * a.M ? a.sh_offset[k] : -1;
  __syncthreads();
  const unsigned char *s = reinterpret_cast<const unsigned char *>(s_words);
  // per face: opacity, sigma, index validation, npts / cumsum
  for (int f = t; f < nf; f += kIoThreads) {
    const int row = shift + f * (int)stride;
    const long long fi = f0 + f;
    a.opacity_out_dev[fi] = a.opacity_offset >= 0 ? lds_f32(s_words, row + a.opacity_offset) : 0.0f;
    a.sigma_out_dev[fi] = a.sigma_offset >= 0 ? lds_f32(s_words, row + a.sigma_offset) : 0.0f;
    bool ok = s[row + a.count_offset] == 3;
    int idx[3];
    for (int j = 0; j < 3; j++) {
      idx[j] = __float_as_int(lds_f32(s_words, row + a.index_offset + 4 * j));
      ok = ok && idx[j] >= 0 && (long long)idx[j] < a.V;
    }
    for (int j = 0; j < 3; j++) s_idx[3 * f + j] = ok ? idx[j] : -1;
    if (!ok && a.bad_dev) atomicAdd(a.bad_dev, 1);
    if (a.npts_out_dev) a.npts_out_dev[fi] = 3;
    if (a.cumsum_out_dev) a.cumsum_out_dev[fi] = (int)(3 * fi);
  }
  // SH: output-major, coefficient j of channel c at shs[face][j][c]
  const int per = 3 * a.M;
  for (int e = t; e < nf * per; e += kIoThreads) {
    const int f = e / per, k = e - f * per;
    const int off = s_off[k];
    a.shs_out_dev[(long long)f0 * per + e] = off >= 0 ? lds_f32(s_words, shift + f * (int)stride + off) : 0.0f;
  }
  __syncthreads();
  // vertices gathered through the indices: output-major [face][corner][xyz]
  for (int e = t; e < nf * 9; e += kIoThreads) {
    const int f = e / 9, r = e - 9 * f, j = r / 3, comp = r - 3 * j;
    const int idx = s_idx[3 * f + j];
    float v = 0.0f;
    if (idx >= 0) {
      const int xo = comp == 0 ? a.xyz_offset[0] : (comp == 1 ? a.xyz_offset[1] : a.xyz_offset[2]);
      const unsigned char *p = a.vert_dev + (long long)idx * a.vert_stride + xo;
      v = U.vert_aligned ? *reinterpret_cast<const float *>(p) : global_f32_bytes(p);
    }
    a.vertices_out_dev[9ll * f0 + e] = v;
  }
}

// Faces per block of the pack kernel: its LDS holds nfield + 1 words per face.
__host__ __device__ constexpr int pack_faces(int M) {
  return (kIoLdsBytes / (4 * (6 + 3 * M))) < kIoMaxFaces ? kIoLdsBytes / (4 * (6 + 3 * M)) : kIoMaxFaces;
}

__global__ __launch_bounds__(kIoThreads) void k_scene_pack(tt_scene_pack_args a) {
  // Per face, nfield + 1 aligned words: word 0 = 3 << 24 (its last byte is the row's count
  // byte), then the fields (3 indices, opacity, sigma, f_dc_0..2, f_rest_{c (M-1) + j - 1}).
  // Byte r of a row is byte (r + 3) & 3 of word (r + 3) >> 2, so any 4 bytes of one row are
  // one funnel shift of two words.
  __shared__ unsigned s_v[kIoLdsBytes / 4 + 2];
  const int t = threadIdx.x;
  const int S = TT_SCENE_FACE_BYTES(a.M);
  const int nfield = 5 + 3 * a.M, nw = nfield + 1;
  const int F = pack_faces(a.M);
  const int f0 = blockIdx.x * F;
  const int nf = min(F, a.P - f0);
  // head words of each face: count word, indices, opacity, sigma
  for (int e = t; e < nf * 6; e += kIoThreads) {
    const int f = e / 6, k = e - 6 * f;
    const long long fi = f0 + f;
    unsigned v = 3u << 24;
    if (k >= 1 && k <= 3) v = (unsigned)(3 * fi + k - 1);
    if (k == 4) v = __float_as_uint(a.opacity_dev[fi]);
    if (k == 5) v = __float_as_uint(a.sigma_dev[fi]);
    s_v[f * nw + k] = v;
  }
  // SH read in source order (coalesced), scattered to its field: coefficient j, channel c
  // -> f_dc_c (j = 0) or f_rest_{c (M-1) + j - 1}
  const int per = 3 * a.M;
  auto put_sh = [&](int e, float v) {
    const int f = e / per, rem = e - f * per, j = rem / 3, c = rem - 3 * j;
    const int k = j == 0 ? 5 + c : 8 + c * (a.M - 1) + (j - 1);
    s_v[f * nw + 1 + k] = __float_as_uint(v);
  };
  const float *sh = a.shs_dev + (long long)f0 * per;
  const int n_sh = nf * per;
  if ((reinterpret_cast<uintptr_t>(sh) & 15u) == 0) {  // 16-byte loads, then the tail
    for (int e4 = t; e4 < n_sh / 4; e4 += kIoThreads) {
      const float4 v = reinterpret_cast<const float4 *>(sh)[e4];
      put_sh(4 * e4, v.x);
      put_sh(4 * e4 + 1, v.y);
      put_sh(4 * e4 + 2, v.z);
      put_sh(4 * e4 + 3, v.w);
    }
    for (int e = (n_sh & ~3) + t; e < n_sh; e += kIoThreads) put_sh(e, sh[e]);
  } else {
    for (int e = t; e < n_sh; e += kIoThreads) put_sh(e, sh[e]);
  }
  __syncthreads();
  const long long b0 = (long long)f0 * S, b1 = b0 + (long long)nf * S;
  // the four bytes at row offset r of face f; bytes past the row's end come from the next
  // row's start (only asked for where that row is in this block)
  auto word_fr = [&](int f, int r) -> unsigned {
    const int u = (r + 3) >> 2, m = (r + 3) & 3;
    const unsigned *w = s_v + f * nw;
    const unsigned lo = __builtin_amdgcn_alignbyte(u + 1 < nw ? w[u + 1] : 0u, w[u], (unsigned)m);
    const int n = S - r;  // bytes of this row from r on
    if (n >= 4) return lo;
    const unsigned hi = __builtin_amdgcn_alignbyte(w[nw + 1], w[nw], 3u);  // next row's bytes 0..3
    return (lo & ((1u << (8 * n)) - 1u)) | hi << (8 * n);
  };
  // the block owns bytes [b0, b1): whole 16-byte granules with vector stores, the partial
  // granules at either end byte by byte (neighbouring blocks own their other bytes)
  const long long q_lo = (b0 + 15) & ~15ll, q_hi = b1 & ~15ll;
  for (long long g = q_lo + 16ll * t; g < q_hi; g += 16ll * kIoThreads) {
    const int o = (int)(g - b0);
    int f = o / S, r = o - f * S;
    unsigned w4[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      w4[k] = word_fr(f, r);
      r += 4;
      if (r >= S) {
        r -= S;
        f++;
      }
    }
    *reinterpret_cast<uint4 *>(a.face_out_dev + g) = make_uint4(w4[0], w4[1], w4[2], w4[3]);
  }
  const long long e_lo = q_lo < q_hi ? q_lo : b1, e_hi = q_lo < q_hi ? q_hi : b1;
  for (long long g = b0 + t; g < b1; g += kIoThreads) {
    if (g >= e_lo && g < e_hi) continue;
    const int o = (int)(g - b0), f = o / S;
    a.face_out_dev[g] = (unsigned char)(word_fr(f, o - f * S) & 0xFFu);
  }
  // vertex rows: triangle i's points (read through cumsum) at rows 3i .. 3i + 2
  for (int e = t; e < nf * 9; e += kIoThreads) {
    const int f = e / 9, r = e - 9 * f;
    a.vert_out_dev[9ll * f0 + e] = a.vertices_dev[3ll * a.cumsum_dev[f0 + f] + r];
  }
}

int fail(const std::string &m) {
  tt::set_last_error(m);
  return TT_ERR_INVALID_ARG;
}

bool aligned4(const void *q) { return (reinterpret_cast<uintptr_t>(q) & 3u) == 0; }
bool aligned16(const void *q) { return (reinterpret_cast<uintptr_t>(q) & 15u) == 0; }

int launch_status(const char *what) {
  const hipError_t e = hipGetLastError();
  if (e == hipSuccess) return TT_OK;
  tt::set_last_error(std::string(what) + " launch failed: " + hipGetErrorString(e));
  return TT_ERR_HIP;
}

}  // namespace

extern "C" int tt_scene_unpack(const tt_scene_unpack_args *a, void *stream) {
  if (!a) return fail("tt_scene_unpack: args is NULL");
  if (a->P < 0 || a->V < 0) return fail("tt_scene_unpack: P and V must be non-negative");
  if (a->M < 1 || a->M > TT_SCENE_MAX_COEFFS) return fail("tt_scene_unpack: M must be in [1, 16]");
  if (a->P == 0) return TT_OK;
  if ((long long)a->P * 3 > INT_MAX) return fail("tt_scene_unpack: 3 P overflows cumsum");
  if (!a->face_dev || !a->vertices_out_dev || !a->opacity_out_dev || !a->sigma_out_dev || !a->shs_out_dev)
    return fail("tt_scene_unpack: face buffer or an output is NULL");
  if (a->V > 0 && !a->vert_dev) return fail("tt_scene_unpack: vertex buffer is NULL");
  if (!aligned16(a->face_dev)) return fail("tt_scene_unpack: the face buffer must be 16-byte aligned");
  if (a->face_stride < 13 || a->face_stride > kIoLdsBytes)
    return fail("tt_scene_unpack: face_stride must be in [13, " + std::to_string(kIoLdsBytes) + "]");
  auto field = [&](int off, int bytes) { return off >= 0 && off + bytes <= a->face_stride; };
  if (!field(a->count_offset, 1) || !field(a->index_offset, 12))
    return fail("tt_scene_unpack: the vertex_indices list lies outside the face row");
  if ((a->opacity_offset >= 0 && !field(a->opacity_offset, 4)) || (a->sigma_offset >= 0 && !field(a->sigma_offset, 4)))
    return fail("tt_scene_unpack: opacity / sigma lie outside the face row");
  for (int k = 0; k < 3 * a->M; k++)
    if (a->sh_offset[k] >= 0 && !field(a->sh_offset[k], 4))
      return fail("tt_scene_unpack: SH field " + std::to_string(k) + " lies outside the face row");
  bool valigned = aligned4(a->vert_dev) && a->vert_stride % 4 == 0;
  for (int c = 0; c < 3; c++) {
    if (a->xyz_offset[c] < 0 || a->xyz_offset[c] + 4 > a->vert_stride)
      return fail("tt_scene_unpack: x, y, z must lie inside the vertex row");
    valigned = valigned && a->xyz_offset[c] % 4 == 0;
  }
  UnpackIn U;
  U.a = *a;
  U.faces_per_block = std::min(kIoMaxFaces, kIoLdsBytes / a->face_stride);
  U.vert_aligned = valigned ? 1 : 0;
  const int blocks = (a->P + U.faces_per_block - 1) / U.faces_per_block;
  hipLaunchKernelGGL(k_scene_unpack, dim3(blocks), dim3(kIoThreads), 0, static_cast<hipStream_t>(stream), U);
  return launch_status("scene unpack");
}

extern "C" int tt_scene_pack(const tt_scene_pack_args *a, void *stream) {
  if (!a) return fail("tt_scene_pack: args is NULL");
  if (a->P < 0) return fail("tt_scene_pack: P must be non-negative");
  if (a->M < 1 || a->M > TT_SCENE_MAX_COEFFS) return fail("tt_scene_pack: M must be in [1, 16]");
  if (a->P == 0) return TT_OK;
  if ((long long)a->P * 3 > INT_MAX) return fail("tt_scene_pack: 3 P overflows the indices");
  if (!a->vertices_dev || !a->cumsum_dev || !a->opacity_dev || !a->sigma_dev || !a->shs_dev || !a->face_out_dev ||
      !a->vert_out_dev)
    return fail("tt_scene_pack: NULL buffer");
  if (!aligned16(a->face_out_dev)) return fail("tt_scene_pack: the face buffer must be 16-byte aligned");
  const int F = pack_faces(a->M);
  const int blocks = (a->P + F - 1) / F;
  hipLaunchKernelGGL(k_scene_pack, dim3(blocks), dim3(kIoThreads), 0, static_cast<hipStream_t>(stream), *a);
  return launch_status("scene pack");
}
