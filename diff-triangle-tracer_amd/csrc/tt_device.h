// Device-side math of the triangle-splat tracer (gfx950).
//
// Everything here is compiled with -ffp-contract=off: the arithmetic rounds
// where the CPU oracle's FP32 path rounds, so hit lists and per-hit decisions
// compare bit-exactly. The formulas restate the reference:
//   window function ........ tritracer/common.h:57-114 (computeWindowFunc)
//   SH colour .............. tritracer/common.h:16-55 (computeColorFromSH)
//   SH / window backward ... tritracer/backward.cu:25-189, vecmath.h:2785-2859
//   sphere helpers ......... tritracer/common.h:116-148, reflect.cu, refract.cu, envlight.cu
// The ray/triangle test (inside optixTrace in the reference) is the watertight
// test of Woop, Benthin and Wald (JCGT 2013), double-sided.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define TT_D __device__ __forceinline__

namespace tt {

constexpr uint32_t kInvalid = 0xFFFFFFFFu;  // reference RayHit::invalid_id (params.h:119)
constexpr float kInfDist = 1e6f;            // reference RayHit::infinite_dist (params.h:118)
constexpr double kStepEps = 1e-9;           // reference STEP_EPSILON (config.h:11), a double literal

// SH constants (auxiliary.h:33-50): real spherical-harmonic normalisations.
constexpr float SH_C0 = 0.28209479177387814f;
constexpr float SH_C1 = 0.4886025119029199f;
constexpr float SH_C2_0 = 1.0925484305920792f, SH_C2_1 = -1.0925484305920792f, SH_C2_2 = 0.31539156525252005f,
                SH_C2_3 = -1.0925484305920792f, SH_C2_4 = 0.5462742152960396f;
constexpr float SH_C3_0 = -0.5900435899266435f, SH_C3_1 = 2.890611442640554f, SH_C3_2 = -0.4570457994644658f,
                SH_C3_3 = 0.3731763325901154f, SH_C3_4 = -0.4570457994644658f, SH_C3_5 = 1.445305721320277f,
                SH_C3_6 = -0.5900435899266435f;

// ------------------------------------------------------------------ vectors
struct v3f {
  float x, y, z;
};

// Pinhole camera ray of pixel (x, y) of a W x H image: cam = eye[3] forward[3] right[3] up[3]
// tan_x tan_y (double). The direction forward + xs right + ys up with
// xs = (2 (x + 0.5) / W - 1) tan_x, ys = (1 - 2 (y + 0.5) / H) tan_y is formed and normalised
// in double in this fixed order and rounded once to float (tt_camera_rays and the fused ray
// generation of the multi-camera launches both use this routine).
constexpr int kCamDoubles = 14;
__device__ __forceinline__ void pinhole_ray(const double *cam, int W, int H, int x, int y, float o[3], float d[3]) {
  const double xs = (2.0 * ((double)x + 0.5) / (double)W - 1.0) * cam[12];
  const double ys = (1.0 - 2.0 * ((double)y + 0.5) / (double)H) * cam[13];
  double v[3];
  for (int i = 0; i < 3; i++) v[i] = (cam[3 + i] + xs * cam[6 + i]) + ys * cam[9 + i];
  const double n = sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
  for (int i = 0; i < 3; i++) {
    d[i] = (float)(v[i] / n);
    o[i] = (float)cam[i];
  }
}
struct v3d {
  double x, y, z;
};
TT_D v3f mk(float x, float y, float z) {
  return v3f{x, y, z};
}
TT_D v3f operator+(v3f a, v3f b) {
  return {a.x + b.x, a.y + b.y, a.z + b.z};
}
TT_D v3f operator-(v3f a, v3f b) {
  return {a.x - b.x, a.y - b.y, a.z - b.z};
}
TT_D v3f operator-(v3f a) {
  return {-a.x, -a.y, -a.z};
}
TT_D v3f operator*(v3f a, float s) {
  return {a.x * s, a.y * s, a.z * s};
}
TT_D v3f operator*(float s, v3f a) {
  return {s * a.x, s * a.y, s * a.z};
}
TT_D v3f hmul(v3f a, v3f b) {
  return {a.x * b.x, a.y * b.y, a.z * b.z};
}
TT_D float dot(v3f a, v3f b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
TT_D v3f cross(v3f a, v3f b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
TT_D float vlen(v3f v) {
  return sqrtf(dot(v, v));
}
TT_D v3f normalize(v3f v) {
  float inv = 1.0f / sqrtf(dot(v, v));
  return v * inv;
}
TT_D v3f divs(v3f a, float s) {
  float inv = 1.0f / s;
  return a * inv;
}
TT_D float sumv(v3f a) {
  return a.x + a.y + a.z;
}
TT_D float comp(v3f v, int i) {
  return i == 0 ? v.x : (i == 1 ? v.y : v.z);
}
TT_D v3d operator+(v3d a, v3d b) {
  return {a.x + b.x, a.y + b.y, a.z + b.z};
}
TT_D v3d operator-(v3d a, v3d b) {
  return {a.x - b.x, a.y - b.y, a.z - b.z};
}
TT_D v3d operator*(v3d a, double s) {
  return {a.x * s, a.y * s, a.z * s};
}
TT_D double dot(v3d a, v3d b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
TT_D v3d cross(v3d a, v3d b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
TT_D v3d todbl(v3f a) {
  return {(double)a.x, (double)a.y, (double)a.z};
}
TT_D v3f toflt(v3d a) {
  return {(float)a.x, (float)a.y, (float)a.z};
}
TT_D v3f ld3(const float *p) {
  return {p[0], p[1], p[2]};
}

// ---------------------------------------------- project-written FP32 functions
// log2: x = m 2^e, m in [sqrt(1/2), sqrt(2)); ln m = 2 atanh((m-1)/(m+1)) to s^9.
// Arguments below the smallest normal count as zero.
TT_D float log2_tt(float x) {
  if (!(x > 0.0f)) return (x == 0.0f) ? -__builtin_huge_valf() : __builtin_nanf("");
  if (x == __builtin_huge_valf()) return x;
  if (x < 1.17549435e-38f) return -__builtin_huge_valf();
  uint32_t b = __float_as_uint(x);
  int e = (int)((b >> 23) & 0xffu) - 127;
  float m = __uint_as_float((b & 0x7fffffu) | 0x3f800000u);
  if (m > 1.41421356f) {
    m = m * 0.5f;
    e = e + 1;
  }
  float f = m - 1.0f;
  float s = f / (2.0f + f);
  float z = s * s;
  float r = z * (0.666666667f + z * (0.4f + z * (0.285714286f + z * 0.222222222f)));
  float ln = (s + s) + s * r;
  return (float)e + ln * 1.44269504f;
}
// exp2: n = rint(y), f = y - n, e^(f ln2) Taylor to degree 7, exact 2^n scaling;
// results below 2^-126 flush to zero.
TT_D float exp2_tt(float y) {
  if (y != y) return y;
  if (y >= 128.0f) return __builtin_huge_valf();
  if (y < -126.0f) return 0.0f;
  float n = rintf(y);
  float f = y - n;
  float g = f * 0.693147181f;
  float p = 1.0f +
            g * (1.0f +
                 g * (0.5f + g * (0.166666667f +
                                  g * (0.0416666667f + g * (0.00833333333f + g * (0.00138888889f + g * 0.000198412698f))))));
  int ni = (int)n;
  int h = ni / 2;
  float s1 = __uint_as_float((uint32_t)(h + 127) << 23);
  float s2 = __uint_as_float((uint32_t)(ni - h + 127) << 23);
  return (p * s1) * s2;
}
TT_D float pow_tt(float x, float y) {
  return exp2_tt(y * log2_tt(x));
}
TT_D float log_tt(float x) {
  return log2_tt(x) * 0.693147181f;
}
TT_D void sincos_tt(float x, float &sn, float &cs) {
  float q = rintf(x * 0.636619772f);
  float r = (x - q * 1.5703125f) - q * 4.83826794e-4f;
  float z = r * r;
  float s = r + r * z * (-0.166666667f + z * (0.00833333333f + z * (-0.000198412698f + z * 2.75573192e-6f)));
  float c = 1.0f + z * (-0.5f + z * (0.0416666667f + z * (-0.00138888889f + z * (2.48015873e-5f + z * -2.75573192e-7f))));
  int qi = ((int)q) & 3;
  if (qi == 0) {
    sn = s, cs = c;
  } else if (qi == 1) {
    sn = c, cs = -s;
  } else if (qi == 2) {
    sn = -s, cs = -c;
  } else {
    sn = -c, cs = s;
  }
}

// ------------------------------------------------------------ intersection
struct RayPre {
  v3f o, d, inv;
  int kx, ky, kz;
  float Sx, Sy, Sz;
};
TT_D RayPre prep_ray(v3f o, v3f d) {
  RayPre r;
  r.o = o;
  r.d = d;
  float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  r.kz = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
  r.kx = r.kz + 1;
  if (r.kx == 3) r.kx = 0;
  r.ky = r.kx + 1;
  if (r.ky == 3) r.ky = 0;
  if (comp(d, r.kz) < 0.0f) {
    int tmp = r.kx;
    r.kx = r.ky;
    r.ky = tmp;
  }
  float dz = comp(d, r.kz);
  r.Sx = comp(d, r.kx) / dz;
  r.Sy = comp(d, r.ky) / dz;
  r.Sz = 1.0f / dz;
  float sx = fabsf(d.x) > 1e-30f ? d.x : copysignf(1e-30f, d.x);
  float sy = fabsf(d.y) > 1e-30f ? d.y : copysignf(1e-30f, d.y);
  float sz = fabsf(d.z) > 1e-30f ? d.z : copysignf(1e-30f, d.z);
  r.inv = mk(1.0f / sx, 1.0f / sy, 1.0f / sz);
  return r;
}

// Woop, Benthin, Wald, "Watertight Ray/Triangle Intersection", JCGT 2(1) 2013.
TT_D bool intersect(const RayPre &r, v3f v0, v3f v1, v3f v2, float &t) {
  v3f A = v0 - r.o, B = v1 - r.o, C = v2 - r.o;
  float Akz = comp(A, r.kz), Bkz = comp(B, r.kz), Ckz = comp(C, r.kz);
  float Ax = comp(A, r.kx) - r.Sx * Akz;
  float Ay = comp(A, r.ky) - r.Sy * Akz;
  float Bx = comp(B, r.kx) - r.Sx * Bkz;
  float By = comp(B, r.ky) - r.Sy * Bkz;
  float Cx = comp(C, r.kx) - r.Sx * Ckz;
  float Cy = comp(C, r.ky) - r.Sy * Ckz;
  float U = Cx * By - Cy * Bx;
  float V = Ax * Cy - Ay * Cx;
  float W = Bx * Ay - By * Ax;
  if (U == 0.0f || V == 0.0f || W == 0.0f) {
    U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
    V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
    W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
  }
  if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
  float det = U + V + W;
  if (det == 0.0f) return false;
  float Az = r.Sz * Akz, Bz = r.Sz * Bkz, Cz = r.Sz * Ckz;
  float T = U * Az + V * Bz + W * Cz;
  t = T / det + 0.0f;  // -0.0 as +0.0: hit keys compare t as unsigned bits
  return true;
}

// Chunk key: a candidate must satisfy t <= far and (t > tlo or (t == tlo and prim > plo)).
// (bitwise | and &: no short-circuit branches in the compiled code)
TT_D bool after_key(float t, uint32_t prim, float tlo, long long plo) {
  return (t > tlo) | ((t == tlo) & ((long long)prim > plo));
}

// Conservative slab test: boxes are padded at build time and the interval is
// widened, so a triangle whose computed t lies in [tlo, thi] is never culled.
// Plane distances as (plane - o) * (1/d): relative error of a few ulp of t itself,
// covered by the widening below (the fma form (plane*(1/d) - o/d) errs by ~ulp of
// o/d, which for a distant origin exceeds the box padding).
TT_D void slab(const RayPre &r, float lx, float hx, float ly, float hy, float lz, float hz, float &tn, float &tf) {
  const float t0x = (lx - r.o.x) * r.inv.x, t1x = (hx - r.o.x) * r.inv.x;
  const float t0y = (ly - r.o.y) * r.inv.y, t1y = (hy - r.o.y) * r.inv.y;
  const float t0z = (lz - r.o.z) * r.inv.z, t1z = (hz - r.o.z) * r.inv.z;
  tn = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fminf(t0z, t1z));
  tf = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fmaxf(t0z, t1z));
  tf = fmaf(fabsf(tf), 1e-6f, tf) + 1e-7f;
  tn = fmaf(fabsf(tn), -1e-6f, tn) - 1e-7f;
}
TT_D bool box_ok(float tn, float tf, float tlo_c, float thi_c) {
  return tn <= tf && tf >= tlo_c && tn <= thi_c;
}
// The test slab() + box_ok() make, with the widening moved out of the per-box work: one
// FMA per box instead of two FMAs and two adds. Widening constants kWideRel / kWideAbs
// exceed twice slab()'s (1e-6, 1e-7), which keeps every box slab() + box_ok() accept
// (tlo >= 0 and thi >= 0 always; a box it accepts has tf >= -2e-7): with the widened
// interval [tn - |tn| e - c, tf + |tf| e + c] meeting [tlo_c, thi_c],
//   tf >= tlo (1 - kWideRel) - kWideAbs, tn <= thi (1 + kWideRel) + kWideAbs and
//   tn <= tf (1 + kWideRel) + kWideAbs
// all hold. The per-node bounds lo_w / hi_w are tlo / thi widened by the same constants.
constexpr float kWideRel = 2.5e-6f, kWideAbs = 2.5e-7f;
TT_D bool box_ok_wide(float tn, float tf, float lo_w, float hi_w) {
  return (tn <= fmaf(tf, 1.0f + kWideRel, kWideAbs)) & (tf >= lo_w) & (tn <= hi_w);
}

// ------------------------------------------------------------- window function
struct Win {
  float Cx, phi_final, phi_max, dist;
  float ed[3];
  v3f n;
  v3f rel[3];  // hit - vertex i
};

// computeWindowFunc (common.h:57-114) in the kernels' FP32 form: relative to
// vertex A, with the three edge-distance numerators ec x (p_i - p_{i+1}) . (hit - p_i)
// formed in FP64 (add/sub/mul only) -- in FP32 that dot product cancels to
// ~eps*|hit - p_i| absolute, which the window's diverging slope near an edge
// amplifies into percent-level gradient error.
//
// Split in two: the terms of the triangle alone (the FP64 edge-plane normals n_i, their
// FP32 inverse lengths and the incenter's signed distance to edge 0) and the terms of one
// hit. A forward computes the first part once per triangle into a 128-byte record
// (window_record) and each hit reads it (window_hit); window32 runs both parts in turn. The
// operations are the same either way, so the results are identical bit for bit.
struct TriWin {
  v3d n[3];     // ec x (p_i - p_{i+1}) in FP64, ec = (B - A) x (C - A)
  float il[3];  // 1 / |toflt(n_i)|
  float dist;   // dot(toflt(n_0) il_0, incenter - A)
};
TT_D void window_tri(v3f A, v3f B, v3f C, TriWin &tw) {
  v3f eAB = B - A, eAC = C - A, eBC = C - B;
  float a = vlen(eBC);
  float b = vlen(eAC);
  float c = vlen(eAB);
  v3f incA = divs(b * eAB + c * eAC, (a + b) + c);
  v3d Ad = todbl(A), Bd = todbl(B), Cd = todbl(C);
  v3d ecd = cross(Bd - Ad, Cd - Ad);
  tw.n[0] = cross(ecd, Ad - Bd);
  tw.n[1] = cross(ecd, Bd - Cd);
  tw.n[2] = cross(ecd, Cd - Ad);
  v3f m0 = toflt(tw.n[0]), m1 = toflt(tw.n[1]), m2 = toflt(tw.n[2]);
  tw.il[0] = 1.0f / sqrtf(dot(m0, m0));
  tw.il[1] = 1.0f / sqrtf(dot(m1, m1));
  tw.il[2] = 1.0f / sqrtf(dot(m2, m2));
  tw.dist = dot(m0 * tw.il[0], incA);
}
// The unit face normal the forward accumulates for the "others" output.
TT_D v3f face_normal(v3f A, v3f B, v3f C) { return normalize(cross(B - A, C - A)); }
// The terms of one hit at distance t along (o, d): everything of Win except n.
TT_D void window_hit(const TriWin &tw, v3f A, v3f B, v3f C, v3f o, v3f d, float t, float sigma, Win &w) {
  v3d Ad = todbl(A), Bd = todbl(B), Cd = todbl(C);
  v3d od = todbl(o), td = todbl(d) * (double)t;
  v3d rA = (od - Ad) + td, rB = (od - Bd) + td, rC = (od - Cd) + td;
  float q0 = (float)dot(tw.n[0], rA), q1 = (float)dot(tw.n[1], rB), q2 = (float)dot(tw.n[2], rC);
  float dist = tw.dist;
  float ed0 = q0 * tw.il[0];
  if (dist > 0.0f) {
    dist = -dist;
    ed0 = -ed0;
  }
  w.ed[0] = ed0;
  w.ed[1] = q1 * tw.il[1];
  w.ed[2] = q2 * tw.il[2];
  w.rel[0] = toflt(rA);
  w.rel[1] = toflt(rB);
  w.rel[2] = toflt(rC);
  float phi = fmaxf(fmaxf(w.ed[0], w.ed[1]), w.ed[2]);
  w.phi_max = phi;
  w.dist = dist;
  w.phi_final = phi / dist;
  w.Cx = fmaxf(0.0f, pow_tt(w.phi_final, sigma));
}
TT_D void window32(v3f A, v3f B, v3f C, v3f o, v3f d, float t, float sigma, Win &w) {
  TriWin tw;
  window_tri(A, B, C, tw);
  w.n = face_normal(A, B, C);
  window_hit(tw, A, B, C, o, d, t, sigma, w);
}

// Window records: 8 float4 (128 B, one cache line) per triangle, written by the forward's
// record pass and read once per consumed hit:
//   [0] n0.x n0.y  [1] n0.z n1.x  [2] n1.y n1.z  [3] n2.x n2.y   (FP64 pairs)
//   [4] n2.z (FP64), A.x, A.y     [5] A.z, B.x, B.y, B.z
//   [6] C.x, C.y, C.z, il0        [7] il1, il2, dist, sigma
TT_D void window_record_store(float4 *q, v3f A, v3f B, v3f C, float sigma) {
  TriWin tw;
  window_tri(A, B, C, tw);
  double2 *qd = reinterpret_cast<double2 *>(q);
  qd[0] = make_double2(tw.n[0].x, tw.n[0].y);
  qd[1] = make_double2(tw.n[0].z, tw.n[1].x);
  qd[2] = make_double2(tw.n[1].y, tw.n[1].z);
  qd[3] = make_double2(tw.n[2].x, tw.n[2].y);
  const double z2 = tw.n[2].z;
  const uint64_t zb = __double_as_longlong(z2);
  q[4] = make_float4(__uint_as_float((uint32_t)zb), __uint_as_float((uint32_t)(zb >> 32)), A.x, A.y);
  q[5] = make_float4(A.z, B.x, B.y, B.z);
  q[6] = make_float4(C.x, C.y, C.z, tw.il[0]);
  q[7] = make_float4(tw.il[1], tw.il[2], tw.dist, sigma);
}
// A record as loaded (decoded by window_record_decode): lets a caller issue the loads of
// the next hit's record before it works on the current one.
struct WinRec {
  float4 q[8];
};
TT_D void window_record_fetch(const float4 *p, WinRec &r) {
#pragma unroll
  for (int j = 0; j < 8; j++) r.q[j] = p[j];
}
TT_D double dbl_of(float lo, float hi) {
  return __longlong_as_double((long long)((uint64_t)__float_as_uint(lo) | (uint64_t)__float_as_uint(hi) << 32));
}
TT_D void window_record_decode(const WinRec &r, TriWin &tw, v3f &A, v3f &B, v3f &C, float &sigma) {
  const float4 a = r.q[0], b = r.q[1], c = r.q[2], e = r.q[3], f = r.q[4], g = r.q[5], h = r.q[6], k = r.q[7];
  tw.n[0] = v3d{dbl_of(a.x, a.y), dbl_of(a.z, a.w), dbl_of(b.x, b.y)};
  tw.n[1] = v3d{dbl_of(b.z, b.w), dbl_of(c.x, c.y), dbl_of(c.z, c.w)};
  tw.n[2] = v3d{dbl_of(e.x, e.y), dbl_of(e.z, e.w), dbl_of(f.x, f.y)};
  A = mk(f.z, f.w, g.x);
  B = mk(g.y, g.z, g.w);
  C = mk(h.x, h.y, h.z);
  tw.il[0] = h.w;
  tw.il[1] = k.x;
  tw.il[2] = k.y;
  tw.dist = k.z;
  sigma = k.w;
}
TT_D void window_record_load(const float4 *q, TriWin &tw, v3f &A, v3f &B, v3f &C, float &sigma) {
  const double2 *qd = reinterpret_cast<const double2 *>(q);
  const double2 a = qd[0], b = qd[1], c = qd[2], e = qd[3];
  const float4 f = q[4], g = q[5], h = q[6], k = q[7];
  tw.n[0] = v3d{a.x, a.y, b.x};
  tw.n[1] = v3d{b.y, c.x, c.y};
  const double z2 = dbl_of(f.x, f.y);
  tw.n[2] = v3d{e.x, e.y, z2};
  A = mk(f.z, f.w, g.x);
  B = mk(g.y, g.z, g.w);
  C = mk(h.x, h.y, h.z);
  tw.il[0] = h.w;
  tw.il[1] = k.x;
  tw.il[2] = k.y;
  tw.dist = k.z;
  sigma = k.w;
}

// --------------------------------------------------------------- SH colour
TT_D v3f shc(const float *sh, int i) {
  return v3f{sh[3 * i], sh[3 * i + 1], sh[3 * i + 2]};
}
// computeColorFromSH (common.h:16-55).
TT_D v3f sh_color(const float *sh, int deg, v3f dir, bool &c0, bool &c1, bool &c2) {
  dir = divs(dir, vlen(dir));
  v3f res = SH_C0 * shc(sh, 0);
  if (deg > 0) {
    float x = dir.x, y = dir.y, z = dir.z;
    res = res - (SH_C1 * y) * shc(sh, 1) + (SH_C1 * z) * shc(sh, 2) - (SH_C1 * x) * shc(sh, 3);
    if (deg > 1) {
      float xx = x * x, yy = y * y, zz = z * z;
      float xy = x * y, yz = y * z, xz = x * z;
      res = res + (SH_C2_0 * xy) * shc(sh, 4) + (SH_C2_1 * yz) * shc(sh, 5) +
            (SH_C2_2 * (2.0f * zz - xx - yy)) * shc(sh, 6) + (SH_C2_3 * xz) * shc(sh, 7) +
            (SH_C2_4 * (xx - yy)) * shc(sh, 8);
      if (deg > 2) {
        res = res + ((SH_C3_0 * y) * (3.0f * xx - yy)) * shc(sh, 9) + ((SH_C3_1 * xy) * z) * shc(sh, 10) +
              ((SH_C3_2 * y) * (4.0f * zz - xx - yy)) * shc(sh, 11) +
              ((SH_C3_3 * z) * (2.0f * zz - 3.0f * xx - 3.0f * yy)) * shc(sh, 12) +
              ((SH_C3_4 * x) * (4.0f * zz - xx - yy)) * shc(sh, 13) + ((SH_C3_5 * z) * (xx - yy)) * shc(sh, 14) +
              ((SH_C3_6 * x) * (xx - 3.0f * yy)) * shc(sh, 15);
      }
    }
  }
  res = res + mk(0.5f, 0.5f, 0.5f);
  c0 = res.x < 0.0f;
  c1 = res.y < 0.0f;
  c2 = res.z < 0.0f;
  return mk(fmaxf(res.x, 0.0f), fmaxf(res.y, 0.0f), fmaxf(res.z, 0.0f));
}

// --------------------------------------------------------- sphere helpers
// hit_sphere (common.h:116-141)
TT_D bool hit_sphere(float rad, v3f cnt, v3f org, v3f dir, float &t) {
  v3f op = org - cnt;
  float a = dot(dir, dir);
  if (a <= 0.0f) return false;
  float b = dot(op, dir);
  float c = dot(op, op) - rad * rad;
  float det = b * b - c * a;
  if (det < 0.0f) return false;
  det = sqrtf(det);
  t = (-b - det) / a;
  if ((double)t < kStepEps) {
    t = (-b + det) / a;
    if ((double)t < kStepEps) return false;
  }
  return true;
}
TT_D v3f reflect3(v3f i, v3f n) {
  return i - (2.0f * n) * dot(n, i);
}

// ---- bounce variant (project-defined; no counterpart in the reference)
// Equirect environment [H][W][3]: direction -> phi = atan2(x, -z), theta = acos(y) in
// double after normalising in double, u = (phi / 2pi + 0.5) W, v = theta / pi H;
// bilinear over texel centres, wrapping in u, clamped in v.
struct EnvTaps {
  int idx[4];
  float w[4];
};
TT_D EnvTaps env_taps(v3f d, int W, int H) {
  constexpr double kInvPi = 0.31830988618379067154, kInv2Pi = 0.15915494309189533577;
  EnvTaps e;
  double dx = d.x, dy = d.y, dz = d.z;
  const double n = sqrt((dx * dx + dy * dy) + dz * dz);
  if (!(n > 0.0)) {
    for (int j = 0; j < 4; j++) e.idx[j] = 0, e.w[j] = j == 0 ? 1.0f : 0.0f;
    return e;
  }
  dx /= n, dy /= n, dz /= n;
  const double phi = atan2(dx, -dz);
  const double theta = acos(fmin(1.0, fmax(-1.0, dy)));
  const double u = (phi * kInv2Pi + 0.5) * (double)W - 0.5;
  const double v = theta * kInvPi * (double)H - 0.5;
  const double x0 = floor(u), y0 = floor(v);
  const double fx = u - x0, fy = v - y0;
  const int ix = (int)x0, iy = (int)y0;
  const int xa = ((ix % W) + W) % W, xb = (xa + 1) % W;
  const int ya = iy < 0 ? 0 : (iy > H - 1 ? H - 1 : iy), yb = iy + 1 < 0 ? 0 : (iy + 1 > H - 1 ? H - 1 : iy + 1);
  e.idx[0] = ya * W + xa, e.idx[1] = ya * W + xb, e.idx[2] = yb * W + xa, e.idx[3] = yb * W + xb;
  e.w[0] = (float)((1.0 - fx) * (1.0 - fy)), e.w[1] = (float)(fx * (1.0 - fy));
  e.w[2] = (float)((1.0 - fx) * fy), e.w[3] = (float)(fx * fy);
  return e;
}
TT_D void env_value(const float *env, int W, int H, v3f d, float out[3]) {
  const EnvTaps e = env_taps(d, W, H);
  for (int c = 0; c < 3; c++)
    out[c] = ((e.w[0] * env[3 * e.idx[0] + c] + e.w[1] * env[3 * e.idx[1] + c]) + e.w[2] * env[3 * e.idx[2] + c]) +
             e.w[3] * env[3 * e.idx[3] + c];
}

// Secondary rays from a primary segment's sums: expected normal N/|N| turned to face
// the ray, expected depth D/A, expected IOR eta/A (A = 1 - T); reflected ray from
// x + n eps; refracted ray (Snell, entering with ratio 1/eta; total internal
// reflection -> the reflected direction) from x - n eps.
struct Bounce {
  bool valid;
  v3f orr, dr, ot, dt;
};
TT_D Bounce bounce_rays(v3f o, v3f d, float T, const float N[3], float D, float eta_sum) {
  Bounce g;
  g.valid = false;
  const float A = 1.0f - T;
  const float nn = (N[0] * N[0] + N[1] * N[1]) + N[2] * N[2];
  if (!(A > 1e-6f) || !(nn > 1e-24f)) return g;
  v3f n = normalize(mk(N[0], N[1], N[2]));
  if (dot(n, d) > 0.0f) n = -n;
  const float eps = (float)kStepEps;
  const float tbar = D / A;
  const v3f x = o + d * tbar;
  const float eta = eta_sum / A;
  const v3f r = reflect3(d, n);
  const float cosi = -dot(d, n);
  const float ratio = 1.0f / eta;
  const float k = 1.0f - (ratio * ratio) * (1.0f - cosi * cosi);
  g.valid = true;
  g.orr = x + n * eps;
  g.dr = r;
  if (k < 0.0f) {
    g.ot = g.orr;
    g.dt = r;
  } else {
    g.ot = x - n * eps;
    g.dt = d * ratio + n * (ratio * cosi - sqrtf(k));
  }
  return g;
}
// refract.cu:166-187, with its formula as written.
TT_D bool refract_quirk(v3f &out, v3f in, v3f n, float eta_org) {
  float cosi = dot(in, n);
  float eta = eta_org;
  if (cosi > 0.0) {
    eta = (float)(1.0 / (double)eta_org);
    n = -n;
  }
  float s2i = (float)fmax(0.0, 1.0 - (double)sqrtf(cosi));
  float s2t = s2i / sqrtf(eta);
  if (s2t >= 1) return false;
  float cost = (float)sqrt(fmax(0.0, 1.0 - (double)s2t));
  out = divs(in, eta) + (-cosi / eta - cost) * n;
  return true;
}
// envlight.cu:25-51: LCG and cosine-weighted hemisphere sample.
TT_D float lcg(uint32_t &u) {
  u = u * 1664525u + 1013904223u;
  return (float)(u & 0x00FFFFFFu) * (1.0f / 16777216.0f);
}
TT_D v3f cos_hemi(uint32_t &seed, v3f nrm) {
  float u = lcg(seed);
  float v = lcg(seed);
  float z = sqrtf(u);
  float rr = sqrtf(fmaxf(0.0f, (float)(1.0 - (double)(z * z))));
  float phi = v * 6.2831853071795864769f;
  float sp, cp;
  sincos_tt(phi, sp, cp);
  v3f dl = mk(rr * cp, rr * sp, z);
  v3f up = fabsf(nrm.z) < 0.999f ? mk(0.0f, 0.0f, 1.0f) : mk(0.0f, 1.0f, 0.0f);
  v3f tt = normalize(cross(up, nrm));
  v3f bb = normalize(cross(nrm, tt));
  v3f wd = dl.x * tt + dl.y * bb + dl.z * nrm;
  return normalize(wd);
}

}  // namespace tt
