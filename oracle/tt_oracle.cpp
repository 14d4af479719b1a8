// CPU oracle for the triangle-splat tracer -- TEST INFRASTRUCTURE ONLY.
// See tt_oracle.h for what it restates and who may load it.
//
// Build: oracle/Makefile (g++ -O2 -fopenmp -ffp-contract=off). Contraction must
// stay off: the f32 path is written to round exactly where the HIP kernels
// round, so that hit lists and per-hit decisions compare bit-exactly.
#include "tt_oracle.h"

#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

namespace {

constexpr uint32_t kInvalid = 0xFFFFFFFFu;  // RayHit::invalid_id, params.h:119
constexpr float kInfDist = 1e6f;            // RayHit::infinite_dist, params.h:118
constexpr double kStepEps = 1e-9;           // STEP_EPSILON, config.h:11 (a double literal)
bool g_abs_vertex = false;  // see add_vertex_hit

// SH constants, auxiliary.h:33-50 (textbook real-SH normalisations; pinned by tests).
constexpr double SH_C0 = 0.28209479177387814;
constexpr double SH_C1 = 0.4886025119029199;
constexpr double SH_C2[5] = {1.0925484305920792, -1.0925484305920792, 0.31539156525252005, -1.0925484305920792,
                             0.5462742152960396};
constexpr double SH_C3[7] = {-0.5900435899266435, 2.890611442640554, -0.4570457994644658, 0.3731763325901154,
                             -0.4570457994644658, 1.445305721320277, -0.5900435899266435};

// ------------------------------------------------------------------ vectors
template <class R>
struct V3 {
  R x, y, z;
};
template <class R>
inline V3<R> mk(R x, R y, R z) {
  return V3<R>{x, y, z};
}
template <class R>
inline V3<R> operator+(V3<R> a, V3<R> b) {
  return {a.x + b.x, a.y + b.y, a.z + b.z};
}
template <class R>
inline V3<R> operator-(V3<R> a, V3<R> b) {
  return {a.x - b.x, a.y - b.y, a.z - b.z};
}
template <class R>
inline V3<R> operator-(V3<R> a) {
  return {-a.x, -a.y, -a.z};
}
template <class R>
inline V3<R> operator*(V3<R> a, R s) {
  return {a.x * s, a.y * s, a.z * s};
}
template <class R>
inline V3<R> operator*(R s, V3<R> a) {
  return {s * a.x, s * a.y, s * a.z};
}
template <class R>
inline V3<R> hmul(V3<R> a, V3<R> b) {
  return {a.x * b.x, a.y * b.y, a.z * b.z};
}
template <class R>
inline R dot(V3<R> a, V3<R> b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <class R>
inline V3<R> cross(V3<R> a, V3<R> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R>
inline R vlen(V3<R> v) {
  return std::sqrt(dot(v, v));
}
// normalize(v) = v * (1 / sqrt(v.v)); the vector-by-scalar divide is a multiply by the reciprocal.
template <class R>
inline V3<R> normalize(V3<R> v) {
  R inv = R(1) / std::sqrt(dot(v, v));
  return v * inv;
}
template <class R>
inline V3<R> divs(V3<R> a, R s) {
  R inv = R(1) / s;
  return a * inv;
}
template <class R>
inline R sumv(V3<R> a) {
  return a.x + a.y + a.z;
}
template <class R>
inline R comp(V3<R> v, int i) {
  return i == 0 ? v.x : (i == 1 ? v.y : v.z);
}
using V3f = V3<float>;
using V3d = V3<double>;

template <class R>
inline V3<R> ld3(const float *p) {
  return {R(p[0]), R(p[1]), R(p[2])};
}
template <class R>
inline V3<R> ld3d(const double *p) {
  return {R(p[0]), R(p[1]), R(p[2])};
}

// ------------------------------------------------ project-written FP32 math
inline uint32_t f2u(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}
inline float u2f(uint32_t u) {
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

// log2: x = m * 2^e with m in [sqrt(1/2), sqrt(2)); ln m = 2 atanh(s), s = (m-1)/(m+1),
// series to s^9. Arguments below the smallest normal are treated as zero.
float tt_log2f(float x) {
  if (!(x > 0.0f)) return (x == 0.0f) ? -INFINITY : NAN;
  if (x == INFINITY) return INFINITY;
  if (x < 1.17549435e-38f) return -INFINITY;
  uint32_t b = f2u(x);
  int e = (int)((b >> 23) & 0xffu) - 127;
  float m = u2f((b & 0x7fffffu) | 0x3f800000u);
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

// exp2: y = n + f, n = rint(y), |f| <= 1/2; e^(f ln2) by its Taylor series to degree 7;
// scaled by 2^n in two exact steps. Results below 2^-126 are flushed to zero.
float tt_exp2f(float y) {
  if (y != y) return y;
  if (y >= 128.0f) return INFINITY;
  if (y < -126.0f) return 0.0f;
  float n = std::rint(y);
  float f = y - n;
  float g = f * 0.693147181f;
  float p = 1.0f +
            g * (1.0f +
                 g * (0.5f + g * (0.166666667f +
                                  g * (0.0416666667f + g * (0.00833333333f + g * (0.00138888889f + g * 0.000198412698f))))));
  int ni = (int)n;
  int h = ni / 2;
  float s1 = u2f((uint32_t)(h + 127) << 23);
  float s2 = u2f((uint32_t)(ni - h + 127) << 23);
  return (p * s1) * s2;
}

// pow(x, y) = exp2(y * log2(x)), the same composition as a fast pow intrinsic:
// x < 0 gives NaN, x == 0 gives 0 for y > 0.
float tt_powf(float x, float y) {
  return tt_exp2f(y * tt_log2f(x));
}
float tt_logf(float x) {
  return tt_log2f(x) * 0.693147181f;
}

// sin/cos: quadrant reduction with a two-part pi/2, Taylor polynomials on [-pi/4, pi/4].
void tt_sincosf(float x, float *sn, float *cs) {
  float q = std::rint(x * 0.636619772f);
  float r = (x - q * 1.5703125f) - q * 4.83826794e-4f;
  float z = r * r;
  float s = r + r * z * (-0.166666667f + z * (0.00833333333f + z * (-0.000198412698f + z * 2.75573192e-6f)));
  float c = 1.0f + z * (-0.5f + z * (0.0416666667f + z * (-0.00138888889f + z * (2.48015873e-5f + z * -2.75573192e-7f))));
  int qi = ((int)q) & 3;
  if (qi == 0) {
    *sn = s;
    *cs = c;
  } else if (qi == 1) {
    *sn = c;
    *cs = -s;
  } else if (qi == 2) {
    *sn = -s;
    *cs = -c;
  } else {
    *sn = -c;
    *cs = s;
  }
}

template <class R>
inline R powR(R x, R y);
template <>
inline float powR<float>(float x, float y) {
  return tt_powf(x, y);
}
template <>
inline double powR<double>(double x, double y) {
  if (x > 0.0) return std::pow(x, y);
  if (x == 0.0 && y > 0.0) return 0.0;
  return NAN;
}
template <class R>
inline R logR(R x);
template <>
inline float logR<float>(float x) {
  return tt_logf(x);
}
template <>
inline double logR<double>(double x) {
  return std::log(x);
}

// --------------------------------------------------------- window function
// common.h:57-114 computeWindowFunc. Optionally the incenter distance is frozen
// (the finite-difference surrogate) -- the reference gradient treats it as constant.
template <class R>
struct Win {
  R Cx, phi_final, phi_max, dist;
  R ed[3];
  V3<R> n;
  V3<R> rel[3];  // hit - vertex i (f32 path: compensated)
};

// The reference's formulation, used by the f64 oracle: edge distances as dot(e, hit) + offset.
template <class R>
inline void window(const V3<R> V[3], V3<R> hit, R sigma, Win<R> &w, const double *frozen_dist) {
  V3<R> A = V[0], B = V[1], C = V[2];
  R a = vlen(B - C);
  R b = vlen(A - C);
  R c = vlen(A - B);
  V3<R> inc = divs(a * A + b * B + c * C, (a + b) + c);
  V3<R> ec = cross(B - A, C - A);
  w.n = normalize(ec);
  R dist = std::numeric_limits<R>::infinity();
  R phi = -std::numeric_limits<R>::infinity();
  for (int i = 0; i < 3; i++) {
    V3<R> p1 = V[i];
    V3<R> p2 = V[(i + 1) % 3];
    V3<R> en = normalize(cross(ec, p1 - p2));
    R off = -dot(en, p1);
    if (dist == std::numeric_limits<R>::infinity()) {
      dist = dot(en, inc) + off;
      if (dist > R(0)) {
        en = -en;
        off = -off;
        dist = -dist;
      }
    }
    R ed = dot(en, hit) + off;
    phi = std::fmax(phi, ed);
    w.ed[i] = ed;
    w.rel[i] = hit - p1;
  }
  if (frozen_dist) dist = R(*frozen_dist);
  w.phi_max = phi;
  w.dist = dist;
  w.phi_final = phi / dist;
  w.Cx = std::fmax(R(0), powR<R>(w.phi_final, sigma));
}

// The FP32 arithmetic of the HIP kernels. Same window function; the three
// edge-distance numerators n_i . (hit - p_i), with n_i = ec x (p_i - p_{i+1}),
// are formed in FP64 from the FP32 inputs (add/sub/mul only, so both sides round
// identically). In FP32 that dot product cancels to ~eps*|hit - p_i| absolute,
// which the window's diverging slope near an edge (sigma < 1) turns into
// percent-level gradient error; everything else is well conditioned in FP32.
inline void window32(const V3f V[3], V3f o, V3f d, float t, float sigma, Win<float> &w) {
  V3f A = V[0], B = V[1], C = V[2];
  V3f eAB = B - A, eAC = C - A, eBC = C - B;
  float a = vlen(eBC);
  float b = vlen(eAC);
  float c = vlen(eAB);
  V3f incA = divs(b * eAB + c * eAC, (a + b) + c);
  V3f ec = cross(eAB, eAC);
  w.n = normalize(ec);
  // FP64 part
  V3d Ad = mk<double>(A.x, A.y, A.z), Bd = mk<double>(B.x, B.y, B.z), Cd = mk<double>(C.x, C.y, C.z);
  V3d od = mk<double>(o.x, o.y, o.z), td = mk<double>(d.x, d.y, d.z) * (double)t;
  V3d rA = (od - Ad) + td, rB = (od - Bd) + td, rC = (od - Cd) + td;
  V3d ecd = cross(Bd - Ad, Cd - Ad);
  V3d n0 = cross(ecd, Ad - Bd), n1 = cross(ecd, Bd - Cd), n2 = cross(ecd, Cd - Ad);
  float q0 = (float)dot(n0, rA), q1 = (float)dot(n1, rB), q2 = (float)dot(n2, rC);
  V3f m0 = mk((float)n0.x, (float)n0.y, (float)n0.z);
  V3f m1 = mk((float)n1.x, (float)n1.y, (float)n1.z);
  V3f m2 = mk((float)n2.x, (float)n2.y, (float)n2.z);
  // FP32 part
  float il0 = 1.0f / std::sqrt(dot(m0, m0));
  float il1 = 1.0f / std::sqrt(dot(m1, m1));
  float il2 = 1.0f / std::sqrt(dot(m2, m2));
  float dist = dot(m0 * il0, incA);
  float ed0 = q0 * il0;
  if (dist > 0.0f) {
    dist = -dist;
    ed0 = -ed0;
  }
  w.ed[0] = ed0;
  w.ed[1] = q1 * il1;
  w.ed[2] = q2 * il2;
  w.rel[0] = mk((float)rA.x, (float)rA.y, (float)rA.z);
  w.rel[1] = mk((float)rB.x, (float)rB.y, (float)rB.z);
  w.rel[2] = mk((float)rC.x, (float)rC.y, (float)rC.z);
  float phi = std::fmax(std::fmax(w.ed[0], w.ed[1]), w.ed[2]);
  w.phi_max = phi;
  w.dist = dist;
  w.phi_final = phi / dist;
  w.Cx = std::fmax(0.0f, tt_powf(w.phi_final, sigma));
}

// ------------------------------------------------------------- SH colour
// common.h:16-55 computeColorFromSH (3DGS sign convention, +0.5, clamp at 0).
template <class R, class P>
inline V3<R> sh_color(const P *sh, int deg, V3<R> dir, bool clamped[3]) {
  dir = divs(dir, vlen(dir));
  auto S = [&](int i) { return V3<R>{R(sh[3 * i]), R(sh[3 * i + 1]), R(sh[3 * i + 2])}; };
  V3<R> res = R(SH_C0) * S(0);
  if (deg > 0) {
    R x = dir.x, y = dir.y, z = dir.z;
    res = res - (R(SH_C1) * y) * S(1) + (R(SH_C1) * z) * S(2) - (R(SH_C1) * x) * S(3);
    if (deg > 1) {
      R xx = x * x, yy = y * y, zz = z * z;
      R xy = x * y, yz = y * z, xz = x * z;
      res = res + (R(SH_C2[0]) * xy) * S(4) + (R(SH_C2[1]) * yz) * S(5) +
            (R(SH_C2[2]) * (R(2) * zz - xx - yy)) * S(6) + (R(SH_C2[3]) * xz) * S(7) +
            (R(SH_C2[4]) * (xx - yy)) * S(8);
      if (deg > 2) {
        res = res + ((R(SH_C3[0]) * y) * (R(3) * xx - yy)) * S(9) + ((R(SH_C3[1]) * xy) * z) * S(10) +
              ((R(SH_C3[2]) * y) * (R(4) * zz - xx - yy)) * S(11) +
              ((R(SH_C3[3]) * z) * (R(2) * zz - R(3) * xx - R(3) * yy)) * S(12) +
              ((R(SH_C3[4]) * x) * (R(4) * zz - xx - yy)) * S(13) + ((R(SH_C3[5]) * z) * (xx - yy)) * S(14) +
              ((R(SH_C3[6]) * x) * (xx - R(3) * yy)) * S(15);
      }
    }
  }
  res = res + mk(R(0.5), R(0.5), R(0.5));
  if (clamped) {
    clamped[0] = res.x < R(0);
    clamped[1] = res.y < R(0);
    clamped[2] = res.z < R(0);
  }
  return mk(std::fmax(res.x, R(0)), std::fmax(res.y, R(0)), std::fmax(res.z, R(0)));
}

// backward.cu:33-102 computeColorFromSHBwd: dL/dsh_k = Y_k(dir) * dL/drgb (masked by the clamp).
template <class R, class Acc>
inline void sh_bwd(int deg, V3<R> dir0, const bool clamped[3], const R dL_dcolor[3], Acc add) {
  V3<R> dir = divs(dir0, vlen(dir0));
  V3<R> g = mk(dL_dcolor[0], dL_dcolor[1], dL_dcolor[2]);
  g.x *= clamped[0] ? R(0) : R(1);
  g.y *= clamped[1] ? R(0) : R(1);
  g.z *= clamped[2] ? R(0) : R(1);
  R x = dir.x, y = dir.y, z = dir.z;
  add(0, R(SH_C0) * g);
  if (deg > 0) {
    add(1, (R(-SH_C1) * y) * g);
    add(2, (R(SH_C1) * z) * g);
    add(3, (R(-SH_C1) * x) * g);
    if (deg > 1) {
      R xx = x * x, yy = y * y, zz = z * z;
      R xy = x * y, yz = y * z, xz = x * z;
      add(4, (R(SH_C2[0]) * xy) * g);
      add(5, (R(SH_C2[1]) * yz) * g);
      add(6, (R(SH_C2[2]) * (R(2) * zz - xx - yy)) * g);
      add(7, (R(SH_C2[3]) * xz) * g);
      add(8, (R(SH_C2[4]) * (xx - yy)) * g);
      if (deg > 2) {
        add(9, ((R(SH_C3[0]) * y) * (R(3) * xx - yy)) * g);
        add(10, ((R(SH_C3[1]) * xy) * z) * g);
        add(11, ((R(SH_C3[2]) * y) * (R(4) * zz - xx - yy)) * g);
        add(12, ((R(SH_C3[3]) * z) * (R(2) * zz - R(3) * xx - R(3) * yy)) * g);
        add(13, ((R(SH_C3[4]) * x) * (R(4) * zz - xx - yy)) * g);
        add(14, ((R(SH_C3[5]) * z) * (xx - yy)) * g);
        add(15, ((R(SH_C3[6]) * x) * (xx - R(3) * yy)) * g);
      }
    }
  }
}

// ---------------------------------------------- 3x3 helpers (row-major [r*3+c])
// Semantics of vecmath.h:2626-2859 (vecmat3x3, skew3x3, dnormvdv, dnorm_dp1/2/3).
template <class R>
inline V3<R> vecmat(V3<R> v, const R A[9]) {
  return {A[0] * v.x + A[3] * v.y + A[6] * v.z, A[1] * v.x + A[4] * v.y + A[7] * v.z,
          A[2] * v.x + A[5] * v.y + A[8] * v.z};
}
template <class R>
inline void skew(V3<R> v, R S[9]) {
  S[0] = R(0), S[1] = -v.z, S[2] = v.y;
  S[3] = v.z, S[4] = R(0), S[5] = -v.x;
  S[6] = -v.y, S[7] = v.x, S[8] = R(0);
}
template <class R>
inline void outer(V3<R> u, V3<R> v, R M[9]) {
  M[0] = u.x * v.x, M[1] = u.x * v.y, M[2] = u.x * v.z;
  M[3] = u.y * v.x, M[4] = u.y * v.y, M[5] = u.y * v.z;
  M[6] = u.z * v.x, M[7] = u.z * v.y, M[8] = u.z * v.z;
}
template <class R>
inline void eye_scaled(R s, R M[9]) {
  for (int i = 0; i < 9; i++) M[i] = ((i % 4) == 0 ? R(1) : R(0)) * s;
}
// d(v/|v|)/dv = (I - n n^T) / |v|; the scale is formed as 1/|v| in double then rounded.
template <class R>
inline void dnormvdv(V3<R> nv, V3<R> v, R out[9]) {
  R op[9];
  outer(nv, nv, op);
  for (int i = 0; i < 9; i++) out[i] = ((i % 4) == 0 ? R(1) : R(0)) - op[i];
  R s = R(1.0 / (double)vlen(v));
  for (int i = 0; i < 9; i++) out[i] *= s;
}
// ne = (p1-p3)x(p2-p3) x (p1-p2) = b (a.c) - a (b.c); Jacobians w.r.t. p1, p2, p3.
template <class R>
inline void dne_dp1(V3<R> p1, V3<R> p2, V3<R> p3, R out[9]) {
  V3<R> b = p2 - p3, c = p1 - p2;
  V3<R> axb = cross(p1 - p3, b);
  R m[9], d[9], s[9];
  outer(b, c, m);
  eye_scaled(dot(b, c), d);
  skew(axb, s);
  for (int i = 0; i < 9; i++) m[i] = m[i] - d[i];
  for (int i = 0; i < 9; i++) out[i] = m[i] + s[i];
}
template <class R>
inline void dne_dp2(V3<R> p1, V3<R> p2, V3<R> p3, R out[9]) {
  V3<R> a = p1 - p3, b = p2 - p3, c = p1 - p2;
  V3<R> axb = cross(p1 - p3, b);
  R d[9], m[9], s[9];
  eye_scaled(dot(a, c), d);
  outer(a, c, m);
  for (int i = 0; i < 9; i++) d[i] = d[i] - m[i];
  skew(axb, s);
  for (int i = 0; i < 9; i++) out[i] = d[i] - s[i];
}
template <class R>
inline void dne_dp3(V3<R> p1, V3<R> p2, V3<R> p3, R out[9]) {
  V3<R> c = p1 - p2;
  R d[9], m[9];
  eye_scaled(dot(c, c), d);
  outer(c, c, m);
  for (int i = 0; i < 9; i++) out[i] = m[i] - d[i];
}

// backward.cu:104-189 computeWindowFuncBwd (hit point and incenter distance held constant).
template <class R, class Acc>
inline void window_bwd(const V3<R> vert[3], const V3<R> rel[3], R dL_dphi, R phi_x, const R ed[3], V3<R> dL_dn3,
                       Acc add, int edge_mask = -1) {
  for (int k = 0; k < 3; k++) {
    // the max edge(s), backward.cu's tolerance test; a replay passes the FP32 pass's choice
    if (edge_mask >= 0 ? !((edge_mask >> k) & 1) : std::fabs(ed[k] - phi_x) > R(1e-6f)) continue;
    V3<R> p1 = vert[k], p2 = vert[(k + 1) % 3], p3 = vert[(k + 2) % 3];
    V3<R> n = cross(p1 - p3, p2 - p3);
    V3<R> ne = cross(n, p1 - p2);
    V3<R> Ne = normalize(ne);
    R J[9];
    dnormvdv(Ne, ne, J);
    V3<R> dL_dNe = dL_dphi * rel[k];
    V3<R> dL_dne = vecmat(dL_dNe, J);
    R j1[9], j2[9], j3[9];
    dne_dp1(p1, p2, p3, j1);
    dne_dp2(p1, p2, p3, j2);
    dne_dp3(p1, p2, p3, j3);
    V3<R> g1 = vecmat(dL_dne, j1) - dL_dphi * Ne;
    V3<R> g2 = vecmat(dL_dne, j2);
    V3<R> g3 = vecmat(dL_dne, j3);
    add(k, g1);
    add((k + 1) % 3, g2);
    add((k + 2) % 3, g3);
  }
  {
    V3<R> p1 = vert[0], p2 = vert[1], p3 = vert[2];
    V3<R> n = cross(p1 - p3, p2 - p3);
    V3<R> N = normalize(n);
    R J[9];
    dnormvdv(N, n, J);
    V3<R> dL_dn = vecmat(dL_dn3, J);
    R s1[9], s2[9], s3[9];
    skew(p3 - p2, s1);
    skew(p1 - p3, s2);
    skew(p2 - p1, s3);
    add(0, vecmat(dL_dn, s1));
    add(1, vecmat(dL_dn, s2));
    add(2, vecmat(dL_dn, s3));
  }
}

// ------------------------------------------------- watertight intersection
struct RayPre {
  V3f o, d;
  int kx, ky, kz;
  float Sx, Sy, Sz;
  V3f invd;
};

inline RayPre prep_ray(V3f o, V3f d) {
  RayPre r;
  r.o = o;
  r.d = d;
  float ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z);
  r.kz = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
  r.kx = r.kz + 1;
  if (r.kx == 3) r.kx = 0;
  r.ky = r.kx + 1;
  if (r.ky == 3) r.ky = 0;
  if (comp(d, r.kz) < 0.0f) std::swap(r.kx, r.ky);
  r.Sx = comp(d, r.kx) / comp(d, r.kz);
  r.Sy = comp(d, r.ky) / comp(d, r.kz);
  r.Sz = 1.0f / comp(d, r.kz);
  auto sd = [](float v) { return std::fabs(v) > 1e-30f ? v : std::copysign(1e-30f, v); };
  r.invd = mk(1.0f / sd(d.x), 1.0f / sd(d.y), 1.0f / sd(d.z));
  return r;
}

// Woop, Benthin, Wald: "Watertight Ray/Triangle Intersection", JCGT 2(1), 2013.
// Double-sided; exact-zero edge functions are recomputed in double.
inline bool intersect(const RayPre &r, V3f v0, V3f v1, V3f v2, float &t) {
  V3f A = v0 - r.o, B = v1 - r.o, C = v2 - r.o;
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
  t = T / det + 0.0f;  // -0.0 as +0.0, as the kernels' hit keys need
  return true;
}

// ------------------------------------------------------------------- scene
struct BNode {
  float lo[3], hi[3];
  int left, right;   // children (internal)
  int start, count;  // leaf range into order[] when count > 0
};

struct Scene {
  const float *verts;
  int V;
  const int *npts;
  const int *cumsum;
  int P;
  const float *shs;
  int M;
  int deg;
  const float *opac;
  const float *sigma;
  std::vector<BNode> nodes;
  std::vector<int> order;
  double bvh_seconds = 0.0;

  V3f vtx(int prim, int j) const { return ld3<float>(verts + 3 * (size_t)(cumsum[prim] + j)); }
};

void build_bvh(Scene &s) {
  auto t0 = std::chrono::steady_clock::now();
  s.nodes.clear();
  s.order.resize(s.P);
  std::vector<float> cen((size_t)s.P * 3);
  std::vector<float> blo((size_t)s.P * 3), bhi((size_t)s.P * 3);
  for (int i = 0; i < s.P; i++) {
    s.order[i] = i;
    for (int a = 0; a < 3; a++) {
      float v0 = s.verts[3 * (size_t)s.cumsum[i] + a];
      float v1 = s.verts[3 * (size_t)(s.cumsum[i] + 1) + a];
      float v2 = s.verts[3 * (size_t)(s.cumsum[i] + 2) + a];
      float lo = std::min(v0, std::min(v1, v2)), hi = std::max(v0, std::max(v1, v2));
      blo[3 * (size_t)i + a] = lo;
      bhi[3 * (size_t)i + a] = hi;
      cen[3 * (size_t)i + a] = 0.5f * (lo + hi);
    }
  }
  if (s.P == 0) return;
  struct Item {
    int node, b, e;
  };
  std::vector<Item> work;
  s.nodes.push_back(BNode{});
  work.push_back({0, 0, s.P});
  while (!work.empty()) {
    Item it = work.back();
    work.pop_back();
    BNode nd;
    float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int a = 0; a < 3; a++) nd.lo[a] = INFINITY, nd.hi[a] = -INFINITY;
    for (int i = it.b; i < it.e; i++) {
      int p = s.order[i];
      for (int a = 0; a < 3; a++) {
        nd.lo[a] = std::min(nd.lo[a], blo[3 * (size_t)p + a]);
        nd.hi[a] = std::max(nd.hi[a], bhi[3 * (size_t)p + a]);
        clo[a] = std::min(clo[a], cen[3 * (size_t)p + a]);
        chi[a] = std::max(chi[a], cen[3 * (size_t)p + a]);
      }
    }
    // conservative padding: the culling must never drop a triangle whose computed t is in range
    for (int a = 0; a < 3; a++) {
      float pad = 1e-5f * (1.0f + std::max(std::fabs(nd.lo[a]), std::fabs(nd.hi[a])));
      nd.lo[a] -= pad;
      nd.hi[a] += pad;
    }
    int n = it.e - it.b;
    if (n <= 4) {
      nd.left = nd.right = -1;
      nd.start = it.b;
      nd.count = n;
      s.nodes[it.node] = nd;
      continue;
    }
    int axis = 0;
    float ext = chi[0] - clo[0];
    for (int a = 1; a < 3; a++)
      if (chi[a] - clo[a] > ext) ext = chi[a] - clo[a], axis = a;
    int mid = it.b + n / 2;
    std::nth_element(s.order.begin() + it.b, s.order.begin() + mid, s.order.begin() + it.e, [&](int x, int y) {
      float cx = cen[3 * (size_t)x + axis], cy = cen[3 * (size_t)y + axis];
      return cx < cy || (cx == cy && x < y);
    });
    nd.start = nd.count = 0;
    nd.left = (int)s.nodes.size();
    nd.right = nd.left + 1;
    s.nodes[it.node] = nd;
    s.nodes.push_back(BNode{});
    s.nodes.push_back(BNode{});
    work.push_back({nd.left, it.b, mid});
    work.push_back({nd.right, mid, it.e});
  }
  s.bvh_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// ---------------------------------------------------------------- k-gather
struct Hit {
  float t;
  uint32_t prim;
};
inline bool hit_less(float t, uint32_t p, const Hit &h) {
  return t < h.t || (t == h.t && p < h.prim);
}
inline bool after_key(float t, uint32_t prim, float tlo, long long plo) {
  return t > tlo || (t == tlo && (long long)prim > plo);
}
inline void kinsert(Hit *buf, int k, float t, uint32_t p) {
  if (!hit_less(t, p, buf[k - 1])) return;
  for (int i = 0; i < k; i++) {
    if (hit_less(t, p, buf[i])) {
      Hit tmp = buf[i];
      buf[i] = Hit{t, p};
      t = tmp.t;
      p = tmp.prim;
    }
  }
}

inline bool box_visit(const BNode &n, const RayPre &r, float tlo, float thi) {
  float tn = -INFINITY, tf = INFINITY;
  const float o[3] = {r.o.x, r.o.y, r.o.z}, inv[3] = {r.invd.x, r.invd.y, r.invd.z};
  for (int a = 0; a < 3; a++) {
    float t0 = (n.lo[a] - o[a]) * inv[a];
    float t1 = (n.hi[a] - o[a]) * inv[a];
    tn = std::max(tn, std::min(t0, t1));
    tf = std::min(tf, std::max(t0, t1));
  }
  tf = tf + std::fabs(tf) * 1e-6f + 1e-7f;
  tn = tn - std::fabs(tn) * 1e-6f - 1e-7f;
  if (tn > tf) return false;
  if (tf < tlo - std::fabs(tlo) * 1e-6f - 1e-7f) return false;
  if (tn > thi + std::fabs(thi) * 1e-6f + 1e-7f) return false;
  return true;
}

// One chunk: the k smallest (t, prim) among hits with t <= far and (t, prim) after the key.
void gather(const Scene &s, const RayPre &r, float tlo, long long plo, int k, float far_t, bool use_bvh, Hit *buf) {
  for (int i = 0; i < k; i++) buf[i] = Hit{kInfDist, kInvalid};
  auto test = [&](int p) {
    float t;
    if (!intersect(r, s.vtx(p, 0), s.vtx(p, 1), s.vtx(p, 2), t)) return;
    if (!(t <= far_t)) return;
    if (!after_key(t, (uint32_t)p, tlo, plo)) return;
    kinsert(buf, k, t, (uint32_t)p);
  };
  if (!use_bvh || s.nodes.empty()) {
    for (int p = 0; p < s.P; p++) test(p);
    return;
  }
  int stack[128];
  int sp = 0;
  stack[sp++] = 0;
  while (sp > 0) {
    const BNode &n = s.nodes[stack[--sp]];
    float thi = std::min(far_t, buf[k - 1].t);
    if (!box_visit(n, r, tlo, thi)) continue;
    if (n.count > 0) {
      for (int i = 0; i < n.count; i++) test(s.order[n.start + i]);
    } else {
      stack[sp++] = n.left;
      stack[sp++] = n.right;
    }
  }
}

// ------------------------------------------------------------- f32 tracing
struct Opts {
  int k, mode, variant;
  bool bvh;
  float cutoff, near_t, far_t;
  float bg[3];
  std::vector<float> extra;
  const float *tstart = nullptr;
  const float *refl = nullptr, *ior = nullptr, *env = nullptr;
  int env_w = 0, env_h = 0;
};
Opts to_opts(const or_opts *o) {
  Opts p;
  p.k = o->k;
  p.mode = o->chunk_mode;
  p.variant = o->variant;
  p.bvh = o->use_bvh != 0;
  p.cutoff = o->cutoff;
  p.near_t = o->near_t;
  p.far_t = o->far_t;
  for (int i = 0; i < 3; i++) p.bg[i] = o->bg[i];
  if (o->extra && o->extra_len > 0) p.extra.assign(o->extra, o->extra + o->extra_len);
  p.tstart = o->variant == 0 ? o->tstart : nullptr;
  p.refl = o->refl;
  p.ior = o->ior;
  p.env = o->env;
  p.env_w = o->env_w;
  p.env_h = o->env_h;
  return p;
}

struct LogHit {
  uint32_t prim;
  float t;
  int code;        // 0 near skip, 1 alpha skip, 2 blended, 3 terminated (not blended)
  double dist;     // incenter distance in double (blended hits only)
  int edges = -1;  // blended hits: the edges the FP32 pass takes as the window's max edge
                   // (bit k: |ed_k - phi| <= 1e-6 in FP32), a discrete decision the FP64 replay keeps
};
struct Segment {
  int pix;
  V3f o, d;
  double weight;  // 1, or 1/samples for envlight hemisphere samples
  std::vector<LogHit> hits;
  int role = 0;   // bounce variant: 0 primary, 1 reflected, 2 refracted
};

// Bounce variant: sums over a segment's blended hits of w*rho, w*(1-rho), w*eta, w*t
// (w = alpha * T, the blending weight).
struct BounceAcc {
  float R = 0.0f, Q = 0.0f, eta = 0.0f, D = 0.0f;
};

inline void atomic_max_f(float *p, float v) {
  int *ip = reinterpret_cast<int *>(p);
  int nv;
  std::memcpy(&nv, &v, 4);
  int cur = __atomic_load_n(ip, __ATOMIC_RELAXED);
  while (nv > cur && !__atomic_compare_exchange_n(ip, &cur, nv, true, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
  }
}
inline void atomic_add_f(float *p, float v) {
#pragma omp atomic
  *p += v;
}

// forward.cu:25-32 angleBetween, :34-74 computeScaling
inline float angle_between(V3f a, V3f b) {
  float len_ab = vlen(a) * vlen(b);
  if (len_ab < 1e-12f) return 0.0f;
  float c = std::fmax(-1.0f, std::fmin(1.0f, dot(a, b) / len_ab));
  return std::acos(c);
}
inline void compute_scaling(const Scene &s, int idx, V3f org, float *scaling) {
  if (scaling[idx] > 0.0f) return;
  int off = s.cumsum[idx];
  int np = s.npts[idx];
  V3f cen = mk(0.f, 0.f, 0.f);
  for (int i = 0; i < np; i++) cen = cen + ld3<float>(s.verts + 3 * (size_t)(off + i));
  cen = cen * (1.0f / (float)np);
  V3f tri2org = cen - org;
  float mx = 0.0f;
  for (int i = 0; i < np; i++) {
    float ang = angle_between(ld3<float>(s.verts + 3 * (size_t)(off + i)) - org, tri2org);
    mx = std::fmax(ang, mx);
  }
  scaling[idx] = mx;
}

struct FwdSinks {
  float *maxb = nullptr, *density = nullptr, *scaling = nullptr;
  V3f scaling_org;
};

// Start key of the bounce variant's secondary rays (this project's model, no reference
// counterpart): in exact chunk mode they begin at (near_t, -1) -- the hits before near_t
// would only be consumed and skipped, so the blended result is the same and the gather
// does not collect them; in compat mode they begin at 0 (chunk boundaries there depend on
// every consumed hit).
inline float bounce_secondary_start(const Opts &op) { return op.mode == 1 ? op.near_t : 0.0f; }

// The chunk loop of forward.cu:76-155 (FULL) and of the variants (colour only).
// Returns with C, T, N updated; counts consumed/blended hits; optionally logs hits.
void trace_segment_f32(const Scene &s, const Opts &op, V3f o, V3f d, float C[3], float &T, float N[3],
                       int &consumed, int &blended, const FwdSinks *sinks, std::vector<LogHit> *log,
                       float tstart = 0.0f, BounceAcc *acc = nullptr) {
  RayPre r = prep_ray(o, d);
  std::vector<Hit> buf(op.k);
  float last = 0.0f;
  float tlo = tstart;  // per-ray start distance (exact mode): the trace begins at key (tstart, -1)
  long long plo = -1;
  while (true) {
    gather(s, r, tlo, plo, op.k, op.far_t, op.bvh, buf.data());
    for (int i = 0; i < op.k; i++) {
      Hit h = buf[i];
      if (h.prim == kInvalid) return;
      consumed++;
      int prim = (int)h.prim;
      float t = h.t;
      last = std::fmax((float)((double)last + kStepEps), t);
      if (t < op.near_t) {
        if (log) log->push_back({h.prim, t, 0, 0.0});
        continue;
      }
      V3f V[3] = {s.vtx(prim, 0), s.vtx(prim, 1), s.vtx(prim, 2)};
      Win<float> w;
      window32(V, o, d, t, s.sigma[prim], w);
      float alpha = std::fmin(0.99f, s.opac[prim] * w.Cx);
      if (alpha < 1.0f / 255.0f) {
        if (log) log->push_back({h.prim, t, 1, 0.0});
        continue;
      }
      float testT = T * (1.0f - alpha);
      if (testT < op.cutoff) {
        if (log) log->push_back({h.prim, t, 3, 0.0});
        return;
      }
      float bw = alpha * T;
      if (sinks) {
        sinks->density[prim] = 1.0f;
        compute_scaling(s, prim, sinks->scaling_org, sinks->scaling);
        atomic_max_f(sinks->maxb + prim, bw);
      }
      if (sinks || acc) {
        N[0] += w.n.x * bw;
        N[1] += w.n.y * bw;
        N[2] += w.n.z * bw;
      }
      V3f col = sh_color<float>(s.shs + (size_t)prim * s.M * 3, s.deg, d, nullptr);
      C[0] += col.x * bw;
      C[1] += col.y * bw;
      C[2] += col.z * bw;
      if (acc) {
        const float rho = op.refl[prim];
        acc->R += bw * rho;
        acc->Q += bw * (1.0f - rho);
        acc->eta += bw * op.ior[prim];
        acc->D += bw * t;
      }
      T = testT;
      blended++;
      if (log) {
        V3d Vd[3] = {ld3<double>(s.verts + 3 * (size_t)s.cumsum[prim]),
                     ld3<double>(s.verts + 3 * (size_t)(s.cumsum[prim] + 1)),
                     ld3<double>(s.verts + 3 * (size_t)(s.cumsum[prim] + 2))};
        Win<double> wd;
        V3d hd = mk<double>(o.x, o.y, o.z) + (double)t * mk<double>(d.x, d.y, d.z);
        window<double>(Vd, hd, (double)s.sigma[prim], wd, nullptr);
        int edges = 0;
        for (int k = 0; k < 3; k++)
          if (!(std::fabs(w.ed[k] - w.phi_max) > 1e-6f)) edges |= 1 << k;
        log->push_back({h.prim, t, 2, wd.dist, edges});
      }
    }
    if (op.mode == 1) {
      tlo = buf[op.k - 1].t;
      plo = (long long)buf[op.k - 1].prim;
    } else if (last > tlo) {
      tlo = last;
      plo = -1;
    } else {
      plo = (long long)buf[op.k - 1].prim;  // no progress in t: resume after the last prim
    }
  }
}

// common.h:116-141 hit_sphere
inline bool hit_sphere(float rad, V3f cnt, V3f org, V3f dir, float &t) {
  V3f op = org - cnt;
  float a = dot(dir, dir);
  if (a <= 0.0f) return false;
  float b = dot(op, dir);
  float c = dot(op, op) - rad * rad;
  float det = b * b - c * a;
  if (det < 0.0f) return false;
  det = std::sqrt(det);
  t = (-b - det) / a;
  if ((double)t < kStepEps) {
    t = (-b + det) / a;
    if ((double)t < kStepEps) return false;
  }
  return true;
}
inline V3f reflect3(V3f i, V3f n) {
  return i - (2.0f * n) * dot(n, i);
}
// refract.cu:166-187, including its formula as written (sin^2 from sqrt(cos), /sqrt(eta)).
inline bool refract_quirk(V3f &out, V3f in, V3f n, float eta_org) {
  float cosi = dot(in, n);
  float eta = eta_org;
  if (cosi > 0.0) {
    eta = (float)(1.0 / (double)eta_org);
    n = -n;
  }
  float s2i = (float)std::fmax(0.0, 1.0 - (double)std::sqrt(cosi));
  float s2t = s2i / std::sqrt(eta);
  if (s2t >= 1) return false;
  float cost = (float)std::sqrt(std::fmax(0.0, 1.0 - (double)s2t));
  out = divs(in, eta) + (-cosi / eta - cost) * n;
  return true;
}
const float kEps32 = (float)kStepEps;

// envlight.cu:25-51
inline float lcg(uint32_t &u) {
  u = u * 1664525u + 1013904223u;
  return (float)(u & 0x00FFFFFFu) * (1.0f / 16777216.0f);
}
inline V3f cos_hemi(uint32_t &seed, V3f nrm) {
  float u = lcg(seed);
  float v = lcg(seed);
  float z = std::sqrt(u);
  float rr = std::sqrt(std::fmax(0.0f, (float)(1.0 - (double)(z * z))));
  float phi = v * 6.2831853071795864769f;
  float sp, cp;
  tt_sincosf(phi, &sp, &cp);
  V3f dl = mk(rr * cp, rr * sp, z);
  V3f up = std::fabs(nrm.z) < 0.999f ? mk(0.0f, 0.0f, 1.0f) : mk(0.0f, 1.0f, 0.0f);
  V3f tt = normalize(cross(up, nrm));
  V3f bb = normalize(cross(nrm, tt));
  V3f wd = dl.x * tt + dl.y * bb + dl.z * nrm;
  return normalize(wd);
}

// ---- bounce variant (no counterpart in the reference; a project-defined model)
// Equirect environment [H][W][3]: direction -> phi = atan2(x, -z), theta = acos(y) (in
// double, after normalising in double), u = (phi / 2pi + 0.5) W, v = theta / pi H,
// bilinear over texel centres, wrapping in u and clamped in v.
constexpr double kInvPi = 0.31830988618379067154, kInv2Pi = 0.15915494309189533577;
struct EnvTaps {
  int idx[4];
  double w[4];
};
inline EnvTaps env_taps(float fx_, float fy_, float fz_, int W, int H) {
  EnvTaps e;
  double dx = fx_, dy = fy_, dz = fz_;
  const double n = std::sqrt((dx * dx + dy * dy) + dz * dz);
  if (!(n > 0.0)) {
    for (int j = 0; j < 4; j++) e.idx[j] = 0, e.w[j] = j == 0 ? 1.0 : 0.0;
    return e;
  }
  dx /= n, dy /= n, dz /= n;
  const double phi = std::atan2(dx, -dz);
  const double theta = std::acos(std::fmin(1.0, std::fmax(-1.0, dy)));
  const double u = (phi * kInv2Pi + 0.5) * (double)W - 0.5;
  const double v = theta * kInvPi * (double)H - 0.5;
  const double x0 = std::floor(u), y0 = std::floor(v);
  const double fx = u - x0, fy = v - y0;
  const int ix = (int)x0, iy = (int)y0;
  const int xa = ((ix % W) + W) % W, xb = (xa + 1) % W;
  const int ya = iy < 0 ? 0 : (iy > H - 1 ? H - 1 : iy), yb = iy + 1 < 0 ? 0 : (iy + 1 > H - 1 ? H - 1 : iy + 1);
  e.idx[0] = ya * W + xa, e.idx[1] = ya * W + xb, e.idx[2] = yb * W + xa, e.idx[3] = yb * W + xb;
  e.w[0] = (1.0 - fx) * (1.0 - fy), e.w[1] = fx * (1.0 - fy), e.w[2] = (1.0 - fx) * fy, e.w[3] = fx * fy;
  return e;
}
inline void env_f32(const Opts &op, V3f d, float out[3]) {
  const EnvTaps e = env_taps(d.x, d.y, d.z, op.env_w, op.env_h);
  const float w0 = (float)e.w[0], w1 = (float)e.w[1], w2 = (float)e.w[2], w3 = (float)e.w[3];
  for (int c = 0; c < 3; c++)
    out[c] = ((w0 * op.env[3 * e.idx[0] + c] + w1 * op.env[3 * e.idx[1] + c]) + w2 * op.env[3 * e.idx[2] + c]) +
             w3 * op.env[3 * e.idx[3] + c];
}
inline void env_f64(const double *env, int W, int H, V3f d, double out[3]) {
  const EnvTaps e = env_taps(d.x, d.y, d.z, W, H);
  for (int c = 0; c < 3; c++) {
    out[c] = 0.0;
    for (int j = 0; j < 4; j++) out[c] += e.w[j] * env[3 * e.idx[j] + c];
  }
}

// The two secondary rays of a pixel from its primary segment: expected normal
// N / |N| turned to face the ray, expected depth D / A (A = 1 - T), expected IOR
// eta / A; reflected ray from x + n eps, refracted ray (Snell, entering with ratio
// 1/eta; total internal reflection -> the reflected direction) from x - n eps.
struct BounceGeom {
  bool valid = false;
  V3f orr, dr, ot, dt;
};
inline BounceGeom bounce_geometry(V3f o, V3f d, float T, const float N[3], const BounceAcc &acc) {
  BounceGeom g;
  const float A = 1.0f - T;
  const float nn = (N[0] * N[0] + N[1] * N[1]) + N[2] * N[2];
  if (!(A > 1e-6f) || !(nn > 1e-24f)) return g;
  V3f n = normalize(mk(N[0], N[1], N[2]));
  if (dot(n, d) > 0.0f) n = -n;
  const float tbar = acc.D / A;
  const V3f x = o + d * tbar;
  const float eta = acc.eta / A;
  const V3f r = reflect3(d, n);
  const float cosi = -dot(d, n);
  const float ratio = 1.0f / eta;
  const float k = 1.0f - (ratio * ratio) * (1.0f - cosi * cosi);
  g.valid = true;
  g.orr = x + n * kEps32;
  g.dr = r;
  if (k < 0.0f) {
    g.ot = g.orr;
    g.dt = r;
  } else {
    g.ot = x - n * kEps32;
    g.dt = d * ratio + n * (ratio * cosi - std::sqrt(k));
  }
  return g;
}

// Per pixel: the segments (traced rays) a variant launches, mirroring the raygens.
template <class F>
void pixel_segments(const Opts &op, int pix, V3f o, V3f d, F &&seg_fn, float C[3], float &T) {
  if (op.variant == 0) {
    seg_fn(o, d, 1.0, false);
    return;
  }
  float rad = op.extra.size() > 0 ? op.extra[0] : 1.0f;
  V3f cnt = mk(op.extra.size() > 3 ? op.extra[1] : 0.f, op.extra.size() > 3 ? op.extra[2] : 0.f,
               op.extra.size() > 3 ? op.extra[3] : 0.f);
  if (op.variant == 1) {  // reflect.cu:25-48
    float st;
    if (hit_sphere(rad, cnt, o, d, st)) {
      o = o + d * st;
      V3f nrm = normalize(o - cnt);
      o = o + nrm * kEps32;
      d = reflect3(normalize(d), nrm);
    }
    seg_fn(o, d, 1.0, false);
  } else if (op.variant == 2) {  // refract.cu:189-213
    float st;
    if (hit_sphere(rad, cnt, o, d, st)) {
      V3f hp = o + d * st;
      V3f nrm = normalize(hp - cnt);
      V3f in = normalize(d);
      float ior = 1.5f / 1.0003f;
      V3f nd = in;
      if (!refract_quirk(nd, in, nrm, ior)) {
        V3f rn = dot(in, nrm) < 0.0f ? nrm : -nrm;
        nd = reflect3(in, rn);
      }
      o = hp + nrm * kEps32;
      d = nd;
    }
    seg_fn(o, d, 1.0, false);
  } else {  // envlight.cu:116-184
    int samples = (int)(op.extra.size() > 4 ? op.extra[4] : 0.0f);
    float st;
    if (hit_sphere(rad, cnt, o, d, st)) {
      o = o + d * st;
      V3f nrm0 = normalize(o - cnt);
      o = o + nrm0 * kEps32;
      float radc[3] = {0, 0, 0};
      for (int i = 0; i < samples; i++) {
        V3f nrm = normalize(o - cnt);
        uint32_t seed = (((uint32_t)pix + 1u) * 9781u) ^ (((uint32_t)i + 1u) * 6271u);
        V3f dn = cos_hemi(seed, nrm);
        V3f on = o + dn * kEps32;
        float Cs[3] = {0, 0, 0}, Ts = 1.0f;
        seg_fn(on, dn, samples > 0 ? 1.0 / samples : 0.0, true, Cs, &Ts);
        for (int c = 0; c < 3; c++) radc[c] += Cs[c] + Ts * op.bg[c];
      }
      float inv = 1.0f / (float)samples;
      for (int c = 0; c < 3; c++) C[c] += radc[c] * inv;
      T = 0.0f;
    } else {
      seg_fn(o, d, 1.0, false);
    }
  }
}

}  // namespace

// ============================================================== C entry points
extern "C" {

void or_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
}
void or_set_abs_vertex(int on) {
  g_abs_vertex = on != 0;
}
int or_get_threads(void) {
  return omp_get_max_threads();
}

float or_log2f(float x) {
  return tt_log2f(x);
}
float or_exp2f(float x) {
  return tt_exp2f(x);
}
float or_powf(float x, float y) {
  return tt_powf(x, y);
}
void or_sincosf(float x, float *s, float *c) {
  tt_sincosf(x, s, c);
}

int or_intersect(const float *o, const float *d, const float *v0, const float *v1, const float *v2, float *t) {
  RayPre r = prep_ray(ld3<float>(o), ld3<float>(d));
  return intersect(r, ld3<float>(v0), ld3<float>(v1), ld3<float>(v2), *t) ? 1 : 0;
}

void or_cos_hemi(unsigned int seed, const float *nrm, int count, float *out) {
  uint32_t u = seed;
  const V3f n = ld3<float>(nrm);
  for (int i = 0; i < count; i++) {
    const V3f w = cos_hemi(u, n);
    out[3 * i] = w.x, out[3 * i + 1] = w.y, out[3 * i + 2] = w.z;
  }
}

void or_window_f32(const float *verts, const float *hit, float sigma, float *out) {
  V3f V[3] = {ld3<float>(verts), ld3<float>(verts + 3), ld3<float>(verts + 6)};
  Win<float> w;
  window32(V, ld3<float>(hit), mk(0.0f, 0.0f, 0.0f), 0.0f, sigma, w);
  out[0] = w.Cx, out[1] = w.phi_final, out[2] = w.phi_max;
  out[3] = w.ed[0], out[4] = w.ed[1], out[5] = w.ed[2];
  out[6] = w.n.x, out[7] = w.n.y, out[8] = w.n.z, out[9] = w.dist;
}
void or_window_f64(const double *verts, const double *hit, double sigma, double *out) {
  V3d V[3] = {ld3d<double>(verts), ld3d<double>(verts + 3), ld3d<double>(verts + 6)};
  Win<double> w;
  window<double>(V, ld3d<double>(hit), sigma, w, nullptr);
  out[0] = w.Cx, out[1] = w.phi_final, out[2] = w.phi_max;
  out[3] = w.ed[0], out[4] = w.ed[1], out[5] = w.ed[2];
  out[6] = w.n.x, out[7] = w.n.y, out[8] = w.n.z, out[9] = w.dist;
}
void or_sh_f32(const float *sh, int deg, const float *dir, float *rgb, int *clamped) {
  bool cl[3];
  V3f c = sh_color<float>(sh, deg, ld3<float>(dir), cl);
  rgb[0] = c.x, rgb[1] = c.y, rgb[2] = c.z;
  for (int i = 0; i < 3; i++) clamped[i] = cl[i];
}
void or_sh_f64(const double *sh, int deg, const double *dir, double *rgb, int *clamped) {
  bool cl[3];
  V3d c = sh_color<double>(sh, deg, ld3d<double>(dir), cl);
  rgb[0] = c.x, rgb[1] = c.y, rgb[2] = c.z;
  for (int i = 0; i < 3; i++) clamped[i] = cl[i];
}

void *or_scene_create(const float *verts, int num_verts, const int *npts, const int *cumsum, int P,
                      const float *shs, int M, int sh_degree, const float *opac, const float *sigma,
                      int build) {
  Scene *s = new Scene();
  s->verts = verts;
  s->V = num_verts;
  s->npts = npts;
  s->cumsum = cumsum;
  s->P = P;
  s->shs = shs;
  s->M = M;
  s->deg = sh_degree;
  s->opac = opac;
  s->sigma = sigma;
  if (build) build_bvh(*s);
  return s;
}
void or_scene_free(void *scene) {
  delete static_cast<Scene *>(scene);
}
double or_scene_bvh_seconds(void *scene) {
  return static_cast<Scene *>(scene)->bvh_seconds;
}

int or_gather(void *scene, const float *o, const float *d, float tlo, long long plo, int k, float far_t,
              int use_bvh, int *out_prim, float *out_t) {
  const Scene &s = *static_cast<Scene *>(scene);
  RayPre r = prep_ray(ld3<float>(o), ld3<float>(d));
  std::vector<Hit> buf(k);
  gather(s, r, tlo, plo, k, far_t, use_bvh != 0, buf.data());
  int n = 0;
  for (int i = 0; i < k; i++) {
    out_prim[i] = (int)buf[i].prim;
    out_t[i] = buf[i].t;
    if (buf[i].prim != kInvalid) n++;
  }
  return n;
}

int or_forward_f32(void *scene, const float *o, const float *d, int R, const or_opts *opt, float *color,
                   float *others, float *max_blending, float *density, float *scaling, int *stats) {
  const Scene &s = *static_cast<Scene *>(scene);
  Opts op = to_opts(opt);
#pragma omp parallel for schedule(dynamic, 64)
  for (int pix = 0; pix < R; pix++) {
    V3f ro = ld3<float>(o + 3 * (size_t)pix), rd = ld3<float>(d + 3 * (size_t)pix);
    float C[3] = {0, 0, 0}, N[3] = {0, 0, 0}, T = 1.0f;
    int consumed = 0, blended = 0;
    FwdSinks sk;
    sk.maxb = max_blending;
    sk.density = density;
    sk.scaling = scaling;
    sk.scaling_org = ro;
    auto seg = [&](V3f so, V3f sd, double, bool env, float *Cs = nullptr, float *Ts = nullptr) {
      if (env)
        trace_segment_f32(s, op, so, sd, Cs, *Ts, N, consumed, blended, nullptr, nullptr);
      else
        trace_segment_f32(s, op, so, sd, C, T, N, consumed, blended, op.variant == 0 ? &sk : nullptr, nullptr,
                          op.tstart ? std::fmax(op.tstart[pix], 0.0f) : 0.0f);
    };
    if (op.variant == 4) {
      BounceAcc acc;
      trace_segment_f32(s, op, ro, rd, C, T, N, consumed, blended, &sk, nullptr, 0.0f, &acc);
      const BounceGeom g = bounce_geometry(ro, rd, T, N, acc);
      float Lr[3] = {0, 0, 0}, Lt[3] = {0, 0, 0};
      if (g.valid) {
        const float t0 = bounce_secondary_start(op);
        float Cr[3] = {0, 0, 0}, Tr = 1.0f, Nr[3] = {0, 0, 0}, Er[3];
        trace_segment_f32(s, op, g.orr, g.dr, Cr, Tr, Nr, consumed, blended, nullptr, nullptr, t0);
        env_f32(op, g.dr, Er);
        float Ct[3] = {0, 0, 0}, Tt = 1.0f, Nt[3] = {0, 0, 0}, Et[3];
        trace_segment_f32(s, op, g.ot, g.dt, Ct, Tt, Nt, consumed, blended, nullptr, nullptr, t0);
        env_f32(op, g.dt, Et);
        for (int c = 0; c < 3; c++) Lr[c] = Cr[c] + Tr * Er[c], Lt[c] = Ct[c] + Tt * Et[c];
      }
      float E[3];
      env_f32(op, rd, E);
      for (int c = 0; c < 3; c++)
        color[(size_t)c * R + pix] = ((C[c] + acc.R * Lr[c]) + acc.Q * Lt[c]) + T * E[c];
    } else {
      pixel_segments(op, pix, ro, rd, seg, C, T);
      for (int c = 0; c < 3; c++) color[(size_t)c * R + pix] = C[c] + T * op.bg[c];
    }
    if ((op.variant == 0 || op.variant == 4) && others) {
      others[0 * (size_t)R + pix] = 0.0f;
      others[1 * (size_t)R + pix] = 1.0f - T;
      for (int c = 0; c < 3; c++) others[(size_t)(2 + c) * R + pix] = N[c];
      others[5 * (size_t)R + pix] = 0.0f;
      others[6 * (size_t)R + pix] = 0.0f;
    }
    if (stats) {
      stats[2 * (size_t)pix] = consumed;
      stats[2 * (size_t)pix + 1] = blended;
    }
  }
  return 0;
}

// backward.cu:191-350 in f32 (forward variant only), gradients accumulated with atomics.
int or_backward_f32(void *scene, const float *o, const float *d, int R, const or_opts *opt, const float *color,
                    const float *others, const float *dL_dpix, const float *dL_dothers, float *dverts,
                    float *dsigma, float *dopac, float *dsh) {
  const Scene &s = *static_cast<Scene *>(scene);
  Opts op = to_opts(opt);
  const int k = op.k;
#pragma omp parallel for schedule(dynamic, 64)
  for (int pix = 0; pix < R; pix++) {
    V3f ro = ld3<float>(o + 3 * (size_t)pix), rd = ld3<float>(d + 3 * (size_t)pix);
    RayPre r = prep_ray(ro, rd);
    float final_depth = others[pix];
    V3f final_normal = mk(others[2 * (size_t)R + pix], others[3 * (size_t)R + pix], others[4 * (size_t)R + pix]);
    float final_color[3], dL_dpixel[3];
    for (int c = 0; c < 3; c++) {
      final_color[c] = color[(size_t)c * R + pix];
      dL_dpixel[c] = dL_dpix[(size_t)c * R + pix];
    }
    float dL_ddepth = dL_dothers[pix];
    V3f dL_dn2 = mk(dL_dothers[2 * (size_t)R + pix], dL_dothers[3 * (size_t)R + pix], dL_dothers[4 * (size_t)R + pix]);
    float C[3] = {0, 0, 0}, N[3] = {0, 0, 0}, D = 0.0f, T = 1.0f;
    float last = 0.0f, tlo = 0.0f;
    long long plo = -1;
    std::vector<Hit> buf(k);
    bool done = false;
    while (!done) {
      gather(s, r, tlo, plo, k, op.far_t, op.bvh, buf.data());
      for (int i = 0; i < k && !done; i++) {
        Hit h = buf[i];
        if (h.prim == kInvalid) {
          done = true;
          break;
        }
        int prim = (int)h.prim;
        float t = h.t;
        last = std::fmax((float)((double)last + kStepEps), t);
        if (t < op.near_t) continue;
        int off = s.cumsum[prim];
        V3f V[3] = {s.vtx(prim, 0), s.vtx(prim, 1), s.vtx(prim, 2)};
        Win<float> w;
        window32(V, ro, rd, t, s.sigma[prim], w);
        float opacity = s.opac[prim];
        float alpha = std::fmin(0.99f, opacity * w.Cx);
        if (alpha < 1.0f / 255.0f) continue;
        float testT = T * (1.0f - alpha);
        if (testT < op.cutoff) {
          done = true;
          break;
        }
        float bw = alpha * T;
        bool cl[3];
        V3f col = sh_color<float>(s.shs + (size_t)prim * s.M * 3, s.deg, rd, cl);
        C[0] += col.x * bw, C[1] += col.y * bw, C[2] += col.z * bw;
        N[0] += w.n.x * bw, N[1] += w.n.y * bw, N[2] += w.n.z * bw;
        D += bw * t;
        float dL_dcolor[3];
        float cc[3] = {col.x, col.y, col.z};
        float dL_dalpha = 0.0f;
        float inv1a = 1.0f / (1.0f - alpha);
        for (int c = 0; c < 3; c++) {
          dL_dcolor[c] = dL_dpixel[c] * bw;
          dL_dalpha += dL_dpixel[c] * (T * cc[c] - (final_color[c] - C[c]) * inv1a);
        }
        dL_dalpha += dL_ddepth * (T * t - (final_depth - D) * inv1a);
        V3f Nf = mk(N[0], N[1], N[2]);
        dL_dalpha += sumv(hmul(dL_dn2, T * w.n - (final_normal - Nf) * inv1a));
        atomic_add_f(dopac + prim, dL_dalpha * w.Cx);
        float dL_dC = opacity * dL_dalpha;
        if (w.phi_final > 0.0f) atomic_add_f(dsigma + prim, (dL_dC * w.Cx) * tt_logf(w.phi_final));
        float dL_dphi = ((s.sigma[prim] / w.phi_max) * dL_dC) * w.Cx;
        V3f dL_dn3 = dL_dn2 * bw;
        window_bwd<float>(V, w.rel, dL_dphi, w.phi_max, w.ed, dL_dn3, [&](int j, V3f g) {
          float *p = dverts + 3 * (size_t)(off + j);
          atomic_add_f(p, g.x), atomic_add_f(p + 1, g.y), atomic_add_f(p + 2, g.z);
        });
        sh_bwd<float>(s.deg, rd, cl, dL_dcolor, [&](int j, V3f g) {
          float *p = dsh + ((size_t)prim * s.M + j) * 3;
          atomic_add_f(p, g.x), atomic_add_f(p + 1, g.y), atomic_add_f(p + 2, g.z);
        });
        T = testT;
      }
      if (done) break;
      if (op.mode == 1) {
        tlo = buf[k - 1].t;
        plo = (long long)buf[k - 1].prim;
      } else if (last > tlo) {
        tlo = last;
        plo = -1;
      } else {
        plo = (long long)buf[k - 1].prim;
      }
    }
  }
  return 0;
}

// ------------------------------------------------------------ log + f64
struct Log {
  const Scene *s;
  Opts op;
  int R;
  std::vector<Segment> segs;
  std::vector<std::vector<int>> pix_segs;  // per pixel: segment indices
  std::vector<int> env_pix;                // 1 where the pixel took the envlight sphere branch
};

void *or_record(void *scene, const float *o, const float *d, int R, const or_opts *opt) {
  const Scene &s = *static_cast<Scene *>(scene);
  Log *lg = new Log();
  lg->s = &s;
  lg->op = to_opts(opt);
  lg->R = R;
  std::vector<std::vector<Segment>> per(R);
  lg->env_pix.assign(R, 0);
  const Opts &op = lg->op;
#pragma omp parallel for schedule(dynamic, 64)
  for (int pix = 0; pix < R; pix++) {
    V3f ro = ld3<float>(o + 3 * (size_t)pix), rd = ld3<float>(d + 3 * (size_t)pix);
    float C[3] = {0, 0, 0}, T = 1.0f;
    int consumed = 0, blended = 0;
    auto seg = [&](V3f so, V3f sd, double wgt, bool env, float *Cs = nullptr, float *Ts = nullptr) {
      Segment sg;
      sg.pix = pix;
      sg.o = so;
      sg.d = sd;
      sg.weight = wgt;
      float Cl[3] = {0, 0, 0}, Tl = 1.0f, Nl[3] = {0, 0, 0};
      trace_segment_f32(s, op, so, sd, Cl, Tl, Nl, consumed, blended, nullptr, &sg.hits,
                        (!env && op.tstart) ? std::fmax(op.tstart[pix], 0.0f) : 0.0f);
      if (env) {
        for (int c = 0; c < 3; c++) Cs[c] = Cl[c];
        *Ts = Tl;
        lg->env_pix[pix] = 1;
      }
      per[pix].push_back(std::move(sg));
    };
    if (op.variant == 4) {
      Segment sg;
      sg.pix = pix, sg.o = ro, sg.d = rd, sg.weight = 1.0, sg.role = 0;
      float Nl[3] = {0, 0, 0};
      BounceAcc acc;
      trace_segment_f32(s, op, ro, rd, C, T, Nl, consumed, blended, nullptr, &sg.hits, 0.0f, &acc);
      per[pix].push_back(std::move(sg));
      const BounceGeom g = bounce_geometry(ro, rd, T, Nl, acc);
      if (g.valid) {
        for (int role = 1; role <= 2; role++) {
          Segment sb;
          sb.pix = pix, sb.o = role == 1 ? g.orr : g.ot, sb.d = role == 1 ? g.dr : g.dt, sb.weight = 1.0;
          sb.role = role;
          float Cl[3] = {0, 0, 0}, Tl = 1.0f, Nb[3] = {0, 0, 0};
          trace_segment_f32(s, op, sb.o, sb.d, Cl, Tl, Nb, consumed, blended, nullptr, &sb.hits,
                            bounce_secondary_start(op));
          per[pix].push_back(std::move(sb));
        }
      }
      continue;
    }
    pixel_segments(op, pix, ro, rd, seg, C, T);
  }
  lg->pix_segs.resize(R);
  for (int pix = 0; pix < R; pix++)
    for (auto &sg : per[pix]) {
      lg->pix_segs[pix].push_back((int)lg->segs.size());
      lg->segs.push_back(std::move(sg));
    }
  return lg;
}
void or_log_free(void *log) {
  delete static_cast<Log *>(log);
}
long long or_log_size(void *log) {
  long long n = 0;
  for (auto &sg : static_cast<Log *>(log)->segs) n += (long long)sg.hits.size();
  return n;
}
long long or_log_segments(void *log) {
  return (long long)static_cast<Log *>(log)->segs.size();
}
void or_log_export(void *log, long long *seg_offsets, int *seg_pix, int *prims, float *ts, int *codes) {
  Log &lg = *static_cast<Log *>(log);
  long long n = 0;
  for (size_t i = 0; i < lg.segs.size(); i++) {
    seg_offsets[i] = n;
    seg_pix[i] = lg.segs[i].pix;
    for (auto &h : lg.segs[i].hits) {
      prims[n] = (int)h.prim;
      ts[n] = h.t;
      codes[n] = h.code;
      n++;
    }
  }
  seg_offsets[lg.segs.size()] = n;
}

namespace {
struct P64 {
  const double *verts, *opac, *sigma, *shs;
};

// Replay of one segment in double with frozen decisions, t and incenter distance
// (refl: also the sums of w*rho and w*(1-rho) of the bounce variant).
void replay_segment_f64(const Log &lg, const Segment &sg, const P64 &p, double C[3], double &T, double N[3],
                        const double *refl = nullptr, double *Rw = nullptr, double *Qw = nullptr) {
  const Scene &s = *lg.s;
  V3d o = mk<double>(sg.o.x, sg.o.y, sg.o.z), d = mk<double>(sg.d.x, sg.d.y, sg.d.z);
  for (const LogHit &h : sg.hits) {
    if (h.code == 3) return;
    if (h.code != 2) continue;
    int prim = (int)h.prim;
    int off = s.cumsum[prim];
    V3d V[3] = {ld3d<double>(p.verts + 3 * (size_t)off), ld3d<double>(p.verts + 3 * (size_t)(off + 1)),
                ld3d<double>(p.verts + 3 * (size_t)(off + 2))};
    V3d hit = o + (double)h.t * d;
    Win<double> w;
    window<double>(V, hit, p.sigma[prim], w, &h.dist);
    double alpha = std::fmin(0.99, p.opac[prim] * w.Cx);
    double bw = alpha * T;
    N[0] += w.n.x * bw, N[1] += w.n.y * bw, N[2] += w.n.z * bw;
    V3d col = sh_color<double>(p.shs + (size_t)prim * s.M * 3, s.deg, d, nullptr);
    C[0] += col.x * bw, C[1] += col.y * bw, C[2] += col.z * bw;
    if (refl) *Rw += bw * refl[prim], *Qw += bw * (1.0 - refl[prim]);
    T = T * (1.0 - alpha);
  }
}

struct B64 {
  const double *refl, *env;
  int W, H;
};

// Bounce variant in double: per pixel the secondary composites L = C + T E(dir), then
// colour = C_p + R L_r + Q L_t + T_p E(d) (R, Q: sums of w rho, w (1 - rho)).
void forward_f64_bounce(const Log &lg, const P64 &p, const B64 &b, double *color, double *others,
                        double *Lbuf = nullptr) {
  const int R = lg.R;
#pragma omp parallel for schedule(dynamic, 64)
  for (int pix = 0; pix < R; pix++) {
    double C[3] = {0, 0, 0}, N[3] = {0, 0, 0}, T = 1.0, Rw = 0.0, Qw = 0.0;
    double L[2][3] = {{0, 0, 0}, {0, 0, 0}};
    const Segment *prim = nullptr;
    for (int si : lg.pix_segs[pix]) {
      const Segment &sg = lg.segs[si];
      if (sg.role == 0) {
        prim = &sg;
        replay_segment_f64(lg, sg, p, C, T, N, b.refl, &Rw, &Qw);
      } else {
        double Cs[3] = {0, 0, 0}, Ns[3] = {0, 0, 0}, Ts = 1.0, E[3];
        replay_segment_f64(lg, sg, p, Cs, Ts, Ns);
        env_f64(b.env, b.W, b.H, sg.d, E);
        for (int c = 0; c < 3; c++) L[sg.role - 1][c] = Cs[c] + Ts * E[c];
      }
    }
    double E0[3] = {0, 0, 0};
    if (prim) env_f64(b.env, b.W, b.H, prim->d, E0);
    for (int c = 0; c < 3; c++) color[(size_t)c * R + pix] = C[c] + Rw * L[0][c] + Qw * L[1][c] + T * E0[c];
    if (Lbuf)
      for (int c = 0; c < 3; c++) Lbuf[6 * (size_t)pix + c] = L[0][c], Lbuf[6 * (size_t)pix + 3 + c] = L[1][c];
    if (others) {
      others[pix] = 0.0;
      others[(size_t)R + pix] = 1.0 - T;
      for (int c = 0; c < 3; c++) others[(size_t)(2 + c) * R + pix] = N[c];
      others[5 * (size_t)R + pix] = 0.0;
      others[6 * (size_t)R + pix] = 0.0;
    }
  }
}

struct G64 {
  double *dverts, *dsigma, *dopac, *dsh, *drefl;
};

// One blended hit's vertex gradient (9 values: the triangle's three vertices), added to the
// gradient rows of the triangle. Test aid (not part of the restated algorithm): with
// g_abs_vertex set, every one of the 9 entries receives instead the hit's largest |value|.
// Summed over the hits, that is the scale FP32 evaluation error is relative to: the terms
// of one hit are differences of quantities of that size (a tiny triangle hit near one edge
// gives two vertices gradients of ~1e3 and the third a residual of ~1e-3), and the hits'
// terms may cancel in the sum.
inline void add_vertex_hit(double *rows, const double hv[9]) {
  if (g_abs_vertex) {
    double m = 0.0;
    for (int i = 0; i < 9; i++) m = std::fmax(m, std::fabs(hv[i]));
    for (int i = 0; i < 9; i++) rows[i] += m;
  } else {
    for (int i = 0; i < 9; i++) rows[i] += hv[i];
  }
}

// Backward of one replayed segment in double (backward.cu:191-350 per hit) for an
// upstream gradient g on the segment's output colour `fin` (composite incl. its
// background), optional depth / normal terms, and (bounce primary) per-hit colours
// extended by rho L_r + (1 - rho) L_t with their reflectivity gradient. Returns the
// segment's final transmittance.
double backward_segment_f64(const Log &lg, const Segment &sg, const P64 &p, const double g[3], const double fin[3],
                            double dL_ddepth, V3d dL_dn2, double final_depth, V3d final_normal, const double *refl,
                            const double *Lr, const double *Lt, const G64 &out) {
  const Scene &s = *lg.s;
  V3d o = mk<double>(sg.o.x, sg.o.y, sg.o.z), d = mk<double>(sg.d.x, sg.d.y, sg.d.z);
  double C[3] = {0, 0, 0}, N[3] = {0, 0, 0}, D = 0.0, T = 1.0;
  for (const LogHit &h : sg.hits) {
    if (h.code == 3) break;
    if (h.code != 2) continue;
    int prim = (int)h.prim;
    int off = s.cumsum[prim];
    V3d V[3] = {ld3d<double>(p.verts + 3 * (size_t)off), ld3d<double>(p.verts + 3 * (size_t)(off + 1)),
                ld3d<double>(p.verts + 3 * (size_t)(off + 2))};
    double t = (double)h.t;
    V3d hit = o + t * d;
    Win<double> w;
    window<double>(V, hit, p.sigma[prim], w, &h.dist);
    double opacity = p.opac[prim];
    double alpha = std::fmin(0.99, opacity * w.Cx);
    double bw = alpha * T;
    bool cl[3];
    V3d col = sh_color<double>(p.shs + (size_t)prim * s.M * 3, s.deg, d, cl);
    double cc[3] = {col.x, col.y, col.z};
    if (Lr) {
      const double rho = refl[prim];
      double dr = 0.0;
      for (int c = 0; c < 3; c++) {
        dr += g[c] * (Lr[c] - Lt[c]);
        cc[c] += rho * Lr[c] + (1.0 - rho) * Lt[c];
      }
      out.drefl[prim] += bw * dr;
    }
    C[0] += cc[0] * bw, C[1] += cc[1] * bw, C[2] += cc[2] * bw;
    N[0] += w.n.x * bw, N[1] += w.n.y * bw, N[2] += w.n.z * bw;
    D += bw * t;
    double dL_dcolor[3];
    double dL_dalpha = 0.0, inv1a = 1.0 / (1.0 - alpha);
    for (int c = 0; c < 3; c++) {
      dL_dcolor[c] = g[c] * bw;
      dL_dalpha += g[c] * (T * cc[c] - (fin[c] - C[c]) * inv1a);
    }
    dL_dalpha += dL_ddepth * (T * t - (final_depth - D) * inv1a);
    V3d Nf = mk(N[0], N[1], N[2]);
    dL_dalpha += sumv(hmul(dL_dn2, T * w.n - (final_normal - Nf) * inv1a));
    out.dopac[prim] += dL_dalpha * w.Cx;
    double dL_dC = opacity * dL_dalpha;
    if (w.phi_final > 0.0) out.dsigma[prim] += dL_dC * w.Cx * std::log(w.phi_final);
    double dL_dphi = p.sigma[prim] / w.phi_max * dL_dC * w.Cx;
    V3d dL_dn3 = dL_dn2 * bw;
    double hv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    window_bwd<double>(
        V, w.rel, dL_dphi, w.phi_max, w.ed, dL_dn3,
        [&](int j, V3d gv) { hv[3 * j] += gv.x, hv[3 * j + 1] += gv.y, hv[3 * j + 2] += gv.z; }, h.edges);
    add_vertex_hit(out.dverts + 3 * (size_t)off, hv);
    sh_bwd<double>(s.deg, d, cl, dL_dcolor, [&](int j, V3d gv) {
      double *q = out.dsh + ((size_t)prim * s.M + j) * 3;
      q[0] += gv.x, q[1] += gv.y, q[2] += gv.z;
    });
    T = T * (1.0 - alpha);
  }
  return T;
}

void env_scatter_f64(double *denv, int W, int H, V3f d, const double g[3]) {
  const EnvTaps e = env_taps(d.x, d.y, d.z, W, H);
  for (int j = 0; j < 4; j++)
    for (int c = 0; c < 3; c++) denv[3 * e.idx[j] + c] += e.w[j] * g[c];
}

void forward_f64(const Log &lg, const P64 &p, double *color, double *others) {
  const int R = lg.R;
  const double *bg = nullptr;
  double bgd[3] = {lg.op.bg[0], lg.op.bg[1], lg.op.bg[2]};
  bg = bgd;
#pragma omp parallel for schedule(dynamic, 64)
  for (int pix = 0; pix < R; pix++) {
    double C[3] = {0, 0, 0}, N[3] = {0, 0, 0}, T = 1.0;
    if (lg.env_pix[pix]) {
      for (int si : lg.pix_segs[pix]) {
        const Segment &sg = lg.segs[si];
        double Cs[3] = {0, 0, 0}, Ns[3] = {0, 0, 0}, Ts = 1.0;
        replay_segment_f64(lg, sg, p, Cs, Ts, Ns);
        for (int c = 0; c < 3; c++) C[c] += (Cs[c] + Ts * bg[c]) * sg.weight;
      }
      T = 0.0;
    } else {
      for (int si : lg.pix_segs[pix]) replay_segment_f64(lg, lg.segs[si], p, C, T, N);
    }
    for (int c = 0; c < 3; c++) color[(size_t)c * R + pix] = C[c] + T * bg[c];
    if (others) {
      others[pix] = 0.0;
      others[(size_t)R + pix] = 1.0 - T;
      for (int c = 0; c < 3; c++) others[(size_t)(2 + c) * R + pix] = N[c];
      others[5 * (size_t)R + pix] = 0.0;
      others[6 * (size_t)R + pix] = 0.0;
    }
  }
}
}  // namespace

int or_forward_f64(void *log, const double *verts, const double *opac, const double *sigma, const double *shs,
                   double *color, double *others) {
  const Log &lg = *static_cast<Log *>(log);
  forward_f64(lg, P64{verts, opac, sigma, shs}, color, others);
  return 0;
}

double or_loss_f64(void *log, const double *verts, const double *opac, const double *sigma, const double *shs,
                   const double *dL_dpix, const double *dL_dothers) {
  const Log &lg = *static_cast<Log *>(log);
  const int R = lg.R;
  std::vector<double> color((size_t)3 * R), others((size_t)7 * R);
  forward_f64(lg, P64{verts, opac, sigma, shs}, color.data(), others.data());
  double L = 0.0;
  for (size_t i = 0; i < (size_t)3 * R; i++) L += dL_dpix[i] * color[i];
  if (dL_dothers)
    for (size_t i = 0; i < (size_t)7 * R; i++) L += dL_dothers[i] * others[i];
  return L;
}

// backward.cu:191-350 in double over the frozen log (forward variant only).
int or_backward_f64(void *log, const double *verts, const double *opac, const double *sigma, const double *shs,
                    const double *dL_dpix, const double *dL_dothers, double *dverts, double *dsigma,
                    double *dopac, double *dsh) {
  const Log &lg = *static_cast<Log *>(log);
  if (lg.op.variant != 0) return -1;
  const Scene &s = *lg.s;
  const int R = lg.R;
  P64 p{verts, opac, sigma, shs};
  std::vector<double> color((size_t)3 * R), others((size_t)7 * R);
  forward_f64(lg, p, color.data(), others.data());
  // serial accumulation keeps the double sums deterministic
  for (int pix = 0; pix < R; pix++) {
    if (lg.pix_segs[pix].empty()) continue;
    const Segment &sg = lg.segs[lg.pix_segs[pix][0]];
    V3d o = mk<double>(sg.o.x, sg.o.y, sg.o.z), d = mk<double>(sg.d.x, sg.d.y, sg.d.z);
    double final_depth = others[pix];
    V3d final_normal = mk(others[2 * (size_t)R + pix], others[3 * (size_t)R + pix], others[4 * (size_t)R + pix]);
    double final_color[3], dL_dpixel[3];
    for (int c = 0; c < 3; c++) {
      final_color[c] = color[(size_t)c * R + pix];
      dL_dpixel[c] = dL_dpix[(size_t)c * R + pix];
    }
    double dL_ddepth = dL_dothers ? dL_dothers[pix] : 0.0;
    V3d dL_dn2 = dL_dothers ? mk(dL_dothers[2 * (size_t)R + pix], dL_dothers[3 * (size_t)R + pix],
                                 dL_dothers[4 * (size_t)R + pix])
                            : mk(0.0, 0.0, 0.0);
    double C[3] = {0, 0, 0}, N[3] = {0, 0, 0}, D = 0.0, T = 1.0;
    for (const LogHit &h : sg.hits) {
      if (h.code == 3) break;
      if (h.code != 2) continue;
      int prim = (int)h.prim;
      int off = s.cumsum[prim];
      V3d V[3] = {ld3d<double>(verts + 3 * (size_t)off), ld3d<double>(verts + 3 * (size_t)(off + 1)),
                  ld3d<double>(verts + 3 * (size_t)(off + 2))};
      double t = (double)h.t;
      V3d hit = o + t * d;
      Win<double> w;
      window<double>(V, hit, sigma[prim], w, &h.dist);
      double opacity = opac[prim];
      double alpha = std::fmin(0.99, opacity * w.Cx);
      double bw = alpha * T;
      bool cl[3];
      V3d col = sh_color<double>(shs + (size_t)prim * s.M * 3, s.deg, d, cl);
      C[0] += col.x * bw, C[1] += col.y * bw, C[2] += col.z * bw;
      N[0] += w.n.x * bw, N[1] += w.n.y * bw, N[2] += w.n.z * bw;
      D += bw * t;
      double dL_dcolor[3], cc[3] = {col.x, col.y, col.z};
      double dL_dalpha = 0.0, inv1a = 1.0 / (1.0 - alpha);
      for (int c = 0; c < 3; c++) {
        dL_dcolor[c] = dL_dpixel[c] * bw;
        dL_dalpha += dL_dpixel[c] * (T * cc[c] - (final_color[c] - C[c]) * inv1a);
      }
      dL_dalpha += dL_ddepth * (T * t - (final_depth - D) * inv1a);
      V3d Nf = mk(N[0], N[1], N[2]);
      dL_dalpha += sumv(hmul(dL_dn2, T * w.n - (final_normal - Nf) * inv1a));
      dopac[prim] += dL_dalpha * w.Cx;
      double dL_dC = opacity * dL_dalpha;
      if (w.phi_final > 0.0) dsigma[prim] += dL_dC * w.Cx * std::log(w.phi_final);
      double dL_dphi = sigma[prim] / w.phi_max * dL_dC * w.Cx;
      V3d dL_dn3 = dL_dn2 * bw;
      double hv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      window_bwd<double>(
          V, w.rel, dL_dphi, w.phi_max, w.ed, dL_dn3,
          [&](int j, V3d g) { hv[3 * j] += g.x, hv[3 * j + 1] += g.y, hv[3 * j + 2] += g.z; }, h.edges);
      add_vertex_hit(dverts + 3 * (size_t)off, hv);
      sh_bwd<double>(s.deg, d, cl, dL_dcolor, [&](int j, V3d g) {
        double *q = dsh + ((size_t)prim * s.M + j) * 3;
        q[0] += g.x, q[1] += g.y, q[2] += g.z;
      });
      T = T * (1.0 - alpha);
    }
  }
  return 0;
}

int or_forward_f64_bounce(void *log, const double *verts, const double *opac, const double *sigma,
                          const double *shs, const double *refl, const double *env, double *color, double *others) {
  const Log &lg = *static_cast<Log *>(log);
  if (lg.op.variant != 4) return -1;
  forward_f64_bounce(lg, P64{verts, opac, sigma, shs}, B64{refl, env, lg.op.env_w, lg.op.env_h}, color, others);
  return 0;
}

double or_loss_f64_bounce(void *log, const double *verts, const double *opac, const double *sigma,
                          const double *shs, const double *refl, const double *env, const double *dL_dpix,
                          const double *dL_dothers) {
  const Log &lg = *static_cast<Log *>(log);
  const int R = lg.R;
  std::vector<double> color((size_t)3 * R), others((size_t)7 * R);
  forward_f64_bounce(lg, P64{verts, opac, sigma, shs}, B64{refl, env, lg.op.env_w, lg.op.env_h}, color.data(),
                     others.data());
  double L = 0.0;
  for (size_t i = 0; i < (size_t)3 * R; i++) L += dL_dpix[i] * color[i];
  if (dL_dothers)
    for (size_t i = 0; i < (size_t)7 * R; i++) L += dL_dothers[i] * others[i];
  return L;
}

// Bounce variant backward in double over the frozen log: the primary segment with the
// extended colours, its environment term, then each secondary segment with upstream
// R g (reflected) or Q g (refracted) and its environment term.
int or_backward_f64_bounce(void *log, const double *verts, const double *opac, const double *sigma,
                           const double *shs, const double *refl, const double *env, const double *dL_dpix,
                           const double *dL_dothers, double *dverts, double *dsigma, double *dopac, double *dsh,
                           double *drefl, double *denv) {
  const Log &lg = *static_cast<Log *>(log);
  if (lg.op.variant != 4) return -1;
  const int R = lg.R, W = lg.op.env_w, H = lg.op.env_h;
  P64 p{verts, opac, sigma, shs};
  B64 b{refl, env, W, H};
  std::vector<double> color((size_t)3 * R), others((size_t)7 * R), L((size_t)6 * R);
  forward_f64_bounce(lg, p, b, color.data(), others.data(), L.data());
  G64 out{dverts, dsigma, dopac, dsh, drefl};
  for (int pix = 0; pix < R; pix++) {
    const Segment *prim = nullptr;
    for (int si : lg.pix_segs[pix])
      if (lg.segs[si].role == 0) prim = &lg.segs[si];
    if (!prim) continue;
    double g[3], fin[3];
    for (int c = 0; c < 3; c++) g[c] = dL_dpix[(size_t)c * R + pix], fin[c] = color[(size_t)c * R + pix];
    double dL_ddepth = dL_dothers ? dL_dothers[pix] : 0.0;
    V3d dL_dn2 = dL_dothers ? mk(dL_dothers[2 * (size_t)R + pix], dL_dothers[3 * (size_t)R + pix],
                                 dL_dothers[4 * (size_t)R + pix])
                            : mk(0.0, 0.0, 0.0);
    V3d final_normal = mk(others[2 * (size_t)R + pix], others[3 * (size_t)R + pix], others[4 * (size_t)R + pix]);
    const double *Lr = &L[6 * (size_t)pix], *Lt = &L[6 * (size_t)pix + 3];
    // R and Q of the primary (the same sums the forward formed)
    double Cq[3] = {0, 0, 0}, Nq[3] = {0, 0, 0}, Tq = 1.0, Rw = 0.0, Qw = 0.0;
    replay_segment_f64(lg, *prim, p, Cq, Tq, Nq, refl, &Rw, &Qw);
    const double Tp = backward_segment_f64(lg, *prim, p, g, fin, dL_ddepth, dL_dn2, others[pix], final_normal, refl,
                                           Lr, Lt, out);
    double ge[3] = {Tp * g[0], Tp * g[1], Tp * g[2]};
    env_scatter_f64(denv, W, H, prim->d, ge);
    for (int si : lg.pix_segs[pix]) {
      const Segment &sg = lg.segs[si];
      if (sg.role == 0) continue;
      const double wgt = sg.role == 1 ? Rw : Qw;
      const double *Ls = sg.role == 1 ? Lr : Lt;
      double gs[3] = {wgt * g[0], wgt * g[1], wgt * g[2]};
      const double Ts = backward_segment_f64(lg, sg, p, gs, Ls, 0.0, mk(0.0, 0.0, 0.0), 0.0, mk(0.0, 0.0, 0.0),
                                             nullptr, nullptr, nullptr, out);
      double gse[3] = {Ts * gs[0], Ts * gs[1], Ts * gs[2]};
      env_scatter_f64(denv, W, H, sg.d, gse);
    }
  }
  return 0;
}

void or_log_segment_rays(void *log, float *o, float *d, int *role) {
  const Log &lg = *static_cast<Log *>(log);
  for (size_t i = 0; i < lg.segs.size(); i++) {
    const Segment &sg = lg.segs[i];
    o[3 * i] = sg.o.x, o[3 * i + 1] = sg.o.y, o[3 * i + 2] = sg.o.z;
    d[3 * i] = sg.d.x, d[3 * i + 1] = sg.d.y, d[3 * i + 2] = sg.d.z;
    role[i] = sg.role;
  }
}

void or_env_taps(const float *d, int env_w, int env_h, int *idx, double *w) {
  const EnvTaps e = env_taps(d[0], d[1], d[2], env_w, env_h);
  for (int j = 0; j < 4; j++) idx[j] = e.idx[j], w[j] = e.w[j];
}

}  // extern "C"
