/*
 * CPU oracle for the triangle-splat tracer -- TEST INFRASTRUCTURE ONLY.
 *
 * This library is the checker that the HIP product path is compared against.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it; the product (libtritracer_hip.so, diff_tritracer._C) never links, loads
 * or calls it.
 *
 * It restates, as scalar C++ with OpenMP over rays, the algorithm of the
 * reference OptiX tracer (waizui/diff-triangle-tracer):
 *   - k-closest hit gather per chunk ........ tritracer/closestk.h:17-149
 *   - forward chunk loop + outputs .......... tritracer/forward.cu:25-215
 *   - backward re-traversal + gradients ..... tritracer/backward.cu:25-391
 *   - window function, SH colour, sphere .... tritracer/common.h:16-148
 *   - derivative helpers (math only) ........ tritracer/vecmath.h:2626-2859
 *   - reflect / refract / envlight variants . tritracer/reflect.cu, refract.cu, envlight.cu
 *   - constants ............................. tritracer/auxiliary.h:9-30, config.h:10-20
 * The ray/triangle test inside optixTrace is closed NVIDIA code; the oracle
 * defines it as the watertight test of Woop, Benthin and Wald (JCGT 2013),
 * double-sided, t in the closed range of the chunk.
 *
 * Three evaluation modes:
 *   f32  : op-for-op the arithmetic the HIP kernels perform (no FMA
 *          contraction, project-written log2/exp2/sin/cos), so hit lists and
 *          per-hit decisions can be compared bit-exactly;
 *   f64  : the discrete decisions (hit set, skips, termination) recorded by
 *          the f32 pass, with all continuous math re-evaluated in double;
 *   loss : the f64 replay as a scalar loss with hit distance t and incenter
 *          distance frozen -- the surrogate whose finite differences pin the
 *          reference's partial vertex gradient.
 *
 * Parity unpinned: the refract direction (refract.cu:34-35, formula kept with
 * its quirks) at oblique incidence; only its straight-through case has an
 * independent closed form. Everything else here is pinned by known answers
 * (closed forms, libm in double, brute force, finite differences, the LCG's
 * published sequence).
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int k;            /* hits per chunk (reference CHUNK_SIZE = 16) */
  int chunk_mode;   /* 0 = compat (inclusive resume at last_hit_t), 1 = exact */
  int variant;      /* 0 forward, 1 reflect, 2 refract, 3 envlight */
  int use_bvh;      /* 1 = oracle's own BVH, 0 = brute force over all triangles */
  float cutoff;     /* transmittance cutoff (reference forward.cu:134 uses 1e-4) */
  float near_t;     /* 0.2  (auxiliary.h:29) */
  float far_t;      /* 1000 (auxiliary.h:30) */
  float bg[3];
  const float *extra; /* variant parameters: r, cx, cy, cz[, samples] */
  int extra_len;      /* number of floats in extra */
  const float *tstart; /* optional [R] (forward variant, exact mode; or_forward_f32 and or_record
                          only): each ray's trace starts at key (tstart[ray], -1) instead of 0 */
  /* variant 4 (bounce): per-triangle reflectivity [P] and index of refraction [P], equirect
     environment map [env_h][env_w][3] */
  const float *refl, *ior, *env;
  int env_w, env_h;
} or_opts;

/* ---- scene ---- */
void *or_scene_create(const float *verts, int num_verts, const int *npts, const int *cumsum, int P,
                      const float *shs, int M, int sh_degree, const float *opac, const float *sigma,
                      int build_bvh);
void or_scene_free(void *scene);
double or_scene_bvh_seconds(void *scene);
void or_set_threads(int n);
int or_get_threads(void);
/* test aid: the FP64 backwards add to each vertex-gradient entry, per blended hit, that hit's
   largest |vertex-gradient component| (1) instead of the gradient itself (0) */
void or_set_abs_vertex(int on);

/* ---- elementary pieces (single evaluations, for known-answer tests) ---- */
float or_log2f(float x);
float or_exp2f(float x);
float or_powf(float x, float y);
void or_sincosf(float x, float *s, float *c);
/* watertight intersection; returns 1 on hit and writes t */
int or_intersect(const float *o, const float *d, const float *v0, const float *v1, const float *v2, float *t);
/* envlight hemisphere samples: `count` successive cosine-weighted directions around the unit
   normal nrm[3] from one LCG state `seed` -> out[3 * count] */
void or_cos_hemi(unsigned int seed, const float *nrm, int count, float *out);
/* window function: verts[9], hit[3] -> out[0]=Cx out[1]=phi_final out[2]=phi_max out[3..5]=edge dists
   out[6..8]=face normal out[9]=incenter dist */
void or_window_f32(const float *verts, const float *hit, float sigma, float *out);
void or_window_f64(const double *verts, const double *hit, double sigma, double *out);
/* SH colour: sh[M*3], dir[3] -> rgb[3], clamped[3] */
void or_sh_f32(const float *sh, int deg, const float *dir, float *rgb, int *clamped);
void or_sh_f64(const double *sh, int deg, const double *dir, double *rgb, int *clamped);

/* one chunk gather: candidates (t,prim) with t<=far and (t>tlo or (t==tlo and prim>plo)) */
int or_gather(void *scene, const float *o, const float *d, float tlo, long long plo, int k, float far_t,
              int use_bvh, int *out_prim, float *out_t);

/* ---- f32 path (mirror of the HIP kernels) ---- */
int or_forward_f32(void *scene, const float *o, const float *d, int R, const or_opts *opt, float *color,
                   float *others, float *max_blending, float *density, float *scaling, int *stats);
int or_backward_f32(void *scene, const float *o, const float *d, int R, const or_opts *opt, const float *color,
                    const float *others, const float *dL_dpix, const float *dL_dothers, float *dverts,
                    float *dsigma, float *dopac, float *dsh);

/* ---- decision log + f64 replay ---- */
void *or_record(void *scene, const float *o, const float *d, int R, const or_opts *opt);
void or_log_free(void *log);
long long or_log_size(void *log);
long long or_log_segments(void *log);
void or_log_export(void *log, long long *seg_offsets, int *seg_pix, int *prims, float *ts, int *codes);
int or_forward_f64(void *log, const double *verts, const double *opac, const double *sigma, const double *shs,
                   double *color, double *others);
int or_backward_f64(void *log, const double *verts, const double *opac, const double *sigma, const double *shs,
                    const double *dL_dpix, const double *dL_dothers, double *dverts, double *dsigma,
                    double *dopac, double *dsh);
double or_loss_f64(void *log, const double *verts, const double *opac, const double *sigma, const double *shs,
                   const double *dL_dpix, const double *dL_dothers);
/* variant 4 (bounce) replays: reflectivity [P] and environment map [env_h][env_w][3] in double;
   gradients also to them (the bounce geometry and the IOR are held fixed) */
int or_forward_f64_bounce(void *log, const double *verts, const double *opac, const double *sigma,
                          const double *shs, const double *refl, const double *env, double *color, double *others);
double or_loss_f64_bounce(void *log, const double *verts, const double *opac, const double *sigma,
                          const double *shs, const double *refl, const double *env, const double *dL_dpix,
                          const double *dL_dothers);
int or_backward_f64_bounce(void *log, const double *verts, const double *opac, const double *sigma,
                           const double *shs, const double *refl, const double *env, const double *dL_dpix,
                           const double *dL_dothers, double *dverts, double *dsigma, double *dopac, double *dsh,
                           double *drefl, double *denv);
/* per segment of the log: ray origin [3], direction [3], role (0 primary, 1 reflected, 2 refracted) */
void or_log_segment_rays(void *log, float *o, float *d, int *role);
/* environment lookup: 4 texel indices (y * env_w + x) and bilinear weights of direction d */
void or_env_taps(const float *d, int env_w, int env_h, int *idx, double *w);

#ifdef __cplusplus
}
#endif
