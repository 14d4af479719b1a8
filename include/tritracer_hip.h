/*
 * tritracer_hip.h -- C ABI of the MI355X (gfx950) triangle-splat tracer.
 *
 * The library (libtritracer_hip.so) replaces everything the reference
 * waizui/diff-triangle-tracer runs through OptiX: the GAS build, optixTrace with
 * its anyhit k-buffer, and the forward/backward/reflect/refract/envlight
 * raygen programs. Its only caller in this project is the C++ drop-in
 * OptixTracer (tracer.h/tracer.cpp), which keeps the reference's torch-facing
 * signatures and turns nonzero status codes into exceptions.
 *
 * Conventions
 *   - Plain pointers and sizes only. Every pointer named *_dev below is device
 *     memory (HBM) on the device that was current at tt_create.
 *   - Every call takes a HIP stream as `void *stream` (NULL = the null stream)
 *     and is asynchronous on it; nothing synchronises the host unless stated.
 *   - Every entry point returns TT_OK (0) or a TT_ERR_* code; tt_last_error()
 *     returns the message of the last failure on the calling thread.
 *   - A tracer object is not thread-safe: one caller, sequential calls (as the
 *     reference's OptixTracer, tracer.h:67-126).
 *
 * Layouts (the reference's, tracer.cpp:16-142, forward.cu:161-215)
 *   ray_o, ray_d       float [H*W][3]  pixel p = y*W + x
 *   vertices           float [num_vertices][3]
 *   cumsum, npts       int   [P]       triangle i = vertices cumsum[i] .. cumsum[i]+2
 *   shs                float [P][M][3]
 *   opacities, sigma   float [P]
 *   color              float [3][H*W]  planar
 *   others             float [7][H*W]  0 depth(=0) 1 alpha 2..4 normal 5 mid-depth(=0) 6 distortion(=0)
 *   dL_dvertices       float [rows][3] rows = npts[P-1] + cumsum[P-1]
 */
#ifndef TRITRACER_HIP_H
#define TRITRACER_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TT_OK 0
#define TT_ERR_INVALID_ARG 1
#define TT_ERR_HIP 2
#define TT_ERR_STATE 3
#define TT_ERR_UNSUPPORTED 4

/* forward-program variants; the reference selects them by PTX file name
   (diff_tritracer/__init__.py:233-236 -> pipeline.cpp:186-195) */
#define TT_VARIANT_FORWARD 0  /* forward.cu  */
#define TT_VARIANT_REFLECT 1  /* reflect.cu  */
#define TT_VARIANT_REFRACT 2  /* refract.cu  */
#define TT_VARIANT_ENVLIGHT 3 /* envlight.cu */
/* no counterpart in the reference: per-triangle reflectivity / IOR, one reflected and one
   refracted ray per pixel from the primary's expected surface, equirect environment light */
#define TT_VARIANT_BOUNCE 4

/* acceleration-structure modes: the reference's rebuild_gas codes
   (diff_tritracer/__init__.py:42, tracer.cpp:181-195) */
#define TT_BUILD 0
#define TT_REFIT 1
#define TT_REUSE 2
/* no counterpart in the reference: refit, then rebuild when the refit tree has degraded (see
   tt_set_rebuild_policy) */
#define TT_REFIT_AUTO 3

/* chunk resume rule between k-closest gathers */
#define TT_CHUNK_COMPAT 0 /* resume at t >= last_hit_t (inclusive; OptiX closed ray interval,
                             closestk.h:29 with forward.cu:111) */
#define TT_CHUNK_EXACT 1  /* resume strictly after the last (t, prim); result independent of k */

typedef struct tt_tracer tt_tracer;

/* Replaces OptixTracer::OptixTracer context/pipeline creation (pipeline.cpp:267-302):
   binds to the current HIP device. */
int tt_create(tt_tracer **out);
/* Replaces OptixTracer::~OptixTracer (pipeline.cpp:304-326). Frees device memory. */
int tt_destroy(tt_tracer *tr);
/* Message of the last failed call on this thread ("" if none). */
const char *tt_last_error(void);
/* Library version string. */
const char *tt_version(void);

/* Replaces OptixTracer::build_tri_accel (pipeline.cpp:206-265) and the index-buffer
   set-up of trace_fwd (tracer.cpp:179-195). mode TT_BUILD builds an LBVH over the P
   triangles, TT_REFIT refits it to moved vertices (same P and cumsum), TT_REUSE
   does nothing. Vertices are copied into the tracer's leaf array. */
int tt_build(tt_tracer *tr, const float *vertices_dev, int num_vertices, const int *cumsum_dev, int P, int mode,
             void *stream);

typedef struct {
  int width, height;
  const float *ray_o_dev, *ray_d_dev;
  const float *vertices_dev;
  int num_vertices;
  const int *npts_dev, *cumsum_dev;
  int P;
  const float *shs_dev;
  int M, sh_degree;
  const float *opacities_dev, *sigma_dev;
  const float *bg_dev;    /* [3] */
  const float *extra_dev; /* variant parameters r, cx, cy, cz[, samples]; may be NULL for forward */
  int extra_len;          /* floats in extra */
  int variant;            /* TT_VARIANT_* */
  int k;                  /* hits per chunk: 8, 16 or 32 (reference CHUNK_SIZE = 16) */
  int chunk_mode;         /* TT_CHUNK_* */
  float cutoff;           /* transmittance cutoff (reference forward.cu:134: 1e-4) */
  float near_t, far_t;    /* 0.2, 1000 (auxiliary.h:29-30) */
  float *color_dev;       /* [3][R] written */
  float *others_dev;      /* [7][R] written by TT_VARIANT_FORWARD only (may be NULL otherwise) */
  float *max_blending_dev;/* [P] atomic max of blending weights; caller zeroes (forward only) */
  float *density_dev;     /* [P] in/out: set to 1 for blended triangles (forward only) */
  float *scaling_dev;     /* [P] in/out: filled where <= 0 (forward only) */
  int *stats_dev;         /* optional [R][2]: consumed hits, blended hits */
  int *trav_stats_dev;    /* optional [R][2]: BVH nodes visited, triangles tested (all chunks) */
  int hit_log_cap;        /* optional debug log of the first hit_log_cap consumed hits per ray */
  int *hit_log_prim_dev;  /* [R][cap] */
  float *hit_log_t_dev;   /* [R][cap] */
  int *hit_log_code_dev;  /* [R][cap] 0 near skip, 1 alpha skip, 2 blended, 3 terminated */
  int *hit_log_count_dev; /* [R] total consumed hits */
  /* Depth intervals (no counterpart in the reference; TT_VARIANT_FORWARD, TT_CHUNK_EXACT).
     interval_in: optional [R][2] (start, cap) per ray. The gather starts at t >= start; the
     first gathers stop at cap and go on past it only while the ray is still open, so the
     result equals a trace that starts at `start` (start <= near_t: the full trace; a start
     below 0 is taken as 0; a cap below start is ignored; a start beyond far_t: an empty
     trace, in both modes below).
     interval_out: optional [R][4] per ray: t of the first and of the last blended hit, t of
     the hit that ended the ray (inf if none), first blended triangle id (int bits); the
     first field is inf for a ray with no blended hit. */
  const float *interval_in_dev;
  float *interval_out_dev;
  int interval_hard_cap;         /* nonzero: the trace ends at cap (hits past it are dropped), except:
                                    a ray whose first hit at or after start lies within
                                    interval_fallback_start of it, or that has no hit up to cap, is
                                    traced again in full; a ray still open at cap that has no blended
                                    hit yet or whose last blended hit lies within
                                    interval_fallback_end of cap goes on past it */
  float interval_fallback_start, interval_fallback_end;
  /* TT_VARIANT_BOUNCE inputs. Per pixel, with the primary's blending weights w = alpha T:
     colour = C + R L_r + Q L_t + T E(d), R = sum w rho, Q = sum w (1 - rho), L = C' + T' E(d')
     of the reflected / refracted ray, E = bilinear equirect lookup; the secondary rays start
     at the expected surface (depth sum w t / (1 - T), normal sum w n turned to face the
     ray, IOR sum w eta / (1 - T), Snell with ratio 1 / IOR). others as TT_VARIANT_FORWARD. */
  const float *reflectivity_dev; /* [P] in [0, 1] */
  const float *ior_dev;          /* [P] */
  const float *envmap_dev;       /* [env_height][env_width][3] */
  int env_width, env_height;
  float *bounce_aux_dev;         /* [R][20] written: the record tt_trace_bwd reads */
  /* Stored blended hits (no counterpart in the reference, whose backward always re-traverses;
     TT_VARIANT_FORWARD and TT_VARIANT_BOUNCE). Optional: the forward writes blended hit i of
     segment s of pixel p as (triangle id, t bits) to hitlist[(s * hitlist_cap + i) * R + p] for
     i < hitlist_cap and the segment's number of blended hits to hitlist_count[s * R + p], or -1
     when it blended more than hitlist_cap. Segments: 0 the primary ray; for TT_VARIANT_BOUNCE
     also 1 the reflected and 2 the refracted ray. Handing both to tt_trace_bwd replaces its
     re-traversal by a read of the same hits (same gradients); segments with count -1 are
     re-traversed. */
  unsigned int *hitlist_dev;     /* [segments][hitlist_cap][R][2] written */
  int *hitlist_count_dev;        /* [segments][R] written */
  int hitlist_cap;
  /* Optional with hitlist_dev: each stored hit's SH colour (the FP32 value the forward blended;
     a channel clamped at zero is stored as -0.0f), at hitlist_color[3 * e + c] for the entry e
     that hitlist[2 * e] holds. Handing it to tt_trace_bwd spares the backward the SH
     coefficients' read and evaluation per hit (same gradients). */
  float *hitlist_color_dev;      /* [segments][hitlist_cap][R][3] written */
  /* Optional hipEvent_t (no counterpart in the reference): the forward waits for it on its
     stream (hipStreamWaitEvent) right before the traversal kernel, after its window-record pass
     and tile probe, which do not read interval_in. A caller that computes interval_in on
     another stream (the proposal network, tt_proposal_forward) records this event there and
     so overlaps that work with those passes. NULL: no wait. */
  void *interval_ready_event;
} tt_fwd_args;

/* Replaces OptixTracer::trace_fwd's launch (tracer.cpp:105-132) and the raygen programs
   forward.cu:161-215, reflect.cu:102-142, refract.cu:287-327, envlight.cu:190-210. */
int tt_trace_fwd(tt_tracer *tr, const tt_fwd_args *a, void *stream);

typedef struct {
  int width, height;
  const float *ray_o_dev, *ray_d_dev;
  const float *vertices_dev;
  int num_vertices;
  const int *npts_dev, *cumsum_dev;
  int P;
  const float *shs_dev;
  int M, sh_degree;
  const float *opacities_dev, *sigma_dev;
  const float *color_dev;      /* forward colour [3][R] */
  const float *others_dev;     /* forward others [7][R] */
  const float *dL_dpix_dev;    /* [3][R] */
  const float *dL_dothers_dev; /* [7][R] */
  int k, chunk_mode;
  float cutoff, near_t, far_t;
  int zero_grads;              /* nonzero: the library zeroes the four gradient buffers first */
  int dL_dvertices_rows;       /* rows of dL_dvertices (for zeroing) */
  float *dL_dvertices_dev;     /* [rows][3] accumulated (FP32 atomics) */
  float *dL_dsigma_dev;        /* [P] */
  float *dL_dopacity_dev;      /* [P] */
  float *dL_dsh_dev;           /* [P][M][3] */
  /* optional [R][4]: the forward's interval_out (TT_CHUNK_EXACT). The re-traversal then
     covers only [first blended hit, last blended hit] of each ray: same gradients. */
  const float *interval_dev;
  /* TT_VARIANT_BOUNCE: the forward's inputs and bounce_aux record; gradients also to the
     reflectivity and the environment map (accumulated, zeroed with zero_grads). The
     secondary rays' geometry and the IOR are held fixed (no gradient through the bounce). */
  int variant;                   /* TT_VARIANT_FORWARD (0) or TT_VARIANT_BOUNCE */
  const float *reflectivity_dev, *envmap_dev;
  int env_width, env_height;
  const float *bounce_aux_dev;
  float *dL_dreflectivity_dev;   /* [P] */
  float *dL_denvmap_dev;         /* [env_height][env_width][3] */
  /* optional: the hit list and counts a tt_trace_fwd of the same rays, scene and variant wrote
     (takes precedence over interval_dev) */
  const unsigned int *hitlist_dev;
  const int *hitlist_count_dev;
  int hitlist_cap;
  const float *hitlist_color_dev; /* optional: the forward's hitlist_color_dev */
  /* nonzero: the caller guarantees that vertices_dev, cumsum_dev and sigma_dev hold the values
     they held at this tracer's last tt_trace_fwd. When that forward was given the same three
     pointers and P, the backward then reads the per-triangle window records that forward formed
     (128 B per triangle: the ray-independent terms of the window function) instead of forming
     them again; otherwise, and when zero, it forms them from its own inputs. No counterpart in
     the reference (which keeps no per-triangle state between its launches). */
  int same_scene_as_forward;
} tt_bwd_args;

/* Replaces OptixTracer::trace_bwd's launch (tracer.cpp:199-222) and backward.cu:191-391:
   re-traverses every ray with the forward's LBVH and scatters per-triangle gradients. */
int tt_trace_bwd(tt_tracer *tr, const tt_bwd_args *a, void *stream);

/* ---- batched multi-camera path (no counterpart in the reference, which traces one
   camera per call; per-pixel results equal per-camera tt_trace_fwd / tt_trace_bwd) ---- */
#define TT_MAX_CAMERAS 32

typedef struct {
  int n_cameras;                /* 1 .. TT_MAX_CAMERAS */
  const int *width, *height;    /* HOST arrays [n_cameras] */
  const float *ray_o_dev, *ray_d_dev; /* all cameras' rays, camera c first at ray sum_{j<c} W_j H_j */
  const float *vertices_dev;
  int num_vertices;
  const int *npts_dev, *cumsum_dev;
  int P;
  const float *shs_dev;
  int M, sh_degree;
  const float *opacities_dev, *sigma_dev;
  const float *bg_dev;          /* [3] */
  int k;                        /* 8, 16, 32 */
  int chunk_mode;               /* TT_CHUNK_EXACT makes results independent of k */
  int block;                    /* lanes per block: 64 (8x8 tile), 128 (16x8) or 256 (16x16) */
  int tiles_per_fetch;          /* tiles a block claims from the work queue at once (>= 1) */
  float cutoff, near_t, far_t;
  /* forward outputs: per camera [3][R_c] / [7][R_c], concatenated in camera order */
  float *color_dev, *others_dev;
  float *max_blending_dev, *density_dev, *scaling_dev; /* [P], shared by all cameras */
  int *stats_dev;               /* optional [sum R_c][2] */
  /* backward inputs (same concatenation as color/others) and gradient outputs */
  const float *dL_dpix_dev, *dL_dothers_dev;
  int zero_grads;               /* tt_multi_backward zeroes the gradient buffers first */
  int dL_dvertices_rows;
  float *dL_dvertices_dev, *dL_dsigma_dev, *dL_dopacity_dev, *dL_dsh_dev;
  /* optional stored blended hits of all cameras, as tt_fwd_args.hitlist_* with R = sum R_c:
     written by tt_multi_forward, read by tt_multi_backward instead of re-traversing */
  unsigned int *hitlist_dev;    /* [hitlist_cap][sum R_c][2] */
  int *hitlist_count_dev;       /* [sum R_c] */
  int hitlist_cap;
  /* Fused ray generation (optional): pinhole cameras in device memory, TT_CAMERA_DOUBLES doubles
     per camera -- eye[3] forward[3] right[3] up[3] tan_x tan_y, the parameters of tt_camera_rays.
     When given, every launch forms each pixel's ray as tt_camera_rays would, bit for bit, and
     ray_o_dev / ray_d_dev are not read (may be NULL). */
  const double *cameras_dev;
  /* Fused loss (optional, forward): target images laid out like color_dev. The forward writes
     the gradient of the batch's mean loss with respect to each colour into loss_grad_dev (laid
     out like color_dev: pass the dL_dpix_dev buffer to feed the backward) and the loss into
     loss_dev[0]. TT_LOSS_L2: mean of (c - t)^2, dL/dc = 2 (c - t) / N; TT_LOSS_L1: mean of
     |c - t|, dL/dc = sign(c - t) / N; N = 3 x the batch's rays. The loss is summed per tile,
     then over the tiles in a fixed order (the same value on every run). */
  const float *target_dev;
  float *loss_grad_dev;
  float *loss_dev;
  int loss_kind;                /* TT_LOSS_* */
  /* optional with hitlist_dev: the stored hits' colours, as tt_fwd_args.hitlist_color_dev
     with R = sum R_c (written by tt_multi_forward, read by tt_multi_backward) */
  float *hitlist_color_dev;     /* [hitlist_cap][sum R_c][3] */
} tt_multi_args;
#define TT_CAMERA_DOUBLES 14
#define TT_LOSS_NONE 0
#define TT_LOSS_L2 1
#define TT_LOSS_L1 2

/* Forward of every camera in ONE persistent launch (blocks pull tiles from 8 work queues). */
int tt_multi_forward(tt_tracer *tr, const tt_multi_args *a, void *stream);
/* Backward of every camera in ONE persistent launch; gradients of all cameras accumulate. */
int tt_multi_backward(tt_tracer *tr, const tt_multi_args *a, void *stream);
/* Allocation-free variant of tt_multi_forward. No counterpart in the reference; for callers
   that cannot allocate device memory (code under hipGraph stream capture, tight serving
   loops, harnesses that forbid hipMalloc). Same launches and the same outputs as
   tt_multi_forward, bit for bit; every scratch buffer (work-queue heads, tile cost estimates
   and order, traversal stack spill) lives in one static device array of
   TT_NOALLOC_SCRATCH_BYTES that each GPU the library runs on holds until the process ends;
   the call makes no hipMalloc / hipFree or other allocating call. Limits: at most
   TT_NOALLOC_MAX_TILES tiles per batch (TT_ERR_INVALID_ARG beyond, nothing launched); the
   persistent grid shrinks until its stack spill fits the array (TT_ERR_UNSUPPORTED if not one
   block fits). Every call on one GPU shares that GPU's array: calls that overlap on one GPU
   (other streams, threads or tracers) race. The tracer's own tile schedule is not written,
   so a following tt_multi_backward takes its tiles in queue order. The acceleration
   structure must have been built with tt_build on the calling (current) device. */
#define TT_NOALLOC_SCRATCH_BYTES (64u << 20)
#define TT_NOALLOC_MAX_TILES (1 << 18)
int tt_multi_forward_noalloc(tt_tracer *tr, const tt_multi_args *a, void *stream);
/* Record one training step into a hipGraph owned by the tracer: what = 1 forward, 2 backward
   (with zeroing when a->zero_grads), 3 both. Buffers must stay valid while the graph is used.
   The graph holds the tracer's state of the moment it was recorded: after tt_build(TT_BUILD)
   (a rebuild, even of the same triangles) or a call that needed larger tracer scratch (a
   bigger one-shot image, another block size), capture again. tt_build(TT_REFIT) keeps the
   graph valid (same nodes, updated boxes). */
int tt_multi_capture(tt_tracer *tr, const tt_multi_args *a, int what, void *stream);
/* Launch the captured graph on `stream`. TT_ERR_STATE (nothing launched) when the graph is
   stale as described above, or, for a backward-only graph, when a forward of another tile
   layout (any tt_trace_fwd, a tt_multi_forward of another batch) rewrote the tile schedule. */
int tt_multi_replay(tt_tracer *tr, void *stream);
/* Autotuner: times forward+backward for every (k, block, tiles_per_fetch) in the given lists
   (k is swept only in TT_CHUNK_EXACT mode, where results do not depend on it), writes the
   median ms of `reps` runs per configuration into results_ms[nk][nb][nt] (-1 = skipped),
   and stores the fastest configuration into *a. Synchronises the stream. */
int tt_multi_autotune(tt_tracer *tr, tt_multi_args *a, const int *ks, int nk, const int *blocks, int nb,
                      const int *tpf, int nt, int reps, float *results_ms, void *stream);

/* ---- learned depth-interval proposal (no counterpart in the reference) ----
   A 102-64-64-64-2 MLP (bf16 weights and activations, fp32 accumulation, matrix cores)
   mapping a ray to (start, end) of the depth range where its blended hits lie; trained
   on the GPU with Adam on the intervals tt_trace_fwd emits (interval_out), and used to
   fill tt_fwd_args.interval_in. Input encoding: (o / enc_scale, d) and sin, cos of
   2^f pi v, f = 0..7, for each of those six values (whole periods removed first, exact in
   fp32, so any ray origin is valid). */
typedef struct tt_proposal tt_proposal;

typedef struct {
  float enc_scale;  /* ray origins are divided by this before encoding (8) */
  float t_scale;    /* outputs are distances divided by this (8) */
  float t_skip;     /* target distance of rays without blended hits (16: beyond the scene) */
  float tau_start;  /* pinball-loss quantile of the start output (0.1: errs early) */
  float tau_end;    /* pinball-loss quantile of the end output (0.9: errs late) */
  float lr, beta1, beta2, eps; /* Adam (1e-3, 0.9, 0.999, 1e-8) */
} tt_proposal_hparams;

#define TT_PROPOSAL_RAW 0      /* outputs [R][2]: the network's two values */
#define TT_PROPOSAL_INTERVAL 1 /* outputs [R][2]: start = max(0, (y0 - margin_start) t_scale),
                                  cap = (y1 + margin_end) t_scale (tt_fwd_args.interval_in) */
#define TT_PROPOSAL_CAP 2      /* outputs [R][2]: start = 0, cap as above (exact with a soft cap) */

void tt_proposal_default_hparams(tt_proposal_hparams *hp);
/* Parameters: W1[64][102] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[2][64] b4[2] (fp32 master copy). */
int tt_proposal_num_params(void);
/* He-uniform weights from `seed`, zero biases, on the current device. hp NULL = defaults. */
int tt_proposal_create(tt_proposal **out, const tt_proposal_hparams *hp, unsigned long long seed);
int tt_proposal_destroy(tt_proposal *p);
/* Host copies of the fp32 parameters (synchronise the device); set resets the Adam state. */
int tt_proposal_set_params(tt_proposal *p, const float *host_params);
int tt_proposal_get_params(tt_proposal *p, float *host_params);
/* Network outputs for R rays (ray_o, ray_d [R][3]) into out [R][2] (mode TT_PROPOSAL_*). */
int tt_proposal_forward(tt_proposal *p, const float *ray_o_dev, const float *ray_d_dev, long long R,
                        float margin_start, float margin_end, int mode, float *out_dev, void *stream);
/* Gradient of the mean pinball loss over the B rays given (B % 32 == 0) into grad [n_params],
   and the loss into *loss_dev (optional). interval: [B][4] from tt_trace_fwd's interval_out. */
int tt_proposal_loss_grad(tt_proposal *p, const float *ray_o_dev, const float *ray_d_dev, const float *interval_dev,
                          int B, float *grad_dev, float *loss_dev, void *stream);
/* `steps` Adam steps, each on `batch` rays (batch % 32 == 0) drawn by hash from the n_rays given;
   loss_hist [steps] (optional) receives each step's loss. Asynchronous. */
int tt_proposal_train(tt_proposal *p, const float *ray_o_dev, const float *ray_d_dev, const float *interval_dev,
                      long long n_rays, int batch, int steps, float *loss_hist_dev, void *stream);

/* Pinhole camera rays into ray_o, ray_d [height][width][3] (the step before tt_trace_fwd; the
   reference takes these tensors from its Python caller). Pixel (x, y) looks along
   forward + xs right + ys up, xs = (2 (x + 0.5) / width - 1) tan_x, ys = (1 - 2 (y + 0.5) / height) tan_y,
   normalised in double and rounded once to float. eye/forward/right/up: double[3]. */
int tt_camera_rays(int width, int height, const double *eye, const double *forward, const double *right,
                   const double *up, double tan_x, double tan_y, float *ray_o_dev, float *ray_d_dev, void *stream);

/* Optimizer step (no counterpart in the reference, whose caller runs torch.optim.Adam on
   the gradients its autograd Function returns): Adam on up to TT_ADAM_MAX_GROUPS parameter
   groups (vertices, opacity, sigma, SH, ...) in one launch, following torch.optim.Adam's
   single-tensor update without weight decay or amsgrad:
     m = lerp(m, g, 1 - beta1); v = v beta2 + ((1 - beta2) g) g;
     p -= lr / (1 - beta1^step) * m / (sqrt(v) / sqrt(1 - beta2^step) + eps),
   with the scalars formed in double and rounded once, as torch does. zero_grads: the
   gradients are set to 0 afterwards (the next backward can accumulate without a separate
   zeroing launch). step counts from 1; m and v start at 0. After updating vertices, refit
   the acceleration structure with tt_build(..., TT_REFIT, ...). */
#define TT_ADAM_MAX_GROUPS 8
typedef struct {
  float *param_dev;       /* [n] updated in place */
  float *grad_dev;        /* [n] read (zeroed when zero_grads) */
  float *exp_avg_dev;     /* [n] first moment, in/out */
  float *exp_avg_sq_dev;  /* [n] second moment, in/out */
  long long n;
  double lr;
} tt_adam_group;
int tt_adam_step(const tt_adam_group *groups, int n_groups, int step, double beta1, double beta2, double eps,
                 int zero_grads, void *stream);

/* Training-statistics consumer (no counterpart in the reference, whose callers prune and
   densify in Python). The forward's per-triangle statistics -- max_blending (the largest
   alpha*T the triangle received, forward.cu:138-143), density_factor (1 once it was
   blended, forward.cu:134) and scaling (its largest angular radius seen from the ray
   origin, forward.cu:34-74) -- become a new triangle set: a triangle whose max_blending is
   below min_blending or whose opacity is below min_opacity is dropped; a kept triangle that
   was blended, has scaling > split_scaling and (when score_dev is given) score >
   split_score is a split candidate; candidates are split into four by midpoint
   subdivision in index order while the new count stays within max_out, the rest are kept
   whole. Survivors keep their order; the children of triangle (a, b, c) with midpoints
   ab, bc, ca are (a, ab, ca), (ab, b, bc), (ca, bc, c), (ab, bc, ca), in that order, all
   with the parent's orientation.
   tt_densify_plan writes action[i] (TT_DENSIFY_*) and offset[i], the exclusive prefix sum
   of the output triangles per input triangle, with offset[P] = the new count; read
   offset[P] back to size the outputs (the only value the host needs), then
   tt_densify_apply writes the new arrays. */
#define TT_DENSIFY_DROP 0
#define TT_DENSIFY_KEEP 1
#define TT_DENSIFY_SPLIT 2
/* scratch ints tt_densify_plan needs for P triangles */
#define TT_DENSIFY_WORK_INTS(P) (2 * (((long long)(P) + 2047) / 2048) + 2)
typedef struct {
  int P;
  const float *max_blending_dev;     /* [P] */
  const float *opacity_dev;          /* [P] */
  const float *density_dev;          /* [P] nonzero = blended, or NULL (every kept triangle counts as blended) */
  const float *scaling_dev;          /* [P], or NULL: no split candidates */
  const float *score_dev;            /* [P] caller's extra split criterion (e.g. accumulated gradient norm), or NULL */
  const unsigned char *action_in_dev; /* [P] caller's own decision (TT_DENSIFY_*), or NULL; replaces the
                                         statistics rules, max_out still applies to its splits */
  float min_blending;
  float min_opacity;
  float split_scaling;
  float split_score;
  long long max_out;                 /* cap on the new count (<= 0: 2^31 - 1); kept triangles are never dropped to
                                        meet it, only splits are withheld */
  unsigned char *action_dev;         /* out [P] (TT_DENSIFY_SPLIT only where the split was granted); NULL if P = 0 */
  int *offset_dev;                   /* out [P + 1] */
  int *work_dev;                     /* scratch, TT_DENSIFY_WORK_INTS(P) ints */
} tt_densify_plan_args;
int tt_densify_plan(const tt_densify_plan_args *args, void *stream);

/* One array remapped by tt_densify_apply. Rows of the input are per triangle, either
   TT_ROWS_TRIANGLE (row i = width floats at i * width: opacity, sigma, SH [P][M][3] with
   width 3M, optimizer moments of these) or TT_ROWS_VERTEX (row i = the 9 floats of the
   triangle's three points at 3 * cumsum[i]: vertices and their moments; width must be 9).
   The output is always dense: row j at j * width (for vertex rows: npts = 3, cumsum = 3j).
   A kept triangle's row is copied; a split triangle's four children get, by on_split:
   TT_SPLIT_COPY the parent's row, TT_SPLIT_ZERO zeros (optimizer moments of new
   triangles), TT_SPLIT_SUBDIVIDE (vertex rows only) the midpoint-subdivided points. */
#define TT_DENSIFY_MAX_ARRAYS 16
#define TT_ROWS_TRIANGLE 0
#define TT_ROWS_VERTEX 1
#define TT_SPLIT_COPY 0
#define TT_SPLIT_ZERO 1
#define TT_SPLIT_SUBDIVIDE 2
typedef struct {
  const float *src_dev;
  float *dst_dev;  /* [P_out * width]; must not overlap src */
  int width;
  int rows;        /* TT_ROWS_* */
  int on_split;    /* TT_SPLIT_* */
} tt_densify_array;
typedef struct {
  int P;
  int P_out;                 /* offset[P] of the plan */
  const int *offset_dev;     /* [P + 1] from tt_densify_plan (output counts 0, 1 or 4 per triangle) */
  const int *cumsum_dev;     /* [P], needed by TT_ROWS_VERTEX arrays */
  int *npts_out_dev;         /* [P_out] set to 3, or NULL */
  int *cumsum_out_dev;       /* [P_out] set to 3j, or NULL */
  int n_arrays;
  tt_densify_array arrays[TT_DENSIFY_MAX_ARRAYS];
} tt_densify_apply_args;
int tt_densify_apply(const tt_densify_apply_args *args, void *stream);

/* Scene I/O (no counterpart in the reference, which neither loads nor saves scenes): the
   byte layout of a binary little-endian PLY scene file on the device. The file holds
     element vertex V    (x, y, z float, other properties allowed)
     element face P      (property list uchar int vertex_indices, float opacity, float sigma,
                          float f_dc_0..2, float f_rest_0..3(M-1)-1, other properties allowed)
   with SH coefficient j >= 1 of colour channel c at f_rest_{c (M-1) + j - 1} (channel-major,
   the usual splat-PLY order) and f_dc_c the DC term of channel c. The host parses the
   header and hands over the two element bodies as separate device buffers.
   tt_scene_unpack turns them into the tracer's layout: vertices [3P][3] gathered through
   the indices (npts = 3, cumsum = 3i), opacity [P], sigma [P], shs [P][M][3]; absent
   properties (offset -1) read as 0. bad_dev (one int, accumulated) counts faces whose
   list count is not 3 or whose index lies outside [0, V); their vertices are written as 0.
   tt_scene_pack writes the canonical rows this build saves: face rows of
   TT_SCENE_FACE_BYTES(M) bytes (count 3, indices 3i..3i+2, opacity, sigma, f_dc, f_rest
   in that order) and vertex rows x, y, z [3P][3] read through cumsum. */
#define TT_SCENE_MAX_COEFFS 16
#define TT_SCENE_FACE_BYTES(M) (1 + 12 + 4 * (2 + 3 * (M)))
typedef struct {
  int P;
  long long V;
  const unsigned char *face_dev;  /* P rows of face_stride bytes, 16-byte aligned */
  int face_stride;
  int count_offset;               /* uchar list count */
  int index_offset;               /* three int32 indices (little endian) */
  int opacity_offset, sigma_offset;
  int sh_offset[TT_SCENE_MAX_COEFFS * 3]; /* byte offset of coefficient j, channel c at [3 j + c], or -1 */
  int M;                          /* coefficients per channel written to shs (1..16) */
  const unsigned char *vert_dev;  /* V rows of vert_stride bytes */
  int vert_stride;
  int xyz_offset[3];
  float *vertices_out_dev;        /* [3P][3] */
  int *npts_out_dev, *cumsum_out_dev; /* [P] or NULL */
  float *opacity_out_dev, *sigma_out_dev; /* [P] */
  float *shs_out_dev;             /* [P][M][3] */
  int *bad_dev;                   /* [1] incremented, or NULL */
} tt_scene_unpack_args;
int tt_scene_unpack(const tt_scene_unpack_args *args, void *stream);
typedef struct {
  int P, M;
  const float *vertices_dev;      /* triangle i at rows cumsum[i] .. cumsum[i] + 2 */
  const int *cumsum_dev;
  const float *opacity_dev, *sigma_dev, *shs_dev; /* [P], [P], [P][M][3] */
  unsigned char *face_out_dev;    /* P * TT_SCENE_FACE_BYTES(M) bytes, 16-byte aligned */
  float *vert_out_dev;            /* [3P][3] */
} tt_scene_pack_args;
int tt_scene_pack(const tt_scene_pack_args *args, void *stream);

/* Per-kernel timing (the counterpart of the reference's debug timing of its launch,
   tracer.cpp:110-137, at kernel granularity): while enabled, every tt_trace_fwd / tt_trace_bwd
   records HIP events on its stream right before and after its traversal kernel(s) -- not the
   tile-schedule kernels around them -- into a ring of TT_TIMING_RING launches per direction.
   Enabling (or disabling) clears the ring. */
#define TT_TIMING_RING 64
int tt_set_kernel_timing(tt_tracer *tr, int enable);
/* Durations in ms of the last min(max_n, TT_TIMING_RING, launches timed) launches of the forward
   (which = 0) or backward (which = 1) traversal kernels, oldest first; *count_out = how many were
   written. Waits for those launches to finish. */
int tt_kernel_times(tt_tracer *tr, int which, float *ms_out, int max_n, int *count_out);

/* Rebuild policy (no counterpart in the reference, whose caller picks build / refit / reuse
   per call, diff_tritracer/__init__.py:42). Quality measure: the SAH cost of the current nodes
   with unit traversal and intersection costs, 1 + the sum over every non-root node of
   SA(node) / SA(root). tt_build(TT_REFIT_AUTO) refits, evaluates that cost, and rebuilds when
   it exceeds max_ratio x the cost right after the last TT_BUILD (default 1.25; <= 0: never);
   it reads the cost back, so it waits for the refit. */
int tt_set_rebuild_policy(tt_tracer *tr, float max_ratio);
/* The SAH cost of the current acceleration structure into cost_dev[0] on `stream`. */
int tt_bvh_cost(tt_tracer *tr, float *cost_dev, void *stream);
/* counts = builds (incl. automatic), refits (incl. TT_REFIT_AUTO), automatic rebuilds since
   tt_create; costs = SAH cost right after the last build and the last evaluated one (-1 if
   none). Waits for the device. */
int tt_build_stats(tt_tracer *tr, int counts[3], float costs[2]);

/* Number of triangles of the current acceleration structure (-1 if none). */
int tt_num_primitives(const tt_tracer *tr);

/* Timing of the last tt_build on its stream in ms (hipEvents), or -1. */
float tt_last_build_ms(const tt_tracer *tr);

#ifdef __cplusplus
}
#endif

#endif /* TRITRACER_HIP_H */
