#!/usr/bin/env python
"""Benchmark of the HIP triangle-splat tracer on one MI355X.

Default workload (the headline): fwd+bwd Mrays/s at 1920x1080 on 1M clustered
triangle splats (64 clusters, seed 1), k = 16 hits per chunk, transmittance
cutoff 1e-3, random N(0,1) image-space loss gradient; exact chunk mode.

  plain        : forward launch (also storing each pixel's blended hits, up to
                 --hitlist-cap) + gradient zeroing + backward launch reading those
                 hits instead of re-traversing (same gradients)
  plain_retrace: the same with the reference's re-traversing backward
  proposal     : the learned depth-interval proposal (bf16 MLP on the matrix
                 cores, trained first on the intervals the tracer emits for 16
                 other camera poses of the same scene; the measured frame is held
                 out) -> forward capped at the proposed end with exact continuation
                 past it (same image as plain, bit for bit) -> backward from the
                 stored hits (same gradients); the network pass that proposes the
                 next step's intervals runs on a second stream beside the backward
  proposal_hard: lossy variant in config: proposed start and a hard cap, with
                 fallbacks to the full trace; its PSNR against plain is reported
  proposal_ceiling: each ray's own emitted interval of this frame as (start, cap), trusted
                 (hard cap, no fallback; rays that blend nothing are not traced) -- what a
                 perfect proposal could save (same image as plain)
  value        : Mrays/s of the proposal step; plain and the ablations are in config

One step = one pass over the frame with inputs resident in HBM; times are medians
over the timed steps. The LBVH build, refit and the proposal's training are timed
separately and reported beside it.

Prints one JSON line. Extra objects:
  roofline      : the dominant traversal kernel's algorithmic bytes per launch (per ray
                  64 + 48 Hc + 204 Hb forward, 104 + 48 Hc + 428 Hb backward) / its median
                  launch time (the library's HIP events right around that kernel on the
                  launch stream) against 8 TB/s
  cpu_baseline  : the scalar C++/OpenMP oracle (same algorithm, FP32) on the frame,
                  timed on the host cores
Other workloads: --workload multiview (2M splats, 8 uneven cameras, one persistent
hipGraph launch, autotuned), --workload proposal (3M splats, 128 clusters: the
learned-proposal configuration), --workload bounce (secondary rays and environment
light), --workload train (fwd + bwd + fused Adam + LBVH refit), --workload densify
(pruning + densification from the forward's statistics), --workload sceneio (scene-file rows
<-> tracer layout on the GPU).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "diff-triangle-tracer_amd"))

HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0  # MI355X dense bf16 MFMA
# Counters of the default workload's traversal kernels (1M splats, 1920x1080, k=16), copied from
# the rocprofv3 runs summarised in PROFILE_SOURCE (they cannot be read inside this process):
# traffic = HBM bytes per launch, 2.0 x FETCH_SIZE + 1.0 x WRITE_SIZE (factors from a 256 MiB
# copy); avg_ms = the profiler's median duration over the timed steps; l2_hit = TCC_HIT / (TCC_HIT + TCC_MISS);
# valu_busy = SQ_ACTIVE_INST_VALU x 4 / (SIMDs x GRBM_GUI_ACTIVE / XCDs).
PROFILE_SOURCE = "profiles/trace_fwd_bwd_1M_1080p.md"
MEASURED = {
    "k_forward<16,0,false,1>": dict(traffic=6.166e9, avg_ms=4.144, l2_hit=0.691, valu_busy=0.68),
    "k_forward<16,0,false,0>": dict(traffic=6.197e9, avg_ms=4.161, l2_hit=0.691, valu_busy=0.68),
    "k_backward_rec<16,0>": dict(traffic=4.692e9, avg_ms=1.788, l2_hit=0.632, valu_busy=0.49),
}
# densify and scene-file kernels at 1M splats (profiles/densify_sceneio_1M.md)
MEASURED_TRAFFIC = {"k_apply": 1.058e9, "k_scene_unpack": 0.493e9}
TRAIN_POSES, HELD_OUT_EXTRA = 16, 3


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or None."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(scene, ray_o, ray_d, stride, cutoff, k):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O

    ro = np.ascontiguousarray(ray_o[::stride, ::stride]).reshape(-1, 3)
    rd = np.ascontiguousarray(ray_d[::stride, ::stride]).reshape(-1, 3)
    R = ro.shape[0]
    t0 = time.perf_counter()
    S = O.Scene(scene.verts, scene.npts, scene.cumsum, scene.shs, scene.sh_degree, scene.opac, scene.sigma)
    t_build = time.perf_counter() - t0
    opts = O.make_opts(k=k, cutoff=cutoff)
    rng = np.random.Generator(np.random.PCG64(11))
    dpix = rng.standard_normal((3, R)).astype(np.float32)
    doth = np.zeros((7, R), np.float32)
    t0 = time.perf_counter()
    f = O.forward_f32(S, ro, rd, opts)
    O.backward_f32(S, ro, rd, opts, f["color"], f["others"], dpix, doth)
    t = time.perf_counter() - t0
    return dict(value=round(R / t / 1e6, 5), unit="Mrays/s", cores=O.get_threads(), kind="port", cpu_model=cpu_model(),
                sample=f"every {stride}th pixel in x and y of the same 1920x1080 frame ({R} rays), full 1M-splat "
                       f"scene, FP32 fwd+bwd of the scalar C++/OpenMP oracle; its own BVH build "
                       f"{t_build:.2f} s excluded", seconds=round(t, 3))




def multiview(args, dev):
    """Batched multi-view step: 2M clustered splats, 8 cameras of uneven resolution
    (9,004,032 rays) rendered fwd+bwd by the persistent tile-queue kernels captured in
    one hipGraph; k x block x tiles-per-fetch chosen by the autotuner (exact chunk
    mode, so every configuration renders the same image)."""
    import numpy as np
    import torch
    from diff_tritracer import capi, scenes

    t0 = time.perf_counter()
    P = args.triangles
    scene, ws, hs, ray_o, ray_d = scenes.multiview_config(seed=2, P=P)
    Rt = ray_o.shape[0]
    print(f"[bench] multiview: {P} splats, {len(ws)} cameras, {Rt} rays, generated in "
          f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    verts, npts, cumsum = T(scene.verts), T(scene.npts), T(scene.cumsum)
    shs, opac, sigma = T(scene.shs), T(scene.opac.reshape(-1)), T(scene.sigma.reshape(-1))
    ro, rd = T(ray_o), T(ray_d)
    bg = torch.zeros(3, device=dev)
    color, others = torch.zeros(3 * Rt, device=dev), torch.zeros(7 * Rt, device=dev)
    maxb, dens, scal = (torch.zeros(P, device=dev) for _ in range(3))
    g = torch.Generator(device="cpu").manual_seed(5)
    dpix = torch.randn(3 * Rt, generator=g).to(dev)
    doth = torch.zeros(7 * Rt, device=dev)
    dverts = torch.zeros(verts.shape[0], 3, device=dev)
    dsig, dop = torch.zeros(P, device=dev), torch.zeros(P, device=dev)
    dsh = torch.zeros(P, shs.shape[1], 3, device=dev)
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    h = capi.TracerHandle()
    h.build(verts, cumsum, P, capi.BUILD, sp)
    cap = args.hitlist_cap
    hl = capi.hitlist_buffers(Rt, cap, dev) if cap > 0 else None
    a = capi.multi_args(ws, hs, ro, rd, verts, npts, cumsum, shs, 3, opac, sigma, bg, color, others, maxb, dens, scal,
                        dL_dpix=dpix, dL_dothers=doth, dverts=dverts, dsigma=dsig, dopac=dop, dsh=dsh, k=16,
                        cutoff=args.cutoff, chunk_mode="exact", block=128, tiles_per_fetch=1, zero_grads=True,
                        hitlist=hl)
    t0 = time.perf_counter()
    if args.no_autotune:
        a.k, a.block, a.tiles_per_fetch = 16, 256, 1
        sweep = {}
    else:
        sweep = h.multi_autotune(a, ks=(8, 16, 32), blocks=(64, 128, 256), tpf=(1, 2, 4), reps=3, stream=sp)
    print(f"[bench] autotuner: {len(sweep)} configurations in {time.perf_counter() - t0:.1f} s; best k={a.k} "
          f"block={a.block} tiles/fetch={a.tiles_per_fetch}", file=sys.stderr, flush=True)
    for key in sorted(sweep, key=sweep.get):
        print(f"[bench]   k={key[0]:2d} block={key[1]:3d} tpf={key[2]}  {sweep[key]:.3f} ms", file=sys.stderr)
    h.multi_capture(a, 3, sp)
    for _ in range(args.warmup):
        h.multi_replay(sp)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(args.steps):
        h.multi_replay(sp)
    stop.record(stream)
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop) / args.steps
    # the fused step: rays formed in the launches from the camera table (no ray buffers) and the
    # loss gradient (mean L2 against a target image) written by the forward for the backward
    cams = capi.camera_table(scenes.multiview_bases(scenes.MULTIVIEW_SIZES, seed=2), dev)
    target = torch.rand(3 * Rt, generator=torch.Generator(device="cpu").manual_seed(6)).to(dev)
    dpix_f, loss = torch.zeros(3 * Rt, device=dev), torch.zeros(1, device=dev)
    af = capi.multi_args(ws, hs, None, None, verts, npts, cumsum, shs, 3, opac, sigma, bg, color, others, maxb, dens,
                         scal, dL_dpix=dpix_f, dL_dothers=doth, dverts=dverts, dsigma=dsig, dopac=dop, dsh=dsh, k=a.k,
                         cutoff=args.cutoff, chunk_mode="exact", block=a.block, tiles_per_fetch=a.tiles_per_fetch,
                         zero_grads=True, hitlist=hl, cameras=cams, target=target, loss_grad=dpix_f, loss=loss)
    h.multi_capture(af, 3, sp)
    for _ in range(args.warmup):
        h.multi_replay(sp)
    torch.cuda.synchronize()
    start.record(stream)
    for _ in range(args.steps):
        h.multi_replay(sp)
    stop.record(stream)
    torch.cuda.synchronize()
    ms_fused = start.elapsed_time(stop) / args.steps
    print(f"[bench] fused step (rays + loss gradient in the launches) {ms_fused:.3f} ms/step, loss "
          f"{float(loss.item()):.5f}", file=sys.stderr, flush=True)
    # the same batch as one forward + one backward launch per camera (same k, exact mode)
    launches, off, hc_keep = [], 0, []
    for wc, hc in zip(ws, hs):
        R = wc * hc
        sl = slice(off, off + R)
        col, oth = color[3 * off:3 * (off + R)], others[7 * off:7 * (off + R)]
        hc_ = capi.hitlist_buffers(R, cap, dev) if cap > 0 else None
        fa = capi.fwd_args(wc, hc, ro[sl], rd[sl], verts, npts, cumsum, shs, 3, opac, sigma, bg, col, oth, maxb, dens,
                           scal, k=a.k, cutoff=args.cutoff, chunk_mode="exact", hitlist=hc_)
        ba = capi.bwd_args(wc, hc, ro[sl], rd[sl], verts, npts, cumsum, shs, 3, opac, sigma, col, oth,
                           dpix[3 * off:3 * (off + R)], doth[7 * off:7 * (off + R)], dverts, dsig, dop, dsh, k=a.k,
                           cutoff=args.cutoff, chunk_mode="exact", zero_grads=False, hitlist=hc_)
        hc_keep.append(hc_)
        launches.append((fa, ba))
        off += R

    def per_camera():  # each camera's backward right after its forward (its schedule and record regions)
        for t in (dverts, dsig, dop, dsh):
            t.zero_()
        for fa, ba in launches:
            h.fwd(fa, sp)
            h.bwd(ba, sp)

    for _ in range(args.warmup):
        per_camera()
    torch.cuda.synchronize()
    start.record(stream)
    for _ in range(args.steps):
        per_camera()
    stop.record(stream)
    torch.cuda.synchronize()
    ms_pc = start.elapsed_time(stop) / args.steps
    print(f"[bench] persistent graph {ms:.3f} ms/step, per-camera launches {ms_pc:.3f} ms/step", file=sys.stderr,
          flush=True)
    line = {
        "metric": "multi-view fwd+bwd Mrays/s (one persistent hipGraph launch)",
        "value": round(Rt / (ms * 1e-3) / 1e6, 4), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded clustered triangle splats, 8 pinhole cameras on random poses at radius 3-5)",
        "config": {"workload": f"multi-view fwd+bwd, {P} clustered splats (64 clusters, seed 2), cameras "
                               + ", ".join(f"{w}x{hh}" for w, hh in zip(ws, hs)),
                   "rays": Rt, "k": a.k, "block": a.block, "tiles_per_fetch": a.tiles_per_fetch, "hitlist_cap": cap,
                   "chunk_mode": "exact", "cutoff": args.cutoff,
                   "fused_step_ms": round(ms_fused, 4),
                   "fused_step_mrays": round(Rt / (ms_fused * 1e-3) / 1e6, 4),
                   "per_camera_launches_ms": round(ms_pc, 4),
                   "per_camera_launches_mrays": round(Rt / (ms_pc * 1e-3) / 1e6, 4),
                   "autotune_ms": {f"k{k_}_b{b}_t{t}": round(v, 3) for (k_, b, t), v in sorted(sweep.items())}},
    }
    print(json.dumps(line), flush=True)


def bounce(args, dev):
    """Secondary-ray variant: 1M clustered splats with per-triangle reflectivity ~ Beta(2, 5)
    and IOR ~ U[1, 1.6], one reflected and one refracted ray per pixel, equirect environment
    1024x512x3 with lognormal radiance, 1920x1080 camera; fwd+bwd with gradients to the
    triangles, the reflectivity and the texels."""
    import numpy as np
    import torch
    from diff_tritracer import capi, scenes

    W, H = args.width, args.height
    scene = scenes.clustered_scene(args.triangles, 64, seed=3)
    rng = np.random.Generator(np.random.PCG64(3))
    refl = rng.beta(2.0, 5.0, scene.P).astype(np.float32)
    ior = rng.uniform(1.0, 1.6, scene.P).astype(np.float32)
    env = np.exp(rng.normal(0.0, 0.5, (512, 1024, 3))).astype(np.float32)
    fr = Frame(dev, scene, args.k, args.cutoff)
    fr.h.build(fr.verts, fr.cumsum, fr.P, capi.BUILD, fr.sp)
    fr.buffers(W, H)
    eye = 4.0 * np.array([0.35, 0.3, 1.0]) / np.linalg.norm([0.35, 0.3, 1.0])
    fr.camera(scenes.camera_basis(W, H, eye))
    T = lambda a: torch.from_numpy(a).to(dev)
    B = dict(reflectivity=T(refl), ior=T(ior), envmap=T(env), aux=torch.zeros(fr.R, 20, device=dev),
             dreflectivity=torch.zeros(fr.P, device=dev), denvmap=torch.zeros(512, 1024, 3, device=dev))
    stats = torch.zeros(fr.R, 2, dtype=torch.int32, device=dev)
    c = capi
    cap = args.hitlist_cap
    hl = c.hitlist_buffers(fr.R, cap, dev, segments=3) if cap > 0 else None
    fa = c.fwd_args(W, H, fr.ro, fr.rd, fr.verts, fr.npts, fr.cumsum, fr.shs, 3, fr.opac, fr.sigma, fr.bg, fr.color,
                    fr.others, fr.maxb, fr.dens, fr.scal, k=args.k, cutoff=args.cutoff, chunk_mode="exact",
                    variant="bounce", bounce=B, hitlist=hl)
    ba = c.bwd_args(W, H, fr.ro, fr.rd, fr.verts, fr.npts, fr.cumsum, fr.shs, 3, fr.opac, fr.sigma, fr.color,
                    fr.others, fr.dpix, fr.doth, fr.dverts, fr.dsig, fr.dop, fr.dsh, k=args.k, cutoff=args.cutoff,
                    chunk_mode="exact", zero_grads=True, bounce=B, hitlist=hl)
    fs = c.fwd_args(W, H, fr.ro, fr.rd, fr.verts, fr.npts, fr.cumsum, fr.shs, 3, fr.opac, fr.sigma, fr.bg, fr.color,
                    fr.others, fr.maxb, fr.dens, fr.scal, k=args.k, cutoff=args.cutoff, chunk_mode="exact",
                    variant="bounce", bounce=B, stats=stats)
    fr.h.fwd(fs, fr.sp)
    torch.cuda.synchronize()
    st = stats.to(torch.int64).sum(0).cpu().numpy()
    sp = fr.sp
    ms, per = timed(fr, [("fwd", lambda: fr.h.fwd(fa, sp)), ("bwd", lambda: fr.h.bwd(ba, sp))], args.steps,
                    args.warmup)
    R = fr.R
    print(f"[bench] bounce {ms:.3f} ms/step " + " ".join(f"{k} {v:.3f}" for k, v in per.items()), file=sys.stderr,
          flush=True)
    line = {
        "metric": "bounce variant fwd+bwd Mrays/s (primary + reflected + refracted ray per pixel)",
        "value": round(R / (ms * 1e-3) / 1e6, 4), "unit": "Mrays/s (primary rays)", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (clustered splats, Beta(2,5) reflectivity, U[1,1.6] IOR, lognormal 1024x512 environment)",
        "config": {"workload": f"bounce variant fwd+bwd, {scene.P} clustered splats (64 clusters, seed 3), one "
                               f"{W}x{H} camera, k={args.k}, cutoff {args.cutoff:g}, exact chunk mode",
                   "fwd_ms": round(per["fwd"], 4), "bwd_ms": round(per["bwd"], 4), "hitlist_cap": cap,
                   "hits_consumed_per_pixel": round(float(st[0]) / R, 3),
                   "hits_blended_per_pixel": round(float(st[1]) / R, 3)},
    }
    print(json.dumps(line), flush=True)


def train(args, dev):
    """Training step on the default frame: forward (storing the hit list) + backward reading it
    + the fused Adam step on vertices, opacity, sigma and SH (zeroing the gradients it
    consumed) + LBVH refit to the moved vertices. value = rays per second of the whole step."""
    import numpy as np
    import torch
    from diff_tritracer import capi, scenes

    W, H = args.width, args.height
    scene = scenes.clustered_scene(args.triangles, 64, seed=1)
    fr = Frame(dev, scene, args.k, args.cutoff, args.hitlist_cap)
    fr.h.build(fr.verts, fr.cumsum, fr.P, capi.BUILD, fr.sp)
    fr.buffers(W, H)
    eye = 4.0 * np.array([0.35, 0.3, 1.0]) / np.linalg.norm([0.35, 0.3, 1.0])
    fr.camera(scenes.camera_basis(W, H, eye))
    fr.dpix.mul_(1e-3)  # a loss gradient small enough that 2000 steps keep the scene sane
    params = [(fr.verts, fr.dverts, 1e-4), (fr.opac, fr.dop, 1e-2), (fr.sigma, fr.dsig, 1e-2), (fr.shs, fr.dsh, 2.5e-3)]
    moments = [(torch.zeros_like(p), torch.zeros_like(p)) for p, _, _ in params]
    groups = [(p, g, m, v, lr) for (p, g, lr), (m, v) in zip(params, moments)]
    fa, ba = fr.fwd_args("exact"), fr.bwd_args("exact")
    sp = fr.sp
    state = {"step": 0}
    fr.zero_grads()

    def adam():
        state["step"] += 1
        capi.adam_step(groups, state["step"], zero_grads=True, stream=sp)

    launches = [("fwd", lambda: fr.h.fwd(fa, sp)), ("bwd", lambda: fr.h.bwd(ba, sp)), ("adam", adam),
                ("refit", lambda: fr.h.build(fr.verts, fr.cumsum, fr.P, capi.REFIT, sp))]
    ms, per = timed(fr, launches, args.steps, args.warmup)
    R = fr.R
    n_params = sum(int(p.numel()) for p, _, _ in params)
    print(f"[bench] train {ms:.3f} ms/step " + " ".join(f"{k} {v:.3f}" for k, v in per.items()) +
          f"  ({R / ms / 1e3:.2f} Mrays/s)", file=sys.stderr, flush=True)
    line = {
        "metric": "training step Mrays/s (fwd + bwd + Adam + LBVH refit)",
        "value": round(R / (ms * 1e-3) / 1e6, 4), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic (seeded clustered triangle splats, pinhole camera)",
        "config": {"workload": f"training step, {scene.P} clustered splats (64 clusters, seed 1), one {W}x{H} "
                               f"camera, k={args.k}, cutoff {args.cutoff:g}, exact chunk mode, hit list "
                               f"{args.hitlist_cap}, Adam on {n_params} parameters (4 groups), refit",
                   "kernel_ms": {k: round(v, 4) for k, v in per.items()},
                   # Adam moves 32 B per parameter: p, g, m, v read; p, m, v and the zeroed g written
                   "adam_gbs": round(8 * 4 * n_params / (per["adam"] * 1e-3) / 1e9, 1)},
    }
    print(json.dumps(line), flush=True)


def refit_drift(args, dev):
    """LBVH quality over a training run: the train workload's step (forward + backward + Adam
    on vertices, opacity, sigma and SH + tt_build(TT_REFIT_AUTO)) for --train-steps steps,
    once with the rebuild policy off (refit only) and once at --rebuild-ratio. Every 100 steps:
    the median forward and backward traversal-kernel time of the last 10 steps and the SAH
    cost (1 + sum of non-root node areas / root area). value = fwd+bwd kernel Mrays/s of the
    policy run's last window."""
    import numpy as np
    import torch
    from diff_tritracer import capi, scenes

    W, H = args.width, args.height
    eye = 4.0 * np.array([0.35, 0.3, 1.0]) / np.linalg.norm([0.35, 0.3, 1.0])
    runs = {}
    for label, ratio in (("refit_only", 0.0), ("policy", args.rebuild_ratio), ("rebuild_every_step", 1e-6)):
        scene = scenes.clustered_scene(args.triangles, 64, seed=1)
        fr = Frame(dev, scene, args.k, args.cutoff, args.hitlist_cap)
        fr.h.build(fr.verts, fr.cumsum, fr.P, capi.BUILD, fr.sp)
        fr.buffers(W, H)
        fr.camera(scenes.camera_basis(W, H, eye))
        fr.dpix.mul_(1e-3)
        params = [(fr.verts, fr.dverts, 1e-4), (fr.opac, fr.dop, 1e-2), (fr.sigma, fr.dsig, 1e-2),
                  (fr.shs, fr.dsh, 2.5e-3)]
        groups = [(p, g, torch.zeros_like(p), torch.zeros_like(p), lr) for p, g, lr in params]
        fa, ba = fr.fwd_args("exact"), fr.bwd_args("exact")
        fr.h.set_rebuild_policy(ratio)
        fr.zero_grads()
        fresh, fresh_cost = capi.TracerHandle(), torch.zeros(1, device=dev)  # what a rebuild would give
        series = []
        t0 = time.perf_counter()
        for step in range(1, args.train_steps + 1):
            if step % 100 == 91:
                fr.h.set_kernel_timing(True)
            fr.h.fwd(fa, fr.sp)
            fr.h.bwd(ba, fr.sp)
            capi.adam_step(groups, step, zero_grads=True, stream=fr.sp)
            fr.h.build(fr.verts, fr.cumsum, fr.P, capi.REFIT_AUTO, fr.sp)
            if step % 100 == 0:
                kf, kb = fr.h.kernel_times("fwd", 10), fr.h.kernel_times("bwd", 10)
                fr.h.set_kernel_timing(False)
                st = fr.h.build_stats()
                fresh.build(fr.verts, fr.cumsum, fr.P, capi.BUILD, fr.sp)
                fresh.bvh_cost(fresh_cost, fr.sp)
                series.append(dict(step=step, fwd_ms=round(_median(kf), 4), bwd_ms=round(_median(kb), 4),
                                   sah=round(st[4], 4), sah_built=round(st[3], 4),
                                   sah_fresh_build=round(float(fresh_cost.item()), 4), auto_rebuilds=st[2]))
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        runs[label] = dict(series=series, wall_s=round(wall, 2), rebuilds=fr.h.build_stats()[2])
        print(f"[bench] {label}: " + " ".join(f"{x['step']}:{x['fwd_ms']:.2f}/{x['sah']:.1f}/{x['sah_fresh_build']:.1f}"
                                            for x in series) +
              f"  rebuilds {runs[label]['rebuilds']} wall {wall:.1f} s", file=sys.stderr, flush=True)
        del fr
        torch.cuda.empty_cache()
    last = runs["policy"]["series"][-1]
    R = W * H
    line = {
        "metric": "traversal Mrays/s after Adam + refit steps (automatic rebuild policy)",
        "value": round(R / ((last["fwd_ms"] + last["bwd_ms"]) * 1e-3) / 1e6, 4), "unit": "Mrays/s", "n_gpus": 1,
        "steps": args.train_steps, "warmup": 0, "ms_per_step": round(last["fwd_ms"] + last["bwd_ms"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded clustered triangle splats, pinhole camera, fixed N(0,1e-3) loss gradient)",
        "config": {"workload": f"{args.train_steps} training steps (fwd + bwd + Adam + TT_REFIT_AUTO) on "
                               f"{args.triangles} clustered splats, one {W}x{H} camera, refit only vs rebuild at "
                               f"SAH ratio {args.rebuild_ratio}", **runs},
    }
    print(json.dumps(line), flush=True)


def densify_bench(args, dev):
    """Pruning + densification of a trained-size scene from real forward statistics: the
    1080p exact forward fills max_blending / density / scaling, then each step runs
    tt_densify_plan (3 launches) and tt_densify_apply on the whole parameter set with Adam
    moments (vertices + 2 moments as vertex rows, opacity, sigma, SH + 2 moments each: 12
    arrays, 708 B per triangle). value = input triangles per second of plan + apply."""
    import numpy as np
    import torch
    from diff_tritracer import capi, scenes

    W, H = args.width, args.height
    scene = scenes.clustered_scene(args.triangles, 64, seed=1)
    fr = Frame(dev, scene, args.k, args.cutoff)
    fr.h.build(fr.verts, fr.cumsum, fr.P, capi.BUILD, fr.sp)
    fr.buffers(W, H)
    eye = 4.0 * np.array([0.35, 0.3, 1.0]) / np.linalg.norm([0.35, 0.3, 1.0])
    fr.camera(scenes.camera_basis(W, H, eye))
    fr.h.fwd(fr.fwd_args("exact", hitlist=False), fr.sp)
    torch.cuda.synchronize()
    P, sp = fr.P, fr.sp
    # thresholds: drop triangles below 1/255 peak blending weight, split the blended
    # triangles in the top 5% of angular size, cap the growth at 10%
    blended = fr.scal[fr.dens > 0]
    split_scaling = float(torch.quantile(blended[:1 << 24].float(), 0.95)) if blended.numel() else float("inf")
    kw = dict(min_blending=1.0 / 255.0, min_opacity=0.005, split_scaling=split_scaling, max_out=int(1.1 * P))
    work = torch.empty(capi.densify_work_ints(P), dtype=torch.int32, device=dev)
    action = torch.empty(P, dtype=torch.uint8, device=dev)
    offset = torch.empty(P + 1, dtype=torch.int32, device=dev)
    pa = capi.DensifyPlanArgs()
    pa.P = P
    pa.max_blending_dev, pa.opacity_dev, pa.density_dev, pa.scaling_dev = (capi.ptr(t) for t in
                                                                           (fr.maxb, fr.opac, fr.dens, fr.scal))
    pa.min_blending, pa.min_opacity, pa.split_scaling = kw["min_blending"], kw["min_opacity"], kw["split_scaling"]
    pa.max_out = kw["max_out"]
    pa.action_dev, pa.offset_dev, pa.work_dev = capi.ptr(action), capi.ptr(offset), capi.ptr(work)
    L = capi.lib()
    capi.check(L.tt_densify_plan(capi.C.byref(pa), sp), "tt_densify_plan")
    P_out = int(offset[-1].item())
    acts = torch.bincount(action.long(), minlength=3).cpu().tolist()
    mom = lambda t: (torch.zeros_like(t).normal_(), torch.zeros_like(t).normal_())
    (vm, vv), (om, ov), (sm, sv), (hm, hv) = mom(fr.verts), mom(fr.opac), mom(fr.sigma), mom(fr.shs)
    srcs = [(fr.verts, 9, capi.ROWS_VERTEX, capi.SPLIT_SUBDIVIDE), (vm, 9, capi.ROWS_VERTEX, capi.SPLIT_ZERO),
            (vv, 9, capi.ROWS_VERTEX, capi.SPLIT_ZERO)]
    for t, m, v, w in ((fr.opac, om, ov, 1), (fr.sigma, sm, sv, 1), (fr.shs, hm, hv, 48)):
        srcs += [(t, w, capi.ROWS_TRIANGLE, capi.SPLIT_COPY), (m, w, capi.ROWS_TRIANGLE, capi.SPLIT_ZERO),
                 (v, w, capi.ROWS_TRIANGLE, capi.SPLIT_ZERO)]
    outs = [torch.empty(P_out * w, device=dev) for _, w, _, _ in srcs]
    npts_o = torch.empty(P_out, dtype=torch.int32, device=dev)
    cs_o = torch.empty(P_out, dtype=torch.int32, device=dev)
    aa = capi.densify_apply_args(P, P_out, offset, fr.cumsum, [(s, o, w, r, p) for (s, w, r, p), o in
                                                              zip(srcs, outs)], npts_o, cs_o)

    def plan():
        capi.check(L.tt_densify_plan(capi.C.byref(pa), sp), "tt_densify_plan")

    def apply():
        capi.check(L.tt_densify_apply(capi.C.byref(aa), sp), "tt_densify_apply")

    ms, per = timed(fr, [("plan", plan), ("apply", apply)], args.steps, args.warmup)
    n_keep, n_split = acts[1] + acts[2], acts[2]
    row = sum(4 * w for _, w, _, _ in srcs)                      # 708 B of all arrays per triangle
    copy_row = sum(4 * w for _, w, _, p in srcs if p != capi.SPLIT_ZERO)  # read for a split parent
    apply_bytes = (acts[1] * row + n_split * copy_row + P_out * (row + 8) + 4 * (2 * P + 1))
    plan_bytes = P * (2 * 16 + 4 + 1)
    achieved = apply_bytes / (per["apply"] * 1e-3) / 1e9
    print(f"[bench] densify {ms:.3f} ms/step plan {per['plan']:.3f} apply {per['apply']:.3f}  P {P} -> {P_out} "
          f"(drop {acts[0]}, keep {acts[1]}, split {acts[2]})  apply {achieved:.0f} GB/s", file=sys.stderr, flush=True)
    line = {
        "metric": "prune + densify Mtriangles/s (plan + apply, parameters and Adam moments)",
        "value": round(P / (ms * 1e-3) / 1e6, 4), "unit": "Mtriangles/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic (seeded clustered triangle splats, statistics of a "
                                                   "1080p exact-mode forward)",
        "config": {"workload": f"prune + split {P} clustered splats (64 clusters, seed 1) -> {P_out}: drop "
                               f"{acts[0]}, keep {acts[1]}, split {acts[2]}; 12 arrays (vertices, opacity, sigma, "
                               f"SH + Adam moments)",
                   "kernel_ms": {k: round(v, 4) for k, v in per.items()},
                   "plan_gbs": round(plan_bytes / (per["plan"] * 1e-3) / 1e9, 1)},
        "roofline": dict(bound="hbm", achieved=round(achieved, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                         frac=round(achieved / HBM_PEAK_GBS, 5),
                         traffic=MEASURED_TRAFFIC["k_apply"] if P == 1_000_000 else None, kernel="k_apply",
                         bytes_per_launch=int(apply_bytes), avg_launch_ms=round(per["apply"], 4)),
    }
    print(json.dumps(line), flush=True)


def sceneio_bench(args, dev):
    """Scene files: the GPU halves of save (tt_scene_pack: tracer layout -> PLY vertex and
    face rows) and load (tt_scene_unpack: rows -> tracer layout, vertices gathered through
    the indices) on a 1M-splat scene with inputs resident in HBM. value = triangles per
    second of pack + unpack. The host legs (file write / read, PCIe) are timed once and
    reported in config."""
    import ctypes
    import tempfile

    import numpy as np
    import torch
    from diff_tritracer import capi, scenes, sceneio

    scene = scenes.clustered_scene(args.triangles, 64, seed=1)
    P = scene.P
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    verts, cumsum, opac, sigma, shs = T(scene.verts), T(scene.cumsum), T(scene.opac), T(scene.sigma), T(scene.shs)
    M = int(shs.shape[1])
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    vert_rows, face_rows, _ = sceneio.pack(verts, cumsum, opac, sigma, shs)
    pa = capi.ScenePackArgs(P, M, capi.ptr(verts), capi.ptr(cumsum), capi.ptr(opac), capi.ptr(sigma), capi.ptr(shs),
                            capi.ptr(face_rows), capi.ptr(vert_rows))
    S = capi.scene_face_bytes(M)
    ua = capi.SceneUnpackArgs()
    ua.P, ua.V, ua.face_stride, ua.M = P, 3 * P, S, M
    ua.count_offset, ua.index_offset, ua.opacity_offset, ua.sigma_offset = 0, 1, 13, 17
    for k in range(capi.SCENE_MAX_COEFFS * 3):
        ua.sh_offset[k] = -1
    for c in range(3):
        ua.sh_offset[c] = 21 + 4 * c
        for j in range(1, M):
            ua.sh_offset[3 * j + c] = 21 + 12 + 4 * (c * (M - 1) + j - 1)
    ua.vert_stride = 12
    ua.xyz_offset[0], ua.xyz_offset[1], ua.xyz_offset[2] = 0, 4, 8
    out_v, out_o, out_s = torch.empty_like(verts), torch.empty_like(opac), torch.empty_like(sigma)
    out_h = torch.empty_like(shs)
    out_n, out_c = torch.empty_like(cumsum), torch.empty_like(cumsum)
    bad = torch.zeros(1, dtype=torch.int32, device=dev)
    ua.face_dev, ua.vert_dev = capi.ptr(face_rows), capi.ptr(vert_rows)
    ua.vertices_out_dev, ua.npts_out_dev, ua.cumsum_out_dev = capi.ptr(out_v), capi.ptr(out_n), capi.ptr(out_c)
    ua.opacity_out_dev, ua.sigma_out_dev, ua.shs_out_dev, ua.bad_dev = (capi.ptr(out_o), capi.ptr(out_s),
                                                                          capi.ptr(out_h), capi.ptr(bad))
    L = capi.lib()

    class _Fr:  # what timed() needs
        pass
    fr = _Fr()
    fr.torch, fr.stream = torch, stream
    launches = [("pack", lambda: capi.check(L.tt_scene_pack(ctypes.byref(pa), sp), "tt_scene_pack")),
                ("unpack", lambda: capi.check(L.tt_scene_unpack(ctypes.byref(ua), sp), "tt_scene_unpack"))]
    ms, per = timed(fr, launches, args.steps, args.warmup)
    assert int(bad.item()) == 0 and torch.equal(out_h, shs) and torch.equal(out_o, opac)
    # host legs, once: save (pack + copy + write) and load (read + copy + unpack) of a real file
    with tempfile.TemporaryDirectory() as d:
        path = d + "/scene.ply"
        t0 = time.perf_counter()
        sceneio.save_ply(path, verts, cumsum, opac, sigma, shs)
        t1 = time.perf_counter()
        sceneio.load_ply(path, dev)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        size = os.path.getsize(path)
    unpack_bytes = P * (S + 36 + 244)   # face rows + gathered vertex rows read, tracer layout written
    pack_bytes = P * (240 + S + 36)     # tracer layout + cumsum read, face and vertex rows written
    ach = unpack_bytes / (per["unpack"] * 1e-3) / 1e9
    print(f"[bench] sceneio pack {per['pack']:.3f} ms ({pack_bytes / (per['pack'] * 1e-3) / 1e9:.0f} GB/s) unpack "
          f"{per['unpack']:.3f} ms ({ach:.0f} GB/s); file {size / 1e6:.0f} MB save {t1 - t0:.2f} s load {t2 - t1:.2f} s",
          file=sys.stderr, flush=True)
    line = {
        "metric": "scene file pack + unpack Mtriangles/s (device side of save / load)",
        "value": round(P / (ms * 1e-3) / 1e6, 4), "unit": "Mtriangles/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u8/fp32 (byte moves)", "data": "synthetic (seeded clustered triangle splats)",
        "config": {"workload": f"PLY scene rows <-> tracer layout, {P} clustered splats (64 clusters, seed 1), "
                               f"M={M}, {S}-byte face rows",
                   "kernel_ms": {k: round(v, 4) for k, v in per.items()},
                   "pack_gbs": round(pack_bytes / (per["pack"] * 1e-3) / 1e9, 1),
                   "file_mb": round(size / 1e6, 1), "save_s": round(t1 - t0, 3), "load_s": round(t2 - t1, 3)},
        "roofline": dict(bound="hbm", achieved=round(ach, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                         frac=round(ach / HBM_PEAK_GBS, 5),
                         traffic=MEASURED_TRAFFIC["k_scene_unpack"] if P == 1_000_000 else None,
                         kernel="k_scene_unpack",
                         bytes_per_launch=int(unpack_bytes), avg_launch_ms=round(per["unpack"], 4)),
    }
    print(json.dumps(line), flush=True)


class Frame:
    """Scene tensors, the tracer handle and per-frame buffers of one workload."""

    def __init__(self, dev, scene, k, cutoff, hitlist_cap=0):
        import torch
        from diff_tritracer import capi

        self.torch, self.capi, self.dev = torch, capi, dev
        T = lambda a: torch.from_numpy(a).to(dev)
        self.scene = scene
        self.P = scene.P
        self.verts, self.npts, self.cumsum = T(scene.verts), T(scene.npts), T(scene.cumsum)
        self.shs, self.opac, self.sigma = T(scene.shs), T(scene.opac.reshape(-1)), T(scene.sigma.reshape(-1))
        self.k, self.cutoff = k, cutoff
        self.hitlist_cap = hitlist_cap
        self.bg = torch.zeros(3, device=dev)
        self.stream = torch.cuda.current_stream()
        self.sp = self.stream.cuda_stream
        self.h = capi.TracerHandle()

    def buffers(self, W, H):
        torch, dev, P = self.torch, self.dev, self.P
        R = W * H
        self.W, self.H, self.R = W, H, R
        self.ro = torch.empty(H, W, 3, device=dev)
        self.rd = torch.empty(H, W, 3, device=dev)
        self.color, self.others = torch.zeros(3, R, device=dev), torch.zeros(7, R, device=dev)
        self.maxb, self.dens, self.scal = (torch.zeros(P, device=dev) for _ in range(3))
        g = torch.Generator(device="cpu").manual_seed(5)
        self.dpix = torch.randn(3, R, generator=g).to(dev)
        self.doth = torch.zeros(7, R, device=dev)
        self.dverts = torch.zeros(self.verts.shape[0], 3, device=dev)
        self.dsig, self.dop = torch.zeros(P, device=dev), torch.zeros(P, device=dev)
        self.dsh = torch.zeros(P, self.shs.shape[1], 3, device=dev)
        self.ival = torch.zeros(R, 4, device=dev)
        self.prop = torch.zeros(R, 2, device=dev)
        self.prop2 = torch.zeros(R, 2, device=dev)
        self.hl = self.capi.hitlist_buffers(R, self.hitlist_cap, dev) if self.hitlist_cap > 0 else None

    def camera(self, basis):
        self.capi.camera_rays(basis, self.W, self.H, self.ro, self.rd, self.sp)

    def fwd_args(self, mode="exact", interval_in=None, interval_out=None, hard=False, stats=None, trav=None,
                 fb_start=0.0, fb_end=0.0, hitlist=True, interval_event=None):
        c = self.capi
        return c.fwd_args(self.W, self.H, self.ro, self.rd, self.verts, self.npts, self.cumsum, self.shs, 3, self.opac,
                          self.sigma, self.bg, self.color, self.others, self.maxb, self.dens, self.scal, k=self.k,
                          cutoff=self.cutoff, chunk_mode=mode, interval_in=interval_in, interval_out=interval_out,
                          hard_cap=hard, fallback_start=fb_start, fallback_end=fb_end, stats=stats,
                          trav_stats=trav, hitlist=self.hl if hitlist else None, interval_event=interval_event)

    def bwd_args(self, mode="exact", interval=None, hitlist=True):
        return self.capi.bwd_args(self.W, self.H, self.ro, self.rd, self.verts, self.npts, self.cumsum, self.shs, 3,
                                  self.opac, self.sigma, self.color, self.others, self.dpix, self.doth, self.dverts,
                                  self.dsig, self.dop, self.dsh, k=self.k, cutoff=self.cutoff, chunk_mode=mode,
                                  zero_grads=False, interval=interval, hitlist=self.hl if hitlist else None,
                                  same_scene=True)  # the frame's scene does not change between fwd and bwd

    def zero_grads(self):
        for t in (self.dverts, self.dsig, self.dop, self.dsh):
            t.zero_()


def _median(xs):
    xs = sorted(xs)
    n = len(xs)
    return 0.5 * (xs[(n - 1) // 2] + xs[n // 2]) if n else float("nan")


def timed(fr, launches, steps, warmup, stats=None, tail=None):
    """Run `launches` (list of (name, fn)) as one step: warmup, then `steps` timed steps
    with HIP events around every launch on the launch stream (torch's current stream, which
    every launch here uses). Returns (median ms/step, {name: median ms}); `stats`, if given,
    receives the mean ms/step over the whole timed region, which ends after `tail()` (work
    that the steps left on another stream joined to the launch stream)."""
    torch = fr.torch
    for _ in range(warmup):
        for _, fn in launches:
            fn()
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(len(launches) + 1)] for _ in range(steps)]
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(fr.stream)
    for i in range(steps):
        evs[i][0].record(fr.stream)
        for j, (_, fn) in enumerate(launches):
            fn()
            evs[i][j + 1].record(fr.stream)
    if tail is not None:
        tail()
    stop.record(fr.stream)
    torch.cuda.synchronize()
    per = {name: _median([e[j].elapsed_time(e[j + 1]) for e in evs]) for j, (name, _) in enumerate(launches)}
    if stats is not None:
        stats["mean_ms"] = start.elapsed_time(stop) / steps
    return _median([e[0].elapsed_time(e[-1]) for e in evs]), per


def psnr(a, b):
    mse = float(((a - b) ** 2).mean())
    return float("inf") if mse == 0.0 else 10.0 * math.log10(1.0 / mse)


def train_proposal(fr, poses, W, H, steps, batch, seed):
    """Emit intervals with the plain exact forward for each training pose (rays made on
    the GPU), train the proposal on them, calibrate the margins on the training rays."""
    import numpy as np
    from diff_tritracer import scenes

    torch, capi = fr.torch, fr.capi
    n = len(poses)
    R = W * H
    O = torch.empty(n, H, W, 3, device=fr.dev)
    D = torch.empty(n, H, W, 3, device=fr.dev)
    IV = torch.empty(n, R, 4, device=fr.dev)
    fr.buffers(W, H)
    t0 = time.perf_counter()
    for i, eye in enumerate(poses):
        b = scenes.camera_basis(W, H, eye)
        capi.camera_rays(b, W, H, O[i], D[i], fr.sp)
        fr.ro, fr.rd = O[i], D[i]
        fr.h.fwd(fr.fwd_args("exact", interval_out=IV[i]), fr.sp)
    torch.cuda.synchronize()
    t_emit = time.perf_counter() - t0
    net = capi.Proposal(seed=seed)
    hist = torch.zeros(steps, device=fr.dev)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(fr.stream)
    net.train(O, D, IV, batch=batch, steps=steps, loss_hist=hist, stream=fr.sp)
    e1.record(fr.stream)
    torch.cuda.synchronize()
    t_train = e0.elapsed_time(e1)
    # margins: smallest widening that covers 99.5% of the training rays' blended ranges
    y = torch.empty(n * R, 2, device=fr.dev)
    net.forward(O, D, y, capi.PROPOSAL_RAW, stream=fr.sp)
    iv = IV.reshape(-1, 4)
    has = torch.isfinite(iv[:, 0])
    stop = torch.where(torch.isfinite(iv[:, 2]), iv[:, 2], iv[:, 1])
    ts = float(net.hp.t_scale)
    over_s = (y[:, 0] - iv[:, 0] / ts)[has]
    under_e = (stop / ts - y[:, 1])[has]
    q = 0.995
    ms = max(0.0, float(torch.quantile(over_s[:: max(1, over_s.numel() // 4_000_000)].float(), q)))
    me = max(0.0, float(torch.quantile(under_e[:: max(1, under_e.numel() // 4_000_000)].float(), q)))
    h = hist.cpu().numpy()
    info = dict(train_poses=n, train_rays=n * R, train_steps=steps, batch=batch, emit_s=round(t_emit, 3),
                train_ms=round(t_train, 2), loss_first=round(float(h[:20].mean()), 5),
                loss_last=round(float(h[-100:].mean()), 5), margin_start=round(ms, 5), margin_end=round(me, 5))
    del O, D, IV, y
    return net, ms, me, info


PROPOSAL_LEGS = ("proposal", "proposal_hard")


def frame_benchmark(args, dev, scene, cam_eye, label):
    import numpy as np
    import torch
    from diff_tritracer import capi, scenes

    W, H = args.width, args.height
    legs = [x for x in args.legs.split(",") if x] if args.legs else None
    want = lambda name: legs is None or name in legs
    fr = Frame(dev, scene, args.k, args.cutoff, args.hitlist_cap)
    P = fr.P
    # LBVH build and refit, timed separately (median of 5 after one warm-up)
    builds, refits = [], []
    for _ in range(6):
        fr.h.build(fr.verts, fr.cumsum, P, capi.BUILD, fr.sp)
        builds.append(fr.h.last_build_ms())
        fr.h.build(fr.verts, fr.cumsum, P, capi.REFIT, fr.sp)
        refits.append(fr.h.last_build_ms())
    build_ms, refit_ms = float(np.median(builds[1:])), float(np.median(refits[1:]))
    print(f"[bench] LBVH build {build_ms:.3f} ms, refit {refit_ms:.3f} ms", file=sys.stderr, flush=True)

    # the proposal: 16 training poses on the sphere of radius 3.5-4.5 around the scene
    net, pinfo, ms, me = None, None, 0.0, 0.0
    if any(want(n) for n in PROPOSAL_LEGS):
        rng = np.random.Generator(np.random.PCG64(101))
        poses = [scenes.orbit_eye(float(rng.uniform(3.5, 4.5)), 101, i) for i in range(TRAIN_POSES)]
        net, ms, me, pinfo = train_proposal(fr, poses, W, H, args.proposal_steps, args.proposal_batch, seed=3)
        print(f"[bench] proposal: {pinfo}", file=sys.stderr, flush=True)

    fr.buffers(W, H)
    fr.camera(scenes.camera_basis(W, H, cam_eye))
    R = fr.R
    # counting pass (plain, exact) for the algorithmic bytes
    stats = torch.zeros(R, 2, dtype=torch.int32, device=dev)
    trav = torch.zeros(R, 2, dtype=torch.int32, device=dev)
    fr.h.fwd(fr.fwd_args("exact", stats=stats, trav=trav), fr.sp)
    torch.cuda.synchronize()
    st = stats.to(torch.int64).sum(0).cpu().numpy()
    Hc, Hb = int(st[0]), int(st[1])
    tv_plain = trav.to(torch.int64).sum(0).cpu().numpy()
    print(f"[bench] hits consumed {Hc / R:.2f}/ray, blended {Hb / R:.2f}/ray; BVH nodes visited "
          f"{tv_plain[0] / R:.1f}/ray, triangles tested {tv_plain[1] / R:.1f}/ray", file=sys.stderr, flush=True)

    sp = fr.sp
    fa_plain = fr.fwd_args("exact")
    ba_plain = fr.bwd_args("exact")
    fa_retrace = fr.fwd_args("exact", hitlist=False)
    ba_retrace = fr.bwd_args("exact", hitlist=False)
    fa_compat = fr.fwd_args("compat")
    ba_compat = fr.bwd_args("compat")
    fa_emit = fr.fwd_args("exact", interval_out=fr.ival)
    # the proposal's forward reads its interval and emits none: the hit-list backward re-traverses nothing
    fa_soft = fr.fwd_args("exact", interval_in=fr.prop)
    ba_bound = fr.bwd_args("exact", interval=fr.ival)
    # the proposal's ceiling: each ray's own emitted interval (plain forward of this frame) as
    # (start, cap) -- start at the first blended (or terminating) hit, cap at the terminating or
    # last blended hit; exact mode, so the image is plain's whatever the interval
    perfect = torch.zeros(R, 2, device=dev)
    if want("proposal_ceiling"):
        fr.h.fwd(fa_emit, sp)
        iv = fr.ival.clone()
        first = torch.where(torch.isfinite(iv[:, 0]), iv[:, 0], iv[:, 2])
        stop_t = torch.where(torch.isfinite(iv[:, 2]), iv[:, 2], iv[:, 1])
        perfect = torch.stack([torch.where(torch.isfinite(first), first, torch.full_like(first, 3.0e38)),
                               torch.where(torch.isfinite(stop_t), stop_t, torch.zeros_like(stop_t))], 1).contiguous()
    # hard mode without fallbacks: the trace ends at the cap and rays that blend nothing (start
    # past the far plane) are not traced; with these intervals the image is plain's
    fa_ceiling = fr.fwd_args("exact", interval_in=perfect, hard=True)
    runs = {
        "plain": [("fwd", lambda: fr.h.fwd(fa_plain, sp)), ("zero", fr.zero_grads), ("bwd", lambda: fr.h.bwd(ba_plain, sp))],
        "plain_retrace": [("fwd", lambda: fr.h.fwd(fa_retrace, sp)), ("zero", fr.zero_grads),
                          ("bwd", lambda: fr.h.bwd(ba_retrace, sp))],
        "plain_compat": [("fwd", lambda: fr.h.fwd(fa_compat, sp)), ("zero", fr.zero_grads),
                         ("bwd", lambda: fr.h.bwd(ba_compat, sp))],
        "handoff_only": [("fwd", lambda: fr.h.fwd(fa_emit, sp)), ("zero", fr.zero_grads),
                         ("bwd", lambda: fr.h.bwd(ba_bound, sp))],
        "proposal_ceiling": [("fwd", lambda: fr.h.fwd(fa_ceiling, sp)), ("zero", fr.zero_grads),
                             ("bwd", lambda: fr.h.bwd(ba_plain, sp))],
    }
    if net is not None:
        ts = float(net.hp.t_scale)
        fa_hard = fr.fwd_args("exact", interval_in=fr.prop2, hard=True, fb_start=ms * ts, fb_end=me * ts)
        # exact: proposed cap only (start 0, continuation past the cap): the image equals plain
        propose = lambda: net.forward(fr.ro, fr.rd, fr.prop, capi.PROPOSAL_CAP, ms, me, stream=sp)
        # lossy: proposed start and hard cap, with the fallbacks
        propose2 = lambda: net.forward(fr.ro, fr.rd, fr.prop2, capi.PROPOSAL_INTERVAL, ms, me, stream=sp)
        # in the timed steps the network runs on a second stream (default priority) beside the
        # backward and proposes the intervals that the next step's forward reads (the views of a
        # training schedule are known ahead, so the next view's proposal overlaps this view's
        # backward); the forward's traversal waits for them (interval_ready_event). Every timed
        # step runs one network pass, one forward and one backward.
        s2 = torch.cuda.Stream()
        fork, ready, ready2 = torch.cuda.Event(), torch.cuda.Event(), torch.cuda.Event()
        for e in (ready, ready2):
            e.record(s2)  # creates the event's handle before the arguments take it
        fa_soft_ev = fr.fwd_args("exact", interval_in=fr.prop, interval_event=ready)
        fa_hard_ev = fr.fwd_args("exact", interval_in=fr.prop2, hard=True, fb_start=ms * ts, fb_end=me * ts,
                                 interval_event=ready2)

        def beside_backward(out, mode, ev, bwd):
            def run():
                fork.record(fr.stream)  # after this step's forward, which has read `out`
                s2.wait_event(fork)
                net.forward(fr.ro, fr.rd, out, mode, ms, me, stream=s2.cuda_stream)
                ev.record(s2)
                bwd()
            return run

        runs["proposal"] = [("fwd", lambda: fr.h.fwd(fa_soft_ev, sp)), ("zero", fr.zero_grads),
                            ("bwd", beside_backward(fr.prop, capi.PROPOSAL_CAP, ready,
                                                    lambda: fr.h.bwd(ba_plain, sp)))]
        runs["proposal_hard"] = [("fwd", lambda: fr.h.fwd(fa_hard_ev, sp)), ("zero", fr.zero_grads),
                                 ("bwd", beside_backward(fr.prop2, capi.PROPOSAL_INTERVAL, ready2,
                                                         lambda: fr.h.bwd(ba_plain, sp)))]
        tails = {"proposal": lambda: fr.stream.wait_event(ready),
                 "proposal_hard": lambda: fr.stream.wait_event(ready2)}
    runs = {k: v for k, v in runs.items() if want(k)}
    if net is None:
        tails = {}
    res, images = {}, {}
    for name, launches in runs.items():
        fr.h.set_kernel_timing(True)
        st_ = {}
        ms_step, per = timed(fr, launches, args.steps, args.warmup, st_, tails.get(name))
        # the traversal kernels alone (library events around them, over the timed steps)
        kf, kb = fr.h.kernel_times("fwd", args.steps), fr.h.kernel_times("bwd", args.steps)
        fr.h.set_kernel_timing(False)
        res[name] = dict(ms=ms_step, mean_ms=st_["mean_ms"], per=per, k_fwd=_median(kf), k_bwd=_median(kb),
                         k_fwd_mean=sum(kf) / max(1, len(kf)), k_bwd_mean=sum(kb) / max(1, len(kb)))
        images[name] = fr.color.clone()
        print(f"[bench] {name:16s} {ms_step:8.3f} ms/step " + " ".join(f"{k} {v:.3f}" for k, v in per.items()) +
              f"  kernels fwd {res[name]['k_fwd']:.3f} bwd {res[name]['k_bwd']:.3f}"
              f"  ({R / ms_step / 1e3:.2f} Mrays/s)", file=sys.stderr, flush=True)
    mr = lambda name: R / (res[name]["ms"] * 1e-3) / 1e6 if name in res else None
    same = lambda a, b: bool(np.array_equal(images[a].cpu().numpy().view(np.uint32),
                                            images[b].cpu().numpy().view(np.uint32)))
    cfg = {}
    if "plain" in res and "proposal_ceiling" in res:
        cfg["proposal_ceiling_image_equals_plain"] = same("proposal_ceiling", "plain")
    tv_prop = None
    if net is not None and "proposal" in res and "plain" in res:
        ref_img = images["plain"].cpu().numpy()
        cfg["proposal_image_equals_plain"] = same("proposal", "plain")
        cfg["psnr_loss_db"] = 0.0 if cfg["proposal_image_equals_plain"] else None
        if "proposal_hard" in res:
            cfg["hard_psnr_vs_plain_db"] = round(psnr(images["proposal_hard"].cpu().numpy(), ref_img), 2)
        # node visits of the proposal-bounded forward (the network's gain in traversal work)
        propose()
        fr.h.fwd(fr.fwd_args("exact", interval_in=fr.prop, stats=stats, trav=trav), sp)
        torch.cuda.synchronize()
        tv_prop = trav.to(torch.int64).sum(0).cpu().numpy()
        # extra held-out poses: equality and hard-mode PSNR
        rng2 = np.random.Generator(np.random.PCG64(202))
        held = []
        for i in range(HELD_OUT_EXTRA):
            eye = scenes.orbit_eye(float(rng2.uniform(3.5, 4.5)), 202, i)
            fr.camera(scenes.camera_basis(W, H, eye))
            fr.h.fwd(fa_plain, sp)
            a = fr.color.cpu().numpy()
            propose()
            fr.h.fwd(fa_soft, sp)
            b = fr.color.cpu().numpy()
            propose2()
            fr.h.fwd(fa_hard, sp)
            c = fr.color.cpu().numpy()
            held.append(dict(equal=bool(np.array_equal(a.view(np.uint32), b.view(np.uint32))),
                             hard_psnr_db=round(psnr(c, a), 2)))
        cfg["held_out_extra"] = held
        fr.camera(scenes.camera_basis(W, H, cam_eye))
    tv_ceil = None
    if "proposal_ceiling" in res:
        fr.h.fwd(fr.fwd_args("exact", interval_in=perfect, hard=True, stats=stats, trav=trav), sp)
        torch.cuda.synchronize()
        tv_ceil = trav.to(torch.int64).sum(0).cpu().numpy()

    # roofline of the dominant traversal kernel of the measured step: algorithmic bytes per
    # launch (per ray 64 + 48 Hc + 204 Hb forward, 104 + 48 Hc + 428 Hb backward, with the
    # plain trace's consumed / blended hits; node traffic, stack spill, hit lists, intervals
    # and gradient records excluded) / its median launch time (library events around it)
    bf = 64 * R + 48 * Hc + 204 * Hb
    bb = 104 * R + 48 * Hc + 428 * Hb
    step = "proposal" if "proposal" in res else next(iter(res))
    rs = res[step]
    # the forward's interval features per step (template bits: 1 interval read, 2 emitted, 4 hard cap)
    ival_bits = {"proposal": 1, "handoff_only": 3, "proposal_ceiling": 5, "proposal_hard": 5}.get(step, 0)
    kfwd = f"k_forward<{args.k},0,false,{ival_bits}>"
    if rs["k_bwd"] >= rs["k_fwd"]:
        dom, dom_bytes, dom_ms = f"k_backward_rec<{args.k},0>", bb, rs["k_bwd"]
    else:
        dom, dom_bytes, dom_ms = kfwd, bf, rs["k_fwd"]
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    default = P == 1_000_000 and (W, H) == (1920, 1080) and label.startswith("1080p")
    prof = MEASURED.get(dom) if default else None
    roofline = dict(bound="hbm", achieved=round(achieved, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 5), traffic=prof["traffic"] if prof else None, kernel=dom,
                    bytes_per_launch=int(dom_bytes), avg_launch_ms=round(dom_ms, 4),
                    avg_launch_ms_mean=round(rs["k_fwd_mean"] if dom == kfwd else rs["k_bwd_mean"], 4),
                    step_frac=round((bf + bb) / ((rs["k_fwd"] + rs["k_bwd"]) * 1e-3) / 1e9 / HBM_PEAK_GBS, 5))
    if prof:  # what bounds the kernel: rocprofv3 counters of this workload
        roofline.update(profiled_avg_ms=prof["avg_ms"], l2_hit=prof["l2_hit"], valu_busy=prof["valu_busy"],
                        counters_from=PROFILE_SOURCE)
    cfg.update({
        "workload": f"{label}: fwd+bwd, {P} clustered triangle splats, one {W}x{H} camera (held out from the "
                    f"proposal's {TRAIN_POSES} training poses), k={args.k}, cutoff {args.cutoff:g}, exact chunk mode, "
                    f"N(0,1) loss gradient",
        "triangles": P, "width": W, "height": H, "k": args.k, "cutoff": args.cutoff, "chunk_mode": "exact",
        "hitlist_cap": args.hitlist_cap, "timing": f"median of {args.steps} steps after {args.warmup} warm-ups",
    })
    for name in res:
        cfg[f"{name}_mrays"] = round(mr(name), 3)
    if "proposal" in res and "plain" in res:
        cfg["speedup_vs_plain"] = round(res["plain"]["ms"] / res["proposal"]["ms"], 4)
    if "proposal_hard" in res and "plain" in res:
        cfg["speedup_hard_vs_plain"] = round(res["plain"]["ms"] / res["proposal_hard"]["ms"], 4)
    if "proposal_ceiling" in res and "plain" in res:
        cfg["speedup_ceiling_vs_plain"] = round(res["plain"]["ms"] / res["proposal_ceiling"]["ms"], 4)
    cfg["ms"] = {k: round(v["ms"], 4) for k, v in res.items()}
    cfg["ms_mean"] = {k: round(v["mean_ms"], 4) for k, v in res.items()}
    cfg["kernel_ms"] = {k: {n: round(x, 4) for n, x in v["per"].items()} for k, v in res.items()}
    cfg["traversal_kernel_ms"] = {k: {"fwd": round(v["k_fwd"], 4), "bwd": round(v["k_bwd"], 4)} for k, v in res.items()}
    if net is not None and "proposal" in res:
        mlp_flops = 2 * 14848 * R
        # the network alone, on the launch stream (in the timed step it runs beside the backward)
        mlp_ms, _ = timed(fr, [("mlp", propose)], args.steps, args.warmup)
        tfl = mlp_flops / (mlp_ms * 1e-3) / 1e12
        # training: forward + backward ~ 3x the forward's MACs per ray (dW and dH per layer)
        train_flops = 3 * 2 * 14848 * args.proposal_batch * args.proposal_steps
        ttf = train_flops / (pinfo["train_ms"] * 1e-3) / 1e12
        cfg.update(mlp_ms=round(mlp_ms, 4), mlp_overlapped_in_step=True,
                   mlp_schedule="the next view's proposal on a second stream beside this view's backward",
                   mlp_tflops=round(tfl, 2), mlp_mfma_frac=round(tfl / BF16_PEAK_TFLOPS, 5),
                   mlp_train_tflops=round(ttf, 3), mlp_train_mfma_frac=round(ttf / BF16_PEAK_TFLOPS, 6),
                   proposal_training=pinfo)
    cfg["nodes_visited_per_ray_plain"] = round(float(tv_plain[0]) / R, 2)
    if tv_prop is not None:
        cfg["nodes_visited_per_ray_proposal"] = round(float(tv_prop[0]) / R, 2)
    if tv_ceil is not None:
        cfg["nodes_visited_per_ray_ceiling"] = round(float(tv_ceil[0]) / R, 2)
    cfg.update({"hits_consumed_per_ray": round(Hc / R, 3), "hits_blended_per_ray": round(Hb / R, 3),
                "lbvh_build_ms": round(build_ms, 3), "lbvh_refit_ms": round(refit_ms, 3)})
    out = dict(value=round(mr(step), 4), ms_per_step=round(rs["ms"], 4), config=cfg, roofline=roofline, step=step)
    return out, fr


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--k", type=int, default=16)
    ap.add_argument("--cutoff", type=float, default=1e-3)
    ap.add_argument("--triangles", type=int, default=0, help="default: 1M (frame), 2M (multiview), 3M (proposal)")
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--proposal-steps", type=int, default=2000)
    ap.add_argument("--proposal-batch", type=int, default=16384)
    ap.add_argument("--cpu-stride", type=int, default=1)
    ap.add_argument("--hitlist-cap", type=int, default=64,
                    help="blended hits per pixel the forward stores for the backward (0: the backward re-traverses)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="--workload multiview: skip the sweep, use k=16, block 256, 1 tile per fetch")
    ap.add_argument("--legs", default="",
                    help="comma-separated subset of the frame workload's steps (plain, plain_retrace, plain_compat, "
                         "handoff_only, proposal_ceiling, proposal, proposal_hard); default all")
    ap.add_argument("--workload", choices=["frame", "multiview", "proposal", "bounce", "train", "densify", "sceneio",
                                           "refit"], default="frame")
    ap.add_argument("--train-steps", type=int, default=1000, help="--workload refit: training steps per run")
    ap.add_argument("--rebuild-ratio", type=float, default=1.25, help="--workload refit: SAH cost ratio that rebuilds")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (calibrates rocprof traffic counters)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the tracer is single-device (rays shard across GPUs with a gradient "
              "all-reduce, not built here). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from diff_tritracer import scenes

    dev = torch.device("cuda:0")
    if args.copy_mib:
        n = args.copy_mib * (1 << 20) // 4
        src = torch.ones(n, device=dev)
        dst = torch.empty(n, device=dev)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": 4 * n, "launches": 3}))
        return
    if args.workload == "multiview":
        args.triangles = args.triangles or 2_000_000
        return multiview(args, dev)
    if args.workload == "bounce":
        args.triangles = args.triangles or 1_000_000
        return bounce(args, dev)
    if args.workload == "train":
        args.triangles = args.triangles or 1_000_000
        return train(args, dev)
    if args.workload == "densify":
        args.triangles = args.triangles or 1_000_000
        return densify_bench(args, dev)
    if args.workload == "sceneio":
        args.triangles = args.triangles or 1_000_000
        return sceneio_bench(args, dev)
    if args.workload == "refit":
        args.triangles = args.triangles or 1_000_000
        return refit_drift(args, dev)
    t0 = time.perf_counter()
    if args.workload == "proposal":
        P = args.triangles or 3_000_000
        scene = scenes.clustered_scene(P, 128, seed=4)
        label = "learned-proposal configuration (128 clusters, seed 4)"
    else:
        P = args.triangles or 1_000_000
        scene = scenes.clustered_scene(P, 64, seed=1)
        label = "1080p frame (64 clusters, seed 1)"
    eye = 4.0 * np.array([0.35, 0.3, 1.0]) / np.linalg.norm([0.35, 0.3, 1.0])
    print(f"[bench] scene {P} splats generated in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    out, fr = frame_benchmark(args, dev, scene, eye, label)

    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline (oracle) ...", file=sys.stderr, flush=True)
        ray_o, ray_d = scenes.look_at_rays(args.width, args.height, eye)
        cpu = cpu_baseline(scene, ray_o, ray_d, args.cpu_stride, args.cutoff, args.k)
        print(f"[bench] cpu baseline {cpu['value']} Mrays/s on {cpu['cores']} threads in {cpu['seconds']} s",
              file=sys.stderr, flush=True)

    line = {
        "metric": "fwd+bwd Mrays/s @1080p, 1M tris; proposal-culled vs plain at equal PSNR",
        "value": out["value"],
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": out["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded clustered triangle splats, pinhole camera rays)",
        "config": dict(out["config"], value_step=out["step"]),
        "roofline": out["roofline"],
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
