#!/usr/bin/env python
"""Benchmark: plain fwd+bwd Mrays/s of the HIP triangle-splat tracer at 1080p on
1M clustered triangle splats (k = 16 hits per chunk, transmittance cutoff 1e-3,
random N(0,1) image-space loss gradient), on one MI355X.

One step = one forward launch + one backward launch (gradient buffers zeroed
inside the step) over the whole 1920x1080 frame, inputs resident in HBM. The
LBVH build and refit are timed separately and reported beside the metric.

Prints one JSON line (rank 0). Extra objects:
  roofline      : the dominant kernel's algorithmic bytes per launch / its average
                  launch time (HIP events on the launch stream) against 8 TB/s
  cpu_baseline  : the scalar C++/OpenMP oracle (same algorithm, FP32) on a
                  bounded ray sample of the same frame, timed on the host cores
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "diff-triangle-tracer_amd"))

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(R, Hc, Hb):
    """Bytes per launch the algorithm must move (ray/pixel I/O + per-hit gathers and
    gradient scatters), BVH node traffic excluded:
      forward : 64 R + 48 Hc + 204 Hb     backward : 104 R + 48 Hc + 428 Hb
    (Hc = consumed hits, Hb = blended hits, summed over the frame)."""
    return 64 * R + 48 * Hc + 204 * Hb, 104 * R + 48 * Hc + 428 * Hb


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
    return dict(value=round(R / t / 1e6, 5), unit="Mrays/s", cores=O.get_threads(), kind="port",
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
    P = args.triangles if args.triangles != 1_000_000 else 2_000_000
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
    a = capi.multi_args(ws, hs, ro, rd, verts, npts, cumsum, shs, 3, opac, sigma, bg, color, others, maxb, dens, scal,
                        dL_dpix=dpix, dL_dothers=doth, dverts=dverts, dsigma=dsig, dopac=dop, dsh=dsh, k=16,
                        cutoff=args.cutoff, chunk_mode="exact", block=128, tiles_per_fetch=1, zero_grads=True)
    t0 = time.perf_counter()
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
    line = {
        "metric": "multi-view fwd+bwd Mrays/s (one persistent hipGraph launch)",
        "value": round(Rt / (ms * 1e-3) / 1e6, 4), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded clustered triangle splats, 8 pinhole cameras on random poses at radius 3-5)",
        "config": {"workload": f"multi-view fwd+bwd, {P} clustered splats (64 clusters, seed 2), cameras "
                               + ", ".join(f"{w}x{hh}" for w, hh in zip(ws, hs)),
                   "rays": Rt, "k": a.k, "block": a.block, "tiles_per_fetch": a.tiles_per_fetch,
                   "chunk_mode": "exact", "cutoff": args.cutoff,
                   "autotune_ms": {f"k{k_}_b{b}_t{t}": round(v, 3) for (k_, b, t), v in sorted(sweep.items())}},
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--k", type=int, default=16)
    ap.add_argument("--cutoff", type=float, default=1e-3)
    ap.add_argument("--triangles", type=int, default=1_000_000)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--cpu-stride", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", choices=["frame", "multiview"], default="frame",
                    help="frame: the 1080p / 1M-splat headline; multiview: 2M splats, 8 uneven cameras in one "
                         "persistent hipGraph launch, autotuned")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (calibrates rocprof traffic counters)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the tracer is single-device (rays shard across GPUs with a gradient "
              "all-reduce, not built here). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from diff_tritracer import capi, scenes

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
        return multiview(args, dev)
    t0 = time.perf_counter()
    scene, ray_o, ray_d = scenes.scale_config(seed=1, P=args.triangles, width=args.width, height=args.height)
    gen_s = time.perf_counter() - t0
    H, W = ray_o.shape[:2]
    R = H * W
    P = scene.P
    print(f"[bench] scene {P} splats, {W}x{H} rays, generated in {gen_s:.1f} s", file=sys.stderr, flush=True)

    T = lambda a, dt=None: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    verts, npts, cumsum = T(scene.verts), T(scene.npts), T(scene.cumsum)
    shs, opac, sigma = T(scene.shs), T(scene.opac.reshape(-1)), T(scene.sigma.reshape(-1))
    ro, rd = T(ray_o), T(ray_d)
    bg = torch.zeros(3, device=dev)
    color = torch.zeros(3, R, device=dev)
    others = torch.zeros(7, R, device=dev)
    maxb = torch.zeros(P, device=dev)
    dens = torch.zeros(P, device=dev)
    scal = torch.zeros(P, device=dev)
    stats = torch.zeros(R, 2, dtype=torch.int32, device=dev)
    trav = torch.zeros(R, 2, dtype=torch.int32, device=dev)
    g = torch.Generator(device="cpu").manual_seed(5)
    dpix = torch.randn(3, R, generator=g).to(dev)
    doth = torch.zeros(7, R, device=dev)
    dverts = torch.zeros(verts.shape[0], 3, device=dev)
    dsig = torch.zeros(P, device=dev)
    dop = torch.zeros(P, device=dev)
    dsh = torch.zeros(P, shs.shape[1], 3, device=dev)

    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    h = capi.TracerHandle()
    # LBVH build and refit, timed separately (median of 5)
    builds, refits = [], []
    for i in range(6):
        h.build(verts, cumsum, P, capi.BUILD, sp)
        builds.append(h.last_build_ms())
        h.build(verts, cumsum, P, capi.REFIT, sp)
        refits.append(h.last_build_ms())
    build_ms = float(np.median(builds[1:]))
    refit_ms = float(np.median(refits[1:]))
    print(f"[bench] LBVH build {build_ms:.3f} ms, refit {refit_ms:.3f} ms", file=sys.stderr, flush=True)

    fa = capi.fwd_args(W, H, ro, rd, verts, npts, cumsum, shs, 3, opac, sigma, bg, color, others, maxb, dens, scal,
                       k=args.k, cutoff=args.cutoff)
    fa_stats = capi.fwd_args(W, H, ro, rd, verts, npts, cumsum, shs, 3, opac, sigma, bg, color, others, maxb, dens,
                             scal, k=args.k, cutoff=args.cutoff, stats=stats, trav_stats=trav)
    ba = capi.bwd_args(W, H, ro, rd, verts, npts, cumsum, shs, 3, opac, sigma, color, others, dpix, doth, dverts,
                       dsig, dop, dsh, k=args.k, cutoff=args.cutoff, zero_grads=False)
    grads = (dverts, dsig, dop, dsh)

    # counting pass for the algorithmic bytes
    h.fwd(fa_stats, sp)
    torch.cuda.synchronize()
    st = stats.to(torch.int64).sum(0).cpu().numpy()
    Hc, Hb = int(st[0]), int(st[1])
    tv = trav.to(torch.int64).sum(0).cpu().numpy()
    print(f"[bench] hits consumed {Hc / R:.2f}/ray, blended {Hb / R:.2f}/ray; BVH nodes visited {tv[0] / R:.1f}/ray, "
          f"triangles tested {tv[1] / R:.1f}/ray (forward)", file=sys.stderr, flush=True)

    def step(ev=None):
        # forward, gradient zeroing (part of the step: the reference allocates zeroed
        # gradients every call), backward; events bracket each kernel on the launch stream
        if ev is not None:
            ev[0].record(stream)
        h.fwd(fa, sp)
        if ev is not None:
            ev[1].record(stream)
        for t in grads:
            t.zero_()
        if ev is not None:
            ev[2].record(stream)
        h.bwd(ba, sp)
        if ev is not None:
            ev[3].record(stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(args.steps)]
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_wall = time.perf_counter()
    start.record(stream)
    for i in range(args.steps):
        step(evs[i])
    stop.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_wall
    total_ms = start.elapsed_time(stop)
    fwd_ms = [e[0].elapsed_time(e[1]) for e in evs]
    zero_ms = [e[1].elapsed_time(e[2]) for e in evs]
    bwd_ms = [e[2].elapsed_time(e[3]) for e in evs]
    ms_per_step = total_ms / args.steps
    mrays = R / (ms_per_step * 1e-3) / 1e6
    print(f"[bench] {ms_per_step:.3f} ms/step (fwd {np.mean(fwd_ms):.3f}, zero {np.mean(zero_ms):.3f}, "
          f"bwd {np.mean(bwd_ms):.3f}); "
          f"wall {wall:.2f} s", file=sys.stderr, flush=True)

    bf, bb = algorithmic_bytes(R, Hc, Hb)
    f_avg, b_avg = float(np.mean(fwd_ms)), float(np.mean(bwd_ms))
    if b_avg >= f_avg:
        dom, dom_bytes, dom_ms = f"k_backward<{args.k}>", bb, b_avg
    else:
        dom, dom_bytes, dom_ms = f"k_forward<{args.k},0>", bf, f_avg
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    roofline = dict(bound="hbm", achieved=round(achieved, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 5), traffic=None, kernel=dom,
                    bytes_per_launch=int(dom_bytes), avg_launch_ms=round(dom_ms, 4),
                    fwd_bwd_frac=round((bf + bb) / ((f_avg + b_avg) * 1e-3) / 1e9 / HBM_PEAK_GBS, 5))

    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline (oracle) ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(scene, ray_o, ray_d, args.cpu_stride, args.cutoff, args.k)
        print(f"[bench] cpu baseline {cpu['value']} Mrays/s on {cpu['cores']} threads in {cpu['seconds']} s",
              file=sys.stderr, flush=True)

    line = {
        "metric": "fwd+bwd Mrays/s @1080p, 1M tris (plain tracer; proposal-culled path not built yet)",
        "value": round(mrays, 4),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded clustered triangle splats, pinhole camera rays)",
        "config": {
            "workload": f"plain fwd+bwd, {P} clustered triangle splats (64 clusters, seed 1), one {W}x{H} camera, "
                        f"k={args.k}, cutoff {args.cutoff:g}, N(0,1) loss gradient",
            "triangles": P, "width": W, "height": H, "k": args.k, "cutoff": args.cutoff, "chunk_mode": "compat",
            "lbvh_build_ms": round(build_ms, 3), "lbvh_refit_ms": round(refit_ms, 3),
            "fwd_ms": round(f_avg, 4), "bwd_ms": round(b_avg, 4), "zero_grads_ms": round(float(np.mean(zero_ms)), 4),
            "hits_consumed_per_ray": round(Hc / R, 3), "hits_blended_per_ray": round(Hb / R, 3),
            "nodes_visited_per_ray": round(float(tv[0]) / R, 2), "triangles_tested_per_ray": round(float(tv[1]) / R, 2),
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
