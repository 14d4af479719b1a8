"""Shared test helpers: run the HIP tracer through the C ABI on numpy inputs,
and run the CPU oracle on the same inputs."""
from __future__ import annotations

import numpy as np

import oracle as O


def to_dev(a, device, dtype=None):
    import torch

    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.to(device)


class GpuScene:
    """Device copies of a SplatScene plus a tracer handle with a built LBVH."""

    def __init__(self, scene, device, build=True):
        import torch
        from diff_tritracer import capi

        self.scene = scene
        self.device = device
        self.verts = to_dev(scene.verts, device)
        self.npts = to_dev(scene.npts, device)
        self.cumsum = to_dev(scene.cumsum, device)
        self.shs = to_dev(scene.shs, device)
        self.opac = to_dev(scene.opac.reshape(-1), device)
        self.sigma = to_dev(scene.sigma.reshape(-1), device)
        self.P = scene.P
        self.h = capi.TracerHandle()
        self.torch = torch
        if build:
            self.h.build(self.verts, self.cumsum, self.P, capi.BUILD)

    def forward(self, ray_o, ray_d, k=16, cutoff=1e-3, chunk_mode="compat", variant="forward", extra=None,
                bg=(0.0, 0.0, 0.0), log_cap=0, stats=True, interval_in=None, interval_out=False, hard_cap=False,
                fallback=(0.0, 0.0), hitlist_cap=0, trav=False):
        from diff_tritracer import capi

        torch = self.torch
        H, W = ray_o.shape[0], ray_o.shape[1]
        R = H * W
        ro, rd = to_dev(ray_o, self.device), to_dev(ray_d, self.device)
        dev = self.device
        color = torch.zeros(3, R, device=dev)
        others = torch.zeros(7, R, device=dev)
        maxb = torch.zeros(self.P, device=dev)
        dens = torch.zeros(self.P, device=dev)
        scal = torch.zeros(self.P, device=dev)
        st = torch.zeros(R, 2, dtype=torch.int32, device=dev) if stats else None
        tv = torch.zeros(R, 2, dtype=torch.int32, device=dev) if trav else None
        bgt = torch.tensor(bg, dtype=torch.float32, device=dev)
        ext = to_dev(np.asarray(extra, np.float32), dev) if extra is not None else None
        log = None
        if log_cap:
            log = (torch.full((R, log_cap), -1, dtype=torch.int32, device=dev), torch.zeros(R, log_cap, device=dev),
                   torch.full((R, log_cap), -1, dtype=torch.int32, device=dev),
                   torch.zeros(R, dtype=torch.int32, device=dev))
        ivi = to_dev(np.asarray(interval_in, np.float32), dev) if interval_in is not None else None
        ivo = torch.full((R, 4), -7.0, device=dev) if interval_out else None
        hl = None
        if hitlist_cap:
            hl = capi.hitlist_buffers(R, hitlist_cap, dev)
            hl[0].fill_(-7)
            hl[1].fill_(-7)
        a = capi.fwd_args(W, H, ro, rd, self.verts, self.npts, self.cumsum, self.shs, 3, self.opac, self.sigma, bgt,
                          color, others, maxb, dens, scal, k=k, cutoff=cutoff, chunk_mode=chunk_mode,
                          variant=variant, extra=ext, stats=st, log=log, interval_in=ivi, interval_out=ivo,
                          hard_cap=hard_cap, fallback_start=fallback[0], fallback_end=fallback[1], hitlist=hl,
                          trav_stats=tv)
        self.h.fwd(a, torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        out = dict(color=color.cpu().numpy(), others=others.cpu().numpy(), max_blending=maxb.cpu().numpy(),
                   density=dens.cpu().numpy(), scaling=scal.cpu().numpy(), _color=color, _others=others)
        if stats:
            out["stats"] = st.cpu().numpy()
        if trav:
            out["trav"] = tv.cpu().numpy()
        if log is not None:
            out["log"] = tuple(x.cpu().numpy() for x in log)
        if ivo is not None:
            out["interval"] = ivo.cpu().numpy()
        if hl is not None:
            out["_hitlist"] = hl
            out["hitlist"] = tuple(x.cpu().numpy() for x in hl)
        return out

    def backward(self, ray_o, ray_d, color, others, dL_dpix, dL_dothers, k=16, cutoff=1e-3, chunk_mode="compat",
                 interval=None, hitlist=None, same_scene=False):
        from diff_tritracer import capi

        torch = self.torch
        H, W = ray_o.shape[0], ray_o.shape[1]
        dev = self.device
        ro, rd = to_dev(ray_o, dev), to_dev(ray_d, dev)
        col = color if isinstance(color, torch.Tensor) else to_dev(color, dev)
        oth = others if isinstance(others, torch.Tensor) else to_dev(others, dev)
        gp = to_dev(np.asarray(dL_dpix, np.float32), dev)
        go = to_dev(np.asarray(dL_dothers, np.float32), dev)
        dverts = torch.empty(self.verts.shape[0], 3, device=dev)
        dsig = torch.empty(self.P, device=dev)
        dop = torch.empty(self.P, device=dev)
        dsh = torch.empty(self.P, self.shs.shape[1], 3, device=dev)
        a = capi.bwd_args(W, H, ro, rd, self.verts, self.npts, self.cumsum, self.shs, 3, self.opac, self.sigma, col,
                          oth, gp, go, dverts, dsig, dop, dsh, k=k, cutoff=cutoff, chunk_mode=chunk_mode,
                          zero_grads=True,
                          interval=to_dev(np.asarray(interval, np.float32), dev) if interval is not None else None,
                          hitlist=hitlist, same_scene=same_scene)
        self.h.bwd(a, torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        return dict(dverts=dverts.cpu().numpy(), dsigma=dsig.cpu().numpy(), dopac=dop.cpu().numpy(),
                    dsh=dsh.cpu().numpy())


def oracle_scene(scene, build_bvh=True):
    return O.Scene(scene.verts, scene.npts, scene.cumsum, scene.shs, scene.sh_degree, scene.opac, scene.sigma,
                   build_bvh=build_bvh)


def log_lists(export, R):
    """Oracle log export -> per-pixel list of (prim, t, code) for single-segment renders."""
    off, spix, prims, ts, codes = export
    out = [None] * R
    for s in range(len(spix)):
        a, b = off[s], off[s + 1]
        out[spix[s]] = (prims[a:b], ts[a:b], codes[a:b])
    return out


def compare_hit_logs(gpu_log, oracle_log_export, R):
    """Bit-exact comparison of the GPU debug hit log with the oracle decision log.
    Returns the number of mismatching rays (0 = identical)."""
    gprim, gt, gcode, gcount = gpu_log
    cap = gprim.shape[1]
    ol = log_lists(oracle_log_export, R)
    bad = 0
    for r in range(R):
        p, t, c = ol[r]
        n = len(p)
        if gcount[r] != n:
            bad += 1
            continue
        m = min(n, cap)
        if not (np.array_equal(gprim[r, :m], p[:m]) and np.array_equal(gt[r, :m].view(np.uint32),
                                                                      t[:m].view(np.uint32))
                and np.array_equal(gcode[r, :m], c[:m])):
            bad += 1
    return bad


def grad_close(g, ref, rel=1e-3):
    """Elementwise |g - ref| <= rel * max(|ref|, floor), floor = rel * max|ref| (entries
    near zero are judged on the tensor's scale). Returns (ok, worst ratio)."""
    g = np.asarray(g, np.float64).reshape(-1)
    ref = np.asarray(ref, np.float64).reshape(-1)
    scale = np.abs(ref).max() if ref.size else 0.0
    if scale == 0.0:
        return bool(np.abs(g).max(initial=0.0) == 0.0), 0.0
    den = np.maximum(np.abs(ref), rel * scale)
    ratio = np.abs(g - ref) / den
    worst = float(ratio.max())
    return worst <= rel, worst


def padded_log(export, R, cap):
    """Oracle decision log (single-segment renders) as padded arrays like the GPU debug log:
    prim [R, cap] (-1 past the end), t [R, cap] (0), code [R, cap] (-1), count [R]."""
    off, spix, prims, ts, codes = export
    cnt = np.diff(off).astype(np.int64)
    count = np.zeros(R, np.int32)
    count[spix] = cnt
    n = int(off[-1])
    pix = np.repeat(spix.astype(np.int64), cnt)
    idx = np.arange(n, dtype=np.int64) - np.repeat(off[:-1].astype(np.int64), cnt)
    m = idx < cap
    P = np.full((R, cap), -1, np.int32)
    T = np.zeros((R, cap), np.float32)
    Cd = np.full((R, cap), -1, np.int32)
    P[pix[m], idx[m]] = prims[m]
    T[pix[m], idx[m]] = ts[m]
    Cd[pix[m], idx[m]] = codes[m]
    return P, T, Cd, count


def compare_hit_logs_fast(gpu_log, export, R):
    """Vectorised compare_hit_logs: the number of rays whose count or first `cap` logged
    hits (prim, t bits, decision) differ (the GPU buffers must start as -1 / 0 / -1)."""
    gprim, gt, gcode, gcount = gpu_log
    P, T, Cd, count = padded_log(export, R, gprim.shape[1])
    bad = (gcount != count) | np.any(gprim != P, axis=1) | np.any(gcode != Cd, axis=1) | \
        np.any(gt.view(np.uint32) != T.view(np.uint32), axis=1)
    return int(bad.sum())
