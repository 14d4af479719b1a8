"""Pruning and densification of a triangle-splat scene from the forward's statistics, on
the GPU (tt_densify_plan + tt_densify_apply).

The reference's forward fills max_blending, density_factor and scaling per triangle
(tritracer/forward.cu:34-74, :134-143; returned by diff_tritracer/__init__.py:95-139) and
leaves their use to its callers. prune_and_split is that use for a training loop: it drops
triangles whose max_blending or opacity is below a floor, splits large blended triangles
into four by midpoint subdivision (within an optional cap on the new count), and rewrites
the parameter tensors -- and the Adam moments of an optimizer over them -- in place.
"""
from __future__ import annotations

import torch

from . import capi


def _check(t, name):
    if t is None:
        return
    if not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()):
        raise RuntimeError(f"{name} must be a contiguous float32 GPU tensor")


def plan(max_blending, opacity, density=None, scaling=None, score=None, action_in=None, min_blending=0.0,
         min_opacity=0.0, split_scaling=float("inf"), split_score=0.0, max_out=0):
    """Decide per triangle: returns (action uint8 [P] (0 drop, 1 keep, 2 split), offset int32
    [P + 1] (first output row of each triangle), new count). Reads the count back (one sync)."""
    for t, n in ((max_blending, "max_blending"), (opacity, "opacity"), (density, "density"), (scaling, "scaling"),
                 (score, "score")):
        _check(t if t is None else t.reshape(-1), n)
    flat = [None if t is None else t.reshape(-1) for t in (max_blending, opacity, density, scaling, score)]
    stream = torch.cuda.current_stream((max_blending if max_blending is not None else action_in).device).cuda_stream
    action, offset = capi.densify_plan(*flat, action_in=action_in, min_blending=min_blending, min_opacity=min_opacity,
                                       split_scaling=split_scaling, split_score=split_score, max_out=max_out,
                                       stream=stream)
    return action, offset, int(offset[-1].item())


def _moments(optimizer, p):
    if optimizer is None:
        return []
    st = optimizer.state.get(p)
    if not st:
        return []
    return [k for k in ("exp_avg", "exp_avg_sq") if k in st and torch.is_tensor(st[k]) and st[k].shape == p.shape]


@torch.no_grad()
def prune_and_split(vertices, cumsum, triangle_params, max_blending, opacity, density=None, scaling=None, score=None,
                    optimizer=None, action_in=None, min_blending=0.0, min_opacity=0.0, split_scaling=float("inf"),
                    split_score=0.0, max_out=0):
    """vertices: [V, 3] (or flat [3 V]) tensor (triangle i's points at rows cumsum[i] .. cumsum[i] + 2);
    triangle_params: [P, ...] tensors (opacity, sigma, SH, ...). Every tensor's .data is
    replaced by the new set (vertices dense: [3 P_out, 3]); children copy the parent's
    per-triangle values; Adam moments (exp_avg, exp_avg_sq) in `optimizer`'s state are
    carried for kept triangles and start at zero for children. Gradients are cleared.
    Returns (action, npts int32 [P_out], cumsum int32 [P_out])."""
    P = int(cumsum.numel())
    dev = vertices.device
    action, offset, P_out = plan(max_blending, opacity, density, scaling, score, action_in, min_blending,
                                 min_opacity, split_scaling, split_score, max_out)
    jobs = []  # (tensor owner, setter, src, width, rows, on_split, new shape)
    _check(vertices, "vertices")
    vshape = (3 * P_out, 3) if vertices.dim() == 2 else (9 * P_out,)  # [V, 3] or flat, as given
    jobs.append((vertices, None, vertices, 9, capi.ROWS_VERTEX, capi.SPLIT_SUBDIVIDE, vshape))
    for k in _moments(optimizer, vertices):
        jobs.append((vertices, k, optimizer.state[vertices][k], 9, capi.ROWS_VERTEX, capi.SPLIT_ZERO, vshape))
    for i, t in enumerate(triangle_params):
        _check(t, f"triangle_params[{i}]")
        if t.shape[0] != P:
            raise RuntimeError(f"triangle_params[{i}] has {t.shape[0]} rows, expected {P}")
        w = t.numel() // P if P else 1
        shape = (P_out, *t.shape[1:])
        jobs.append((t, None, t, w, capi.ROWS_TRIANGLE, capi.SPLIT_COPY, shape))
        for k in _moments(optimizer, t):
            jobs.append((t, k, optimizer.state[t][k], w, capi.ROWS_TRIANGLE, capi.SPLIT_ZERO, shape))
    outs = [torch.empty(shape, dtype=torch.float32, device=dev) for *_, shape in jobs]
    npts = torch.empty(P_out, dtype=torch.int32, device=dev)
    cs = torch.empty(P_out, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    for c in range(0, len(jobs), capi.DENSIFY_MAX_ARRAYS):
        arrays = [(src, out, w, rows, on_split) for (_, _, src, w, rows, on_split, _), out in
                  zip(jobs[c:c + capi.DENSIFY_MAX_ARRAYS], outs[c:c + capi.DENSIFY_MAX_ARRAYS])]
        first = c == 0
        capi.densify_apply(P, P_out, offset, cumsum, arrays, npts if first else None, cs if first else None,
                           stream=stream)
    for (owner, key, *_), out in zip(jobs, outs):
        if key is None:
            owner.data = out
            owner.grad = None
        else:
            optimizer.state[owner][key] = out
    return action, npts, cs
