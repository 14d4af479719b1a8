"""diff_tritracer: differentiable triangle-splat ray tracer on MI355X (gfx950).

Drop-in for the reference package of the same name (waizui/diff-triangle-tracer,
diff_tritracer/__init__.py): the same ``TracerSettings`` fields, the same
``Tracer(forward_program, backward_program)`` module with the same forward
arguments and return tuple, and the same autograd wiring -- the gradient of the
flat triangle-point buffer, of sigma viewed as [-1, 1], of the SH coefficients
and of the opacities; no gradient for rays, background or precomputed colours.

Underneath, ``_C.OptixTracer`` (csrc/tracer.h) runs hand-written HIP kernels
through the C ABI of libtritracer_hip.so instead of OptiX. Program names keep
their reference spelling: "forward.ptx", "reflect.ptx", "refract.ptx",
"envlight.ptx" select the forward program, "backward.ptx" the backward.
"""
from __future__ import annotations

import os
from typing import NamedTuple

import torch
import torch.nn as nn

from . import _C  # the native extension; importing the package fails loudly without it

PACKAGE_DIR = os.path.dirname(os.path.abspath(__file__))


class TracerSettings(NamedTuple):
    """Per-call settings (reference __init__.py:37-43).

    rebuild_gas: 0 build the acceleration structure, 1 refit it to moved
    vertices (same triangles), 2 reuse it as is; 3 (added here) refit and rebuild
    once the refit tree's SAH cost exceeds the tracer's policy ratio
    (``Tracer.set_rebuild_policy``, default 1.25) x its cost after the last build.
    extra_params: float tensor
    (r, cx, cy, cz[, samples]) of the sphere used by the reflect / refract /
    envlight programs. scale_modifier is accepted and unused, as in the reference.
    """
    bg: torch.Tensor
    scale_modifier: float
    sh_degree: int
    debug: bool
    rebuild_gas: int
    extra_params: torch.Tensor


def _cpu_snapshot(values):
    return tuple(v.detach().cpu().clone() if isinstance(v, torch.Tensor) else v for v in values)


def _call_with_snapshot(fn, args, debug: bool, dump_name: str, stage: str):
    """Run fn(*args); in debug mode save a CPU copy of the arguments if it raises
    (the reference's snapshot_fw.dump / snapshot_bw.dump behaviour)."""
    if not debug:
        return fn(*args)
    snapshot = _cpu_snapshot(args)
    try:
        return fn(*args)
    except Exception:
        torch.save(snapshot, dump_name)
        print(f"\nAn error occurred in {stage}. Arguments saved to {dump_name} for debugging.")
        raise


class _Tracer(torch.autograd.Function):
    @staticmethod
    def forward(ctx, tracer, ray_o, ray_d, triangles_points, sigma, num_points_per_triangle,
                cumsum_of_points_per_triangle, number_of_points, shs, colors_precomp, opacities, scaling,
                density_factor, tracer_settings: TracerSettings):
        args = (ray_o, ray_d, triangles_points, sigma, num_points_per_triangle, cumsum_of_points_per_triangle,
                number_of_points, shs, tracer_settings.sh_degree, colors_precomp, opacities, scaling, density_factor,
                tracer_settings.bg, tracer_settings.extra_params, tracer_settings.rebuild_gas,
                tracer_settings.debug)
        # the hit list is only worth writing when a backward can read it
        store = any(ctx.needs_input_grad)
        if not store:
            tracer.set_store_hits(False)
        try:
            res = _call_with_snapshot(tracer.trace_forward, args, tracer_settings.debug, "snapshot_fw.dump", "forward")
        finally:
            if not store:
                tracer.set_store_hits(True)
        color, others, gpu_time, scaling, density_factor, max_blending = res.data

        # stored blended hits and their colours (None: the backward re-traverses)
        ctx.hits = (res.hits, res.hit_counts, res.hit_colors)
        ctx.tracer = tracer
        ctx.tracer_settings = tracer_settings
        ctx.number_of_points = number_of_points
        ctx.save_for_backward(ray_o, ray_d, triangles_points, sigma, num_points_per_triangle,
                              cumsum_of_points_per_triangle, color, others, colors_precomp, opacities, shs, scaling)
        return color, gpu_time, scaling, density_factor, others, max_blending

    @staticmethod
    def backward(ctx, grad_color, _grad_time, _grad_scaling, _grad_density, grad_others, _grad_max_blending):
        (ray_o, ray_d, triangles_points, sigma, num_points_per_triangle, cumsum_of_points_per_triangle, color,
         others, _colors_precomp, opacities, shs, scaling) = ctx.saved_tensors
        st = ctx.tracer_settings
        args = (ray_o, ray_d, st.bg, triangles_points, sigma, num_points_per_triangle, cumsum_of_points_per_triangle,
                color, others, ctx.number_of_points, shs, st.sh_degree, opacities, grad_color, grad_others, scaling,
                st.debug, *ctx.hits)
        res = _call_with_snapshot(ctx.tracer.trace_backward, args, st.debug, "snapshot_bw.dump", "backward")
        grad_triangles, grad_sigma, grad_opacities, grad_sh = res.data
        # gradient layouts as returned by the reference (__init__.py:205-227)
        return (None, None, None, grad_triangles.flatten(0), grad_sigma.view(-1, 1), None, None, None, grad_sh,
                None, grad_opacities, None, None, None)


class Tracer(nn.Module):
    """Differentiable triangle-splat tracer (reference __init__.py:232-283).

    ``Tracer("forward.ptx", "backward.ptx")`` for plain rendering; pass
    "reflect.ptx", "refract.ptx" or "envlight.ptx" as the first argument for the
    sphere-bounce programs. ``set_options(k, cutoff, chunk_mode, hit_cache)`` changes
    the hits gathered per chunk (8/16/32), the transmittance cutoff (reference: 1e-4),
    the chunk resume rule ("compat" or "exact") and the blended hits per pixel the
    forward stores for the backward (0: the backward re-traverses, as the reference).
    """

    def __init__(self, forwad_ptx="forward.ptx", backward_ptx="backward.ptx") -> None:
        super().__init__()
        self.tracer = _C.OptixTracer(PACKAGE_DIR, forwad_ptx, backward_ptx)

    def set_options(self, k: int = 16, cutoff: float = 1e-4, chunk_mode="compat", hit_cache: int = 64):
        mode = {"compat": 0, "exact": 1}[chunk_mode] if isinstance(chunk_mode, str) else int(chunk_mode)
        self.tracer.set_options(int(k), float(cutoff), mode, int(hit_cache))
        return self

    def set_rebuild_policy(self, max_ratio: float = 1.25):
        """rebuild_gas = 3: rebuild when the refit tree's SAH cost exceeds max_ratio x the built one."""
        self.tracer.set_rebuild_policy(float(max_ratio))
        return self

    def build_stats(self):
        """(builds, refits, automatic rebuilds, SAH cost after the last build, last evaluated cost)."""
        return self.tracer.build_stats()

    def forward(self, ray_o, ray_d, triangles_points, sigma, num_points_per_triangle, cumsum_of_points_per_triangle,
                number_of_points, opacities, scaling, density_factor, shs, colors_precomp,
                tracer_settings: TracerSettings):
        if (shs is None) == (colors_precomp is None):
            raise Exception("Please provide excatly one of either SHs or precomputed colors!")
        device = ray_o.device
        if shs is None:
            shs = torch.empty(0, device=device)
        if colors_precomp is None:
            colors_precomp = torch.empty(0, device=device)
        grad = torch.is_grad_enabled()  # no backward under no_grad: the forward stores no hit list
        if not grad:
            self.tracer.set_store_hits(False)
        try:
            return _Tracer.apply(self.tracer, ray_o, ray_d, triangles_points, sigma, num_points_per_triangle,
                                 cumsum_of_points_per_triangle, number_of_points, shs, colors_precomp, opacities,
                                 scaling, density_factor, tracer_settings)
        finally:
            if not grad:
                self.tracer.set_store_hits(True)


from .optim import FusedAdam  # noqa: E402  (the optimizer step on the library's Adam kernel)

__all__ = ["TracerSettings", "Tracer", "_Tracer", "PACKAGE_DIR", "_C", "FusedAdam"]
