"""ctypes binding of the C ABI in include/tritracer_hip.h (libtritracer_hip.so).

This is the binding a caller without the torch extension would write; the
benchmark and the GPU parity tests drive the kernels through it so that what
they measure is exactly the C ABI. Device buffers are passed as raw pointers
(here taken from torch tensors on the current HIP device); every call is
asynchronous on the given stream.
"""
from __future__ import annotations

import ctypes as C
import os

from . import PACKAGE_DIR

LIB_PATH = os.path.join(PACKAGE_DIR, "libtritracer_hip.so")

TT_OK = 0
VARIANTS = {"forward": 0, "reflect": 1, "refract": 2, "envlight": 3, "bounce": 4}
BUILD, REFIT, REUSE, REFIT_AUTO = 0, 1, 2, 3
CHUNK = {"compat": 0, "exact": 1}

EXPORTED = ["tt_create", "tt_destroy", "tt_last_error", "tt_version", "tt_build", "tt_trace_fwd", "tt_trace_bwd",
            "tt_num_primitives", "tt_last_build_ms", "tt_multi_forward", "tt_multi_backward", "tt_multi_forward_noalloc",
            "tt_multi_capture",
            "tt_multi_replay", "tt_multi_autotune", "tt_proposal_default_hparams", "tt_proposal_num_params",
            "tt_proposal_create", "tt_proposal_destroy", "tt_proposal_set_params", "tt_proposal_get_params",
            "tt_proposal_forward", "tt_proposal_loss_grad", "tt_proposal_train", "tt_camera_rays", "tt_adam_step",
            "tt_densify_plan", "tt_densify_apply", "tt_scene_unpack", "tt_scene_pack", "tt_set_kernel_timing",
            "tt_kernel_times", "tt_set_rebuild_policy", "tt_bvh_cost", "tt_build_stats"]
MAX_CAMERAS = 32
TIMING_RING = 64

_vp = C.c_void_p
_fp = C.c_void_p  # device pointers travel as void*


class FwdArgs(C.Structure):
    _fields_ = [
        ("width", C.c_int), ("height", C.c_int),
        ("ray_o_dev", _fp), ("ray_d_dev", _fp),
        ("vertices_dev", _fp), ("num_vertices", C.c_int),
        ("npts_dev", _fp), ("cumsum_dev", _fp), ("P", C.c_int),
        ("shs_dev", _fp), ("M", C.c_int), ("sh_degree", C.c_int),
        ("opacities_dev", _fp), ("sigma_dev", _fp),
        ("bg_dev", _fp), ("extra_dev", _fp), ("extra_len", C.c_int),
        ("variant", C.c_int), ("k", C.c_int), ("chunk_mode", C.c_int),
        ("cutoff", C.c_float), ("near_t", C.c_float), ("far_t", C.c_float),
        ("color_dev", _fp), ("others_dev", _fp), ("max_blending_dev", _fp), ("density_dev", _fp),
        ("scaling_dev", _fp), ("stats_dev", _fp), ("trav_stats_dev", _fp),
        ("hit_log_cap", C.c_int), ("hit_log_prim_dev", _fp), ("hit_log_t_dev", _fp), ("hit_log_code_dev", _fp),
        ("hit_log_count_dev", _fp), ("interval_in_dev", _fp), ("interval_out_dev", _fp),
        ("interval_hard_cap", C.c_int), ("interval_fallback_start", C.c_float), ("interval_fallback_end", C.c_float),
        ("reflectivity_dev", _fp), ("ior_dev", _fp), ("envmap_dev", _fp), ("env_width", C.c_int),
        ("env_height", C.c_int), ("bounce_aux_dev", _fp),
        ("hitlist_dev", _fp), ("hitlist_count_dev", _fp), ("hitlist_cap", C.c_int), ("hitlist_color_dev", _fp),
        ("interval_ready_event", _fp),
    ]


class BwdArgs(C.Structure):
    _fields_ = [
        ("width", C.c_int), ("height", C.c_int),
        ("ray_o_dev", _fp), ("ray_d_dev", _fp),
        ("vertices_dev", _fp), ("num_vertices", C.c_int),
        ("npts_dev", _fp), ("cumsum_dev", _fp), ("P", C.c_int),
        ("shs_dev", _fp), ("M", C.c_int), ("sh_degree", C.c_int),
        ("opacities_dev", _fp), ("sigma_dev", _fp),
        ("color_dev", _fp), ("others_dev", _fp), ("dL_dpix_dev", _fp), ("dL_dothers_dev", _fp),
        ("k", C.c_int), ("chunk_mode", C.c_int),
        ("cutoff", C.c_float), ("near_t", C.c_float), ("far_t", C.c_float),
        ("zero_grads", C.c_int), ("dL_dvertices_rows", C.c_int),
        ("dL_dvertices_dev", _fp), ("dL_dsigma_dev", _fp), ("dL_dopacity_dev", _fp), ("dL_dsh_dev", _fp),
        ("interval_dev", _fp), ("variant", C.c_int), ("reflectivity_dev", _fp), ("envmap_dev", _fp),
        ("env_width", C.c_int), ("env_height", C.c_int), ("bounce_aux_dev", _fp), ("dL_dreflectivity_dev", _fp),
        ("dL_denvmap_dev", _fp),
        ("hitlist_dev", _fp), ("hitlist_count_dev", _fp), ("hitlist_cap", C.c_int), ("hitlist_color_dev", _fp),
        ("same_scene_as_forward", C.c_int),
    ]


class MultiArgs(C.Structure):
    _fields_ = [
        ("n_cameras", C.c_int), ("width", C.POINTER(C.c_int)), ("height", C.POINTER(C.c_int)),
        ("ray_o_dev", _fp), ("ray_d_dev", _fp),
        ("vertices_dev", _fp), ("num_vertices", C.c_int),
        ("npts_dev", _fp), ("cumsum_dev", _fp), ("P", C.c_int),
        ("shs_dev", _fp), ("M", C.c_int), ("sh_degree", C.c_int),
        ("opacities_dev", _fp), ("sigma_dev", _fp), ("bg_dev", _fp),
        ("k", C.c_int), ("chunk_mode", C.c_int), ("block", C.c_int), ("tiles_per_fetch", C.c_int),
        ("cutoff", C.c_float), ("near_t", C.c_float), ("far_t", C.c_float),
        ("color_dev", _fp), ("others_dev", _fp), ("max_blending_dev", _fp), ("density_dev", _fp),
        ("scaling_dev", _fp), ("stats_dev", _fp),
        ("dL_dpix_dev", _fp), ("dL_dothers_dev", _fp), ("zero_grads", C.c_int), ("dL_dvertices_rows", C.c_int),
        ("dL_dvertices_dev", _fp), ("dL_dsigma_dev", _fp), ("dL_dopacity_dev", _fp), ("dL_dsh_dev", _fp),
        ("hitlist_dev", _fp), ("hitlist_count_dev", _fp), ("hitlist_cap", C.c_int),
        ("cameras_dev", _fp), ("target_dev", _fp), ("loss_grad_dev", _fp), ("loss_dev", _fp), ("loss_kind", C.c_int),
        ("hitlist_color_dev", _fp),
    ]


CAMERA_DOUBLES = 14
LOSS = {"none": 0, "l2": 1, "l1": 2}


class AdamGroup(C.Structure):
    _fields_ = [("param_dev", _fp), ("grad_dev", _fp), ("exp_avg_dev", _fp), ("exp_avg_sq_dev", _fp),
                ("n", C.c_longlong), ("lr", C.c_double)]


DENSIFY_DROP, DENSIFY_KEEP, DENSIFY_SPLIT = 0, 1, 2
ROWS_TRIANGLE, ROWS_VERTEX = 0, 1
SPLIT_COPY, SPLIT_ZERO, SPLIT_SUBDIVIDE = 0, 1, 2
DENSIFY_MAX_ARRAYS = 16


def densify_work_ints(P: int) -> int:
    """TT_DENSIFY_WORK_INTS(P): scratch ints of tt_densify_plan."""
    return 2 * ((int(P) + 2047) // 2048) + 2


class DensifyPlanArgs(C.Structure):
    _fields_ = [("P", C.c_int), ("max_blending_dev", _fp), ("opacity_dev", _fp), ("density_dev", _fp),
                ("scaling_dev", _fp), ("score_dev", _fp), ("action_in_dev", _fp), ("min_blending", C.c_float),
                ("min_opacity", C.c_float), ("split_scaling", C.c_float), ("split_score", C.c_float),
                ("max_out", C.c_longlong), ("action_dev", _fp), ("offset_dev", _fp), ("work_dev", _fp)]


class DensifyArray(C.Structure):
    _fields_ = [("src_dev", _fp), ("dst_dev", _fp), ("width", C.c_int), ("rows", C.c_int), ("on_split", C.c_int)]


class DensifyApplyArgs(C.Structure):
    _fields_ = [("P", C.c_int), ("P_out", C.c_int), ("offset_dev", _fp), ("cumsum_dev", _fp),
                ("npts_out_dev", _fp), ("cumsum_out_dev", _fp), ("n_arrays", C.c_int),
                ("arrays", DensifyArray * DENSIFY_MAX_ARRAYS)]


SCENE_MAX_COEFFS = 16


def scene_face_bytes(M: int) -> int:
    """TT_SCENE_FACE_BYTES(M): bytes of a face row as tt_scene_pack writes it."""
    return 1 + 12 + 4 * (2 + 3 * int(M))


class SceneUnpackArgs(C.Structure):
    _fields_ = [("P", C.c_int), ("V", C.c_longlong), ("face_dev", _fp), ("face_stride", C.c_int),
                ("count_offset", C.c_int), ("index_offset", C.c_int), ("opacity_offset", C.c_int),
                ("sigma_offset", C.c_int), ("sh_offset", C.c_int * (SCENE_MAX_COEFFS * 3)), ("M", C.c_int),
                ("vert_dev", _fp), ("vert_stride", C.c_int), ("xyz_offset", C.c_int * 3),
                ("vertices_out_dev", _fp), ("npts_out_dev", _fp), ("cumsum_out_dev", _fp),
                ("opacity_out_dev", _fp), ("sigma_out_dev", _fp), ("shs_out_dev", _fp), ("bad_dev", _fp)]


class ScenePackArgs(C.Structure):
    _fields_ = [("P", C.c_int), ("M", C.c_int), ("vertices_dev", _fp), ("cumsum_dev", _fp), ("opacity_dev", _fp),
                ("sigma_dev", _fp), ("shs_dev", _fp), ("face_out_dev", _fp), ("vert_out_dev", _fp)]


class ProposalHparams(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("enc_scale", "t_scale", "t_skip", "tau_start", "tau_end", "lr", "beta1",
                                        "beta2", "eps")]


PROPOSAL_RAW, PROPOSAL_INTERVAL, PROPOSAL_CAP = 0, 1, 2

_lib = None


def lib():
    """Load libtritracer_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with diff-triangle-tracer_amd/build_native.py")
        L = C.CDLL(LIB_PATH)
        L.tt_create.argtypes = [C.POINTER(_vp)]
        L.tt_destroy.argtypes = [_vp]
        L.tt_last_error.restype = C.c_char_p
        L.tt_version.restype = C.c_char_p
        L.tt_build.argtypes = [_vp, _fp, C.c_int, _fp, C.c_int, C.c_int, _vp]
        L.tt_trace_fwd.argtypes = [_vp, C.POINTER(FwdArgs), _vp]
        L.tt_trace_bwd.argtypes = [_vp, C.POINTER(BwdArgs), _vp]
        L.tt_num_primitives.argtypes = [_vp]
        L.tt_last_build_ms.restype = C.c_float
        L.tt_last_build_ms.argtypes = [_vp]
        L.tt_multi_forward.argtypes = [_vp, C.POINTER(MultiArgs), _vp]
        L.tt_multi_backward.argtypes = [_vp, C.POINTER(MultiArgs), _vp]
        L.tt_multi_forward_noalloc.argtypes = [_vp, C.POINTER(MultiArgs), _vp]
        L.tt_set_kernel_timing.argtypes = [_vp, C.c_int]
        L.tt_set_rebuild_policy.argtypes = [_vp, C.c_float]
        L.tt_bvh_cost.argtypes = [_vp, _fp, _vp]
        L.tt_build_stats.argtypes = [_vp, C.POINTER(C.c_int), C.POINTER(C.c_float)]
        L.tt_kernel_times.argtypes = [_vp, C.c_int, C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_int)]
        L.tt_multi_capture.argtypes = [_vp, C.POINTER(MultiArgs), C.c_int, _vp]
        L.tt_multi_replay.argtypes = [_vp, _vp]
        L.tt_multi_autotune.argtypes = [_vp, C.POINTER(MultiArgs), C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int),
                                        C.c_int, C.POINTER(C.c_int), C.c_int, C.c_int, C.POINTER(C.c_float), _vp]
        _d3 = C.POINTER(C.c_double)
        L.tt_camera_rays.argtypes = [C.c_int, C.c_int, _d3, _d3, _d3, _d3, C.c_double, C.c_double, _fp, _fp, _vp]
        L.tt_adam_step.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, _vp]
        L.tt_densify_plan.argtypes = [C.POINTER(DensifyPlanArgs), _vp]
        L.tt_densify_apply.argtypes = [C.POINTER(DensifyApplyArgs), _vp]
        L.tt_scene_unpack.argtypes = [C.POINTER(SceneUnpackArgs), _vp]
        L.tt_scene_pack.argtypes = [C.POINTER(ScenePackArgs), _vp]
        L.tt_proposal_default_hparams.argtypes = [C.POINTER(ProposalHparams)]
        L.tt_proposal_default_hparams.restype = None
        L.tt_proposal_create.argtypes = [C.POINTER(_vp), C.POINTER(ProposalHparams), C.c_ulonglong]
        L.tt_proposal_destroy.argtypes = [_vp]
        L.tt_proposal_set_params.argtypes = [_vp, C.POINTER(C.c_float)]
        L.tt_proposal_get_params.argtypes = [_vp, C.POINTER(C.c_float)]
        L.tt_proposal_forward.argtypes = [_vp, _fp, _fp, C.c_longlong, C.c_float, C.c_float, C.c_int, _fp, _vp]
        L.tt_proposal_loss_grad.argtypes = [_vp, _fp, _fp, _fp, C.c_int, _fp, _fp, _vp]
        L.tt_proposal_train.argtypes = [_vp, _fp, _fp, _fp, C.c_longlong, C.c_int, C.c_int, _fp, _vp]
        _lib = L
    return _lib


def last_error() -> str:
    return lib().tt_last_error().decode()


def check(rc: int, what: str):
    if rc != TT_OK:
        raise RuntimeError(f"{what} failed ({rc}): {last_error()}")


def ptr(t) -> int | None:
    """Device address of a torch tensor (None for None / empty)."""
    if t is None or t.numel() == 0:
        return None
    return t.data_ptr()


def camera_table(bases, device):
    """Pinhole cameras for the fused ray generation of the multi-camera launches: a float64
    [n, 14] device tensor (eye, forward, right, up, tan_x, tan_y per camera) from
    scenes.camera_basis() dicts."""
    import torch

    rows = [list(b["eye"]) + list(b["forward"]) + list(b["right"]) + list(b["up"]) + [b["tan_x"], b["tan_y"]]
            for b in bases]
    return torch.tensor(rows, dtype=torch.float64, device=device)


def camera_rays(basis, width, height, ray_o, ray_d, stream=None):
    """Pinhole rays on the GPU into ray_o, ray_d ([H, W, 3] float32 tensors) from a
    scenes.camera_basis() dict; equal to scenes.look_at_rays bit for bit."""
    def d3(v):
        return (C.c_double * 3)(*[float(x) for x in v])

    check(lib().tt_camera_rays(int(width), int(height), d3(basis["eye"]), d3(basis["forward"]), d3(basis["right"]),
                               d3(basis["up"]), float(basis["tan_x"]), float(basis["tan_y"]), ptr(ray_o), ptr(ray_d),
                               stream), "tt_camera_rays")


def default_hparams(**kw) -> ProposalHparams:
    hp = ProposalHparams()
    lib().tt_proposal_default_hparams(C.byref(hp))
    for k, v in kw.items():
        setattr(hp, k, float(v))
    return hp


class Proposal:
    """The learned depth-interval proposal (tt_proposal_* of the C ABI)."""

    def __init__(self, seed=0, **hparams):
        self.hp = default_hparams(**hparams)
        self.h = _vp()
        check(lib().tt_proposal_create(C.byref(self.h), C.byref(self.hp), int(seed)), "tt_proposal_create")
        self.n_params = lib().tt_proposal_num_params()

    def close(self):
        if self.h:
            lib().tt_proposal_destroy(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        import numpy as np

        out = np.empty(self.n_params, np.float32)
        check(lib().tt_proposal_get_params(self.h, out.ctypes.data_as(C.POINTER(C.c_float))), "tt_proposal_get_params")
        return out

    def set_params(self, params):
        import numpy as np

        a = np.ascontiguousarray(params, np.float32).reshape(-1)
        if a.size != self.n_params:
            raise ValueError(f"expected {self.n_params} parameters, got {a.size}")
        check(lib().tt_proposal_set_params(self.h, a.ctypes.data_as(C.POINTER(C.c_float))), "tt_proposal_set_params")

    def forward(self, ray_o, ray_d, out, mode=PROPOSAL_RAW, margin_start=0.0, margin_end=0.0, stream=None):
        R = ray_o.numel() // 3
        check(lib().tt_proposal_forward(self.h, ptr(ray_o), ptr(ray_d), R, float(margin_start), float(margin_end),
                                        int(mode), ptr(out), stream), "tt_proposal_forward")

    def loss_grad(self, ray_o, ray_d, interval, grad, loss=None, stream=None):
        B = ray_o.numel() // 3
        check(lib().tt_proposal_loss_grad(self.h, ptr(ray_o), ptr(ray_d), ptr(interval), B, ptr(grad), ptr(loss),
                                          stream), "tt_proposal_loss_grad")

    def train(self, ray_o, ray_d, interval, batch, steps, loss_hist=None, stream=None):
        n = ray_o.numel() // 3
        check(lib().tt_proposal_train(self.h, ptr(ray_o), ptr(ray_d), ptr(interval), n, int(batch), int(steps),
                                      ptr(loss_hist), stream), "tt_proposal_train")


class TracerHandle:
    """Owns a tt_tracer*."""

    def __init__(self):
        h = _vp()
        check(lib().tt_create(C.byref(h)), "tt_create")
        self.h = h

    def close(self):
        if self.h:
            lib().tt_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def build(self, verts, cumsum, P, mode=BUILD, stream=None):
        check(lib().tt_build(self.h, ptr(verts), int(verts.numel() // 3), ptr(cumsum), int(P), int(mode), stream),
              "tt_build")

    def last_build_ms(self) -> float:
        return float(lib().tt_last_build_ms(self.h))

    def num_primitives(self) -> int:
        return int(lib().tt_num_primitives(self.h))

    def fwd(self, args: FwdArgs, stream=None):
        check(lib().tt_trace_fwd(self.h, C.byref(args), stream), "tt_trace_fwd")

    def bwd(self, args: BwdArgs, stream=None):
        check(lib().tt_trace_bwd(self.h, C.byref(args), stream), "tt_trace_bwd")

    def multi_fwd(self, args: "MultiArgs", stream=None):
        check(lib().tt_multi_forward(self.h, C.byref(args), stream), "tt_multi_forward")

    def set_rebuild_policy(self, max_ratio: float):
        """tt_build(REFIT_AUTO) rebuilds once the refit tree's SAH cost exceeds max_ratio x the built one."""
        check(lib().tt_set_rebuild_policy(self.h, float(max_ratio)), "tt_set_rebuild_policy")

    def bvh_cost(self, out, stream=None):
        """SAH cost of the current acceleration structure into the float32 tensor out[0]."""
        check(lib().tt_bvh_cost(self.h, ptr(out), stream), "tt_bvh_cost")

    def build_stats(self):
        """(builds, refits, automatic rebuilds, SAH cost after the last build, last evaluated cost)."""
        n = (C.c_int * 3)()
        c = (C.c_float * 2)()
        check(lib().tt_build_stats(self.h, n, c), "tt_build_stats")
        return n[0], n[1], n[2], c[0], c[1]

    def set_kernel_timing(self, enable=True):
        """Record HIP events around every traversal kernel (ring of TIMING_RING per direction)."""
        check(lib().tt_set_kernel_timing(self.h, 1 if enable else 0), "tt_set_kernel_timing")

    def kernel_times(self, which: str = "fwd", n: int = TIMING_RING):
        """Durations (ms) of the last n timed forward ('fwd') or backward ('bwd') traversal kernels."""
        buf = (C.c_float * max(1, n))()
        cnt = C.c_int(0)
        check(lib().tt_kernel_times(self.h, 0 if which == "fwd" else 1, buf, int(n), C.byref(cnt)), "tt_kernel_times")
        return [buf[i] for i in range(cnt.value)]

    def multi_fwd_noalloc(self, args: "MultiArgs", stream=None):
        """tt_multi_forward with its scratch in the library's static device array (no allocation)."""
        check(lib().tt_multi_forward_noalloc(self.h, C.byref(args), stream), "tt_multi_forward_noalloc")

    def multi_bwd(self, args: "MultiArgs", stream=None):
        check(lib().tt_multi_backward(self.h, C.byref(args), stream), "tt_multi_backward")

    def multi_capture(self, args: "MultiArgs", what=3, stream=None):
        check(lib().tt_multi_capture(self.h, C.byref(args), int(what), stream), "tt_multi_capture")

    def multi_replay(self, stream=None):
        check(lib().tt_multi_replay(self.h, stream), "tt_multi_replay")

    def multi_autotune(self, args: "MultiArgs", ks=(8, 16, 32), blocks=(64, 128, 256), tpf=(1, 2, 4), reps=3,
                       stream=None):
        """Times fwd+bwd for every configuration, stores the fastest in `args`; returns
        {(k, block, tiles_per_fetch): median ms} (skipped configurations omitted)."""
        ka, ba, ta = (C.c_int * len(ks))(*ks), (C.c_int * len(blocks))(*blocks), (C.c_int * len(tpf))(*tpf)
        res = (C.c_float * (len(ks) * len(blocks) * len(tpf)))()
        check(lib().tt_multi_autotune(self.h, C.byref(args), ka, len(ks), ba, len(blocks), ta, len(tpf), int(reps),
                                      res, stream), "tt_multi_autotune")
        out = {}
        for i, k in enumerate(ks):
            for j, b in enumerate(blocks):
                for l, t in enumerate(tpf):
                    v = res[(i * len(blocks) + j) * len(tpf) + l]
                    if v >= 0:
                        out[(k, b, t)] = float(v)
        return out


def adam_step(groups, step, beta1=0.9, beta2=0.999, eps=1e-8, zero_grads=False, stream=None):
    """One Adam step (torch.optim.Adam semantics) on up to 8 groups of (param, grad, exp_avg,
    exp_avg_sq, lr), float32 device tensors of equal sizes per group."""
    arr = (AdamGroup * max(1, len(groups)))()
    for i, (p, g, m, v, lr) in enumerate(groups):
        arr[i] = AdamGroup(ptr(p), ptr(g), ptr(m), ptr(v), int(p.numel()), float(lr))
    check(lib().tt_adam_step(C.cast(arr, C.c_void_p), len(groups), int(step), float(beta1), float(beta2), float(eps),
                             1 if zero_grads else 0, stream), "tt_adam_step")


def densify_plan(max_blending, opacity, density=None, scaling=None, score=None, action_in=None, min_blending=0.0,
                 min_opacity=0.0, split_scaling=float("inf"), split_score=0.0, max_out=0, stream=None):
    """tt_densify_plan on [P] float32 device tensors: returns (action uint8 [P], offset int32
    [P + 1]) on the device; offset[P] is the new triangle count."""
    import torch
    P = int(max_blending.numel() if max_blending is not None else action_in.numel())
    dev = (max_blending if max_blending is not None else action_in).device
    action = torch.empty(P, dtype=torch.uint8, device=dev)
    offset = torch.empty(P + 1, dtype=torch.int32, device=dev)
    work = torch.empty(densify_work_ints(P), dtype=torch.int32, device=dev)
    a = DensifyPlanArgs()
    a.P = P
    a.max_blending_dev, a.opacity_dev, a.density_dev = ptr(max_blending), ptr(opacity), ptr(density)
    a.scaling_dev, a.score_dev, a.action_in_dev = ptr(scaling), ptr(score), ptr(action_in)
    a.min_blending, a.min_opacity = float(min_blending), float(min_opacity)
    a.split_scaling, a.split_score, a.max_out = float(split_scaling), float(split_score), int(max_out)
    a.action_dev, a.offset_dev, a.work_dev = ptr(action), offset.data_ptr(), work.data_ptr()
    check(lib().tt_densify_plan(C.byref(a), stream), "tt_densify_plan")
    return action, offset


def densify_apply_args(P, P_out, offset, cumsum, arrays, npts_out=None, cumsum_out=None) -> DensifyApplyArgs:
    """arrays: list of (src, dst, width, rows, on_split) with float32 device tensors."""
    if len(arrays) > DENSIFY_MAX_ARRAYS:
        raise ValueError(f"at most {DENSIFY_MAX_ARRAYS} arrays per call")
    a = DensifyApplyArgs()
    a.P, a.P_out = int(P), int(P_out)
    a.offset_dev, a.cumsum_dev = ptr(offset), ptr(cumsum)
    a.npts_out_dev, a.cumsum_out_dev = ptr(npts_out), ptr(cumsum_out)
    a.n_arrays = len(arrays)
    for k, (src, dst, width, rows, on_split) in enumerate(arrays):
        a.arrays[k] = DensifyArray(ptr(src), ptr(dst), int(width), int(rows), int(on_split))
    return a


def densify_apply(P, P_out, offset, cumsum, arrays, npts_out=None, cumsum_out=None, stream=None):
    a = densify_apply_args(P, P_out, offset, cumsum, arrays, npts_out, cumsum_out)
    check(lib().tt_densify_apply(C.byref(a), stream), "tt_densify_apply")


def hitlist_buffers(R, cap, device, segments=1, colors=True):
    """Hit list of a forward for its backward: (hits int32 [segments, cap, R, 2], counts int32
    [segments, R][, colours float32 [segments, cap, R, 3]]); segments = 3 for the bounce variant
    (primary, reflected, refracted). With the colours the backward reads each hit's colour
    instead of evaluating its SH coefficients again."""
    import torch
    out = (torch.empty(int(segments), int(cap), int(R), 2, dtype=torch.int32, device=device),
           torch.empty(int(segments), int(R), dtype=torch.int32, device=device))
    if colors:
        out += (torch.empty(int(segments), int(cap), int(R), 3, dtype=torch.float32, device=device),)
    return out


def _set_hitlist(a, hitlist):
    hits, counts = hitlist[0], hitlist[1]
    a.hitlist_dev, a.hitlist_count_dev, a.hitlist_cap = ptr(hits), ptr(counts), int(hits.shape[-3])
    a.hitlist_color_dev = ptr(hitlist[2]) if len(hitlist) > 2 else None


def fwd_args(W, H, ray_o, ray_d, verts, npts, cumsum, shs, sh_degree, opac, sigma, bg, color, others, maxb, density,
             scaling, k=16, cutoff=1e-4, chunk_mode="compat", variant="forward", extra=None, stats=None,
             log=None, near=0.2, far=1000.0, trav_stats=None, interval_in=None, interval_out=None,
             hard_cap=False, fallback_start=0.0, fallback_end=0.0, bounce=None, hitlist=None,
             interval_event=None) -> FwdArgs:
    """bounce: dict(reflectivity [P], ior [P], envmap [H, W, 3], aux [R, 20]) for variant "bounce".
    hitlist: (hits int32 [cap, R, 2], counts int32 [R][, colours float32 [cap, R, 3]]) written for the
    backward (see hitlist_buffers)."""
    a = FwdArgs()
    a.width, a.height = int(W), int(H)
    a.ray_o_dev, a.ray_d_dev = ptr(ray_o), ptr(ray_d)
    a.vertices_dev, a.num_vertices = ptr(verts), int(verts.numel() // 3)
    a.npts_dev, a.cumsum_dev, a.P = ptr(npts), ptr(cumsum), int(cumsum.numel())
    a.shs_dev = ptr(shs)
    a.M = int(shs.shape[1]) if shs is not None and shs.numel() else 0
    a.sh_degree = int(sh_degree)
    a.opacities_dev, a.sigma_dev, a.bg_dev = ptr(opac), ptr(sigma), ptr(bg)
    a.extra_dev = ptr(extra)
    a.extra_len = int(extra.numel()) if extra is not None else 0
    a.variant = VARIANTS[variant] if isinstance(variant, str) else int(variant)
    a.k = int(k)
    a.chunk_mode = CHUNK[chunk_mode] if isinstance(chunk_mode, str) else int(chunk_mode)
    a.cutoff, a.near_t, a.far_t = float(cutoff), float(near), float(far)
    a.color_dev, a.others_dev = ptr(color), ptr(others)
    a.max_blending_dev, a.density_dev, a.scaling_dev = ptr(maxb), ptr(density), ptr(scaling)
    a.stats_dev = ptr(stats)
    a.trav_stats_dev = ptr(trav_stats)
    a.interval_in_dev, a.interval_out_dev = ptr(interval_in), ptr(interval_out)
    a.interval_hard_cap = 1 if hard_cap else 0
    # a torch.cuda.Event recorded where interval_in is produced: the traversal waits for it
    a.interval_ready_event = interval_event.cuda_event if interval_event is not None else None
    a.interval_fallback_start, a.interval_fallback_end = float(fallback_start), float(fallback_end)
    if bounce is not None:
        a.reflectivity_dev, a.ior_dev = ptr(bounce["reflectivity"]), ptr(bounce["ior"])
        a.envmap_dev, a.bounce_aux_dev = ptr(bounce["envmap"]), ptr(bounce["aux"])
        a.env_height, a.env_width = int(bounce["envmap"].shape[0]), int(bounce["envmap"].shape[1])
    if hitlist is not None:
        _set_hitlist(a, hitlist)
    if log is not None:
        prim, t, code, count = log
        a.hit_log_cap = int(prim.shape[-1])
        a.hit_log_prim_dev, a.hit_log_t_dev, a.hit_log_code_dev, a.hit_log_count_dev = (
            ptr(prim), ptr(t), ptr(code), ptr(count))
    return a


def bwd_args(W, H, ray_o, ray_d, verts, npts, cumsum, shs, sh_degree, opac, sigma, color, others, dL_dpix,
             dL_dothers, dverts, dsigma, dopac, dsh, k=16, cutoff=1e-4, chunk_mode="compat", zero_grads=True,
             near=0.2, far=1000.0, interval=None, bounce=None, hitlist=None, same_scene=False) -> BwdArgs:
    """bounce: dict(reflectivity, envmap, aux, dreflectivity [P], denvmap [H, W, 3]) selects the
    bounce variant's backward. hitlist: the forward's (hits, counts): read instead of re-traversing.
    same_scene: verts, cumsum and sigma hold what they held at the tracer's last forward (the
    backward then reuses that forward's window records when the pointers match)."""
    a = BwdArgs()
    a.width, a.height = int(W), int(H)
    a.ray_o_dev, a.ray_d_dev = ptr(ray_o), ptr(ray_d)
    a.vertices_dev, a.num_vertices = ptr(verts), int(verts.numel() // 3)
    a.npts_dev, a.cumsum_dev, a.P = ptr(npts), ptr(cumsum), int(cumsum.numel())
    a.shs_dev = ptr(shs)
    a.M = int(shs.shape[1]) if shs is not None and shs.numel() else 0
    a.sh_degree = int(sh_degree)
    a.opacities_dev, a.sigma_dev = ptr(opac), ptr(sigma)
    a.color_dev, a.others_dev, a.dL_dpix_dev, a.dL_dothers_dev = ptr(color), ptr(others), ptr(dL_dpix), ptr(dL_dothers)
    a.k = int(k)
    a.chunk_mode = CHUNK[chunk_mode] if isinstance(chunk_mode, str) else int(chunk_mode)
    a.cutoff, a.near_t, a.far_t = float(cutoff), float(near), float(far)
    a.zero_grads = 1 if zero_grads else 0
    a.dL_dvertices_rows = int(dverts.numel() // 3)
    a.dL_dvertices_dev, a.dL_dsigma_dev, a.dL_dopacity_dev, a.dL_dsh_dev = ptr(dverts), ptr(dsigma), ptr(dopac), ptr(dsh)
    a.interval_dev = ptr(interval)
    a.same_scene_as_forward = 1 if same_scene else 0
    if hitlist is not None:
        _set_hitlist(a, hitlist)
    if bounce is not None:
        a.variant = VARIANTS["bounce"]
        a.reflectivity_dev, a.envmap_dev = ptr(bounce["reflectivity"]), ptr(bounce["envmap"])
        a.env_height, a.env_width = int(bounce["envmap"].shape[0]), int(bounce["envmap"].shape[1])
        a.bounce_aux_dev = ptr(bounce["aux"])
        a.dL_dreflectivity_dev, a.dL_denvmap_dev = ptr(bounce["dreflectivity"]), ptr(bounce["denvmap"])
    return a


def multi_args(widths, heights, ray_o, ray_d, verts, npts, cumsum, shs, sh_degree, opac, sigma, bg, color, others,
               maxb, density, scaling, dL_dpix=None, dL_dothers=None, dverts=None, dsigma=None, dopac=None, dsh=None,
               k=16, cutoff=1e-3, chunk_mode="exact", block=128, tiles_per_fetch=1, zero_grads=True, stats=None,
               near=0.2, far=1000.0, hitlist=None, cameras=None, target=None, loss_grad=None, loss=None,
               loss_kind="l2") -> MultiArgs:
    """Arguments of the batched multi-camera path; per-camera buffers are concatenated
    in camera order (rays [sum R_c][3], colour [3][R_c] per camera, others [7][R_c]).
    hitlist: (hits [cap, sum R_c, 2], counts [sum R_c]) written by the forward, read by the backward.
    cameras: camera_table() -- rays generated in the launches (ray_o / ray_d may be None).
    target, loss_grad, loss: the fused loss -- target images like colour, dL/dcolour written
    to loss_grad (pass dL_dpix to feed the backward), the mean loss to loss[0]."""
    a = MultiArgs()
    n = len(widths)
    a.n_cameras = n
    a._w = (C.c_int * n)(*[int(w) for w in widths])
    a._h = (C.c_int * n)(*[int(h) for h in heights])
    a.width, a.height = a._w, a._h
    a.ray_o_dev, a.ray_d_dev = ptr(ray_o), ptr(ray_d)
    a.vertices_dev, a.num_vertices = ptr(verts), int(verts.numel() // 3)
    a.npts_dev, a.cumsum_dev, a.P = ptr(npts), ptr(cumsum), int(cumsum.numel())
    a.shs_dev = ptr(shs)
    a.M = int(shs.shape[1]) if shs is not None and shs.numel() else 0
    a.sh_degree = int(sh_degree)
    a.opacities_dev, a.sigma_dev, a.bg_dev = ptr(opac), ptr(sigma), ptr(bg)
    a.k = int(k)
    a.chunk_mode = CHUNK[chunk_mode] if isinstance(chunk_mode, str) else int(chunk_mode)
    a.block, a.tiles_per_fetch = int(block), int(tiles_per_fetch)
    a.cutoff, a.near_t, a.far_t = float(cutoff), float(near), float(far)
    a.color_dev, a.others_dev = ptr(color), ptr(others)
    a.max_blending_dev, a.density_dev, a.scaling_dev = ptr(maxb), ptr(density), ptr(scaling)
    a.stats_dev = ptr(stats)
    a.dL_dpix_dev, a.dL_dothers_dev = ptr(dL_dpix), ptr(dL_dothers)
    a.zero_grads = 1 if zero_grads else 0
    a.dL_dvertices_rows = int(dverts.numel() // 3) if dverts is not None else 0
    a.dL_dvertices_dev, a.dL_dsigma_dev, a.dL_dopacity_dev, a.dL_dsh_dev = ptr(dverts), ptr(dsigma), ptr(dopac), ptr(dsh)
    if hitlist is not None:
        _set_hitlist(a, hitlist)
    a.cameras_dev = ptr(cameras)
    a.target_dev, a.loss_grad_dev, a.loss_dev = ptr(target), ptr(loss_grad), ptr(loss)
    a.loss_kind = LOSS[loss_kind] if target is not None else 0
    return a
