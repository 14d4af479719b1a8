"""ctypes front end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module, and only as the checker. The product path (diff_tritracer) never does.

The oracle restates the reference tracer's algorithm (waizui/diff-triangle-tracer,
tritracer/forward.cu, backward.cu, closestk.h, common.h, reflect/refract/envlight.cu)
in scalar C++; see oracle/tt_oracle.h for the per-function citations.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

VARIANTS = {"forward": 0, "reflect": 1, "refract": 2, "envlight": 3, "bounce": 4}
CODE_NEAR, CODE_ALPHA, CODE_BLEND, CODE_TERM = 0, 1, 2, 3


def build(quiet: bool = True) -> str:
    """Compile liboracle.so with oracle/Makefile (g++, OpenMP)."""
    out = subprocess.run(["make", "-C", HERE, "-s", "-B"], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if not quiet:
        print(out.stdout + out.stderr)
    return LIB_PATH


class _Opts(C.Structure):
    _fields_ = [
        ("k", C.c_int),
        ("chunk_mode", C.c_int),
        ("variant", C.c_int),
        ("use_bvh", C.c_int),
        ("cutoff", C.c_float),
        ("near_t", C.c_float),
        ("far_t", C.c_float),
        ("bg", C.c_float * 3),
        ("extra", C.POINTER(C.c_float)),
        ("extra_len", C.c_int),
        ("tstart", C.POINTER(C.c_float)),
        ("refl", C.POINTER(C.c_float)),
        ("ior", C.POINTER(C.c_float)),
        ("env", C.POINTER(C.c_float)),
        ("env_w", C.c_int),
        ("env_h", C.c_int),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, fp, ip, dp = C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_int), C.POINTER(C.c_double)
        L.or_scene_create.restype = vp
        L.or_scene_create.argtypes = [fp, C.c_int, ip, ip, C.c_int, fp, C.c_int, C.c_int, fp, fp, C.c_int]
        L.or_scene_free.argtypes = [vp]
        L.or_scene_bvh_seconds.restype = C.c_double
        L.or_scene_bvh_seconds.argtypes = [vp]
        L.or_set_threads.argtypes = [C.c_int]
        L.or_set_abs_vertex.argtypes = [C.c_int]
        L.or_get_threads.restype = C.c_int
        for n in ("or_log2f", "or_exp2f"):
            getattr(L, n).restype = C.c_float
            getattr(L, n).argtypes = [C.c_float]
        L.or_powf.restype = C.c_float
        L.or_powf.argtypes = [C.c_float, C.c_float]
        L.or_sincosf.argtypes = [C.c_float, fp, fp]
        L.or_intersect.restype = C.c_int
        L.or_intersect.argtypes = [fp, fp, fp, fp, fp, fp]
        L.or_cos_hemi.argtypes = [C.c_uint, fp, C.c_int, fp]
        L.or_window_f32.argtypes = [fp, fp, C.c_float, fp]
        L.or_window_f64.argtypes = [dp, dp, C.c_double, dp]
        L.or_sh_f32.argtypes = [fp, C.c_int, fp, fp, ip]
        L.or_sh_f64.argtypes = [dp, C.c_int, dp, dp, ip]
        L.or_gather.restype = C.c_int
        L.or_gather.argtypes = [vp, fp, fp, C.c_float, C.c_longlong, C.c_int, C.c_float, C.c_int, ip, fp]
        L.or_forward_f32.restype = C.c_int
        L.or_forward_f32.argtypes = [vp, fp, fp, C.c_int, C.POINTER(_Opts), fp, fp, fp, fp, fp, ip]
        L.or_backward_f32.restype = C.c_int
        L.or_backward_f32.argtypes = [vp, fp, fp, C.c_int, C.POINTER(_Opts), fp, fp, fp, fp, fp, fp, fp, fp]
        L.or_record.restype = vp
        L.or_record.argtypes = [vp, fp, fp, C.c_int, C.POINTER(_Opts)]
        L.or_log_free.argtypes = [vp]
        L.or_log_size.restype = C.c_longlong
        L.or_log_size.argtypes = [vp]
        L.or_log_segments.restype = C.c_longlong
        L.or_log_segments.argtypes = [vp]
        L.or_log_export.argtypes = [vp, C.POINTER(C.c_longlong), ip, ip, fp, ip]
        L.or_forward_f64.restype = C.c_int
        L.or_forward_f64.argtypes = [vp, dp, dp, dp, dp, dp, dp]
        L.or_backward_f64.restype = C.c_int
        L.or_backward_f64.argtypes = [vp, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp]
        L.or_loss_f64.restype = C.c_double
        L.or_loss_f64.argtypes = [vp, dp, dp, dp, dp, dp, dp]
        L.or_forward_f64_bounce.restype = C.c_int
        L.or_forward_f64_bounce.argtypes = [vp, dp, dp, dp, dp, dp, dp, dp, dp]
        L.or_loss_f64_bounce.restype = C.c_double
        L.or_loss_f64_bounce.argtypes = [vp, dp, dp, dp, dp, dp, dp, dp, dp]
        L.or_backward_f64_bounce.restype = C.c_int
        L.or_backward_f64_bounce.argtypes = [vp, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp]
        L.or_env_taps.argtypes = [fp, C.c_int, C.c_int, ip, dp]
        L.or_log_segment_rays.argtypes = [vp, fp, fp, ip]
        _lib = L
    return _lib


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def _i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def set_threads(n: int) -> None:
    lib().or_set_threads(int(n))


def get_threads() -> int:
    return int(lib().or_get_threads())


# --------------------------------------------------------------- elementary
def log2f(x):
    return np.float32(lib().or_log2f(float(x)))


def exp2f(x):
    return np.float32(lib().or_exp2f(float(x)))


def powf(x, y):
    return np.float32(lib().or_powf(float(x), float(y)))


def sincosf(x):
    s, c = C.c_float(), C.c_float()
    lib().or_sincosf(float(x), C.byref(s), C.byref(c))
    return np.float32(s.value), np.float32(c.value)


def cos_hemi(seed, normal, count):
    """`count` successive envlight hemisphere directions (envlight.cu:25-51) around the unit
    `normal`, drawn from one LCG state `seed`: float32 [count, 3]."""
    n = _f32(normal)
    out = np.zeros((int(count), 3), np.float32)
    lib().or_cos_hemi(int(seed) & 0xFFFFFFFF, n.ctypes.data_as(C.POINTER(C.c_float)), int(count),
                      out.ctypes.data_as(C.POINTER(C.c_float)))
    return out


def intersect(o, d, v0, v1, v2):
    o, d, v0, v1, v2 = (_f32(a) for a in (o, d, v0, v1, v2))
    t = np.zeros(1, np.float32)
    hit = lib().or_intersect(_f(o), _f(d), _f(v0), _f(v1), _f(v2), _f(t))
    return (t[0] if hit else None)


def window(verts, hit, sigma, f64=False):
    """Returns dict(Cx, phi_final, phi_max, edge_dists, normal, dist)."""
    if f64:
        v = np.ascontiguousarray(verts, np.float64).reshape(9)
        h = np.ascontiguousarray(hit, np.float64)
        out = np.zeros(10, np.float64)
        lib().or_window_f64(_d(v), _d(h), float(sigma), _d(out))
    else:
        v = _f32(verts).reshape(9)
        h = _f32(hit)
        out = np.zeros(10, np.float32)
        lib().or_window_f32(_f(v), _f(h), float(sigma), _f(out))
    return dict(Cx=out[0], phi_final=out[1], phi_max=out[2], edge_dists=out[3:6].copy(),
                normal=out[6:9].copy(), dist=out[9])


def sh_color(sh, deg, direction, f64=False):
    cl = np.zeros(3, np.int32)
    if f64:
        s = np.ascontiguousarray(sh, np.float64).reshape(-1)
        dr = np.ascontiguousarray(direction, np.float64)
        rgb = np.zeros(3, np.float64)
        lib().or_sh_f64(_d(s), int(deg), _d(dr), _d(rgb), _i(cl))
    else:
        s = _f32(sh).reshape(-1)
        dr = _f32(direction)
        rgb = np.zeros(3, np.float32)
        lib().or_sh_f32(_f(s), int(deg), _f(dr), _f(rgb), _i(cl))
    return rgb, cl.astype(bool)


# -------------------------------------------------------------------- scene
class Scene:
    """Holds the (float32/int32) scene arrays and the oracle's own BVH.

    verts: [V,3]; npts/cumsum: [P]; shs: [P,M,3]; opac, sigma: [P] (or [P,1]).
    """

    def __init__(self, verts, npts, cumsum, shs, sh_degree, opac, sigma, build_bvh=True):
        self.verts = _f32(verts).reshape(-1, 3)
        self.npts = _i32(npts).reshape(-1)
        self.cumsum = _i32(cumsum).reshape(-1)
        self.P = int(self.cumsum.shape[0])
        self.shs = _f32(shs).reshape(self.P, -1, 3) if self.P else np.zeros((0, 16, 3), np.float32)
        self.M = int(self.shs.shape[1])
        self.deg = int(sh_degree)
        self.opac = _f32(opac).reshape(-1)
        self.sigma = _f32(sigma).reshape(-1)
        self.h = lib().or_scene_create(_f(self.verts), self.verts.shape[0], _i(self.npts), _i(self.cumsum), self.P,
                                       _f(self.shs), self.M, self.deg, _f(self.opac), _f(self.sigma),
                                       1 if build_bvh else 0)

    @property
    def bvh_seconds(self) -> float:
        return float(lib().or_scene_bvh_seconds(self.h))

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_scene_free(self.h)
            self.h = None


def make_opts(k=16, chunk_mode="compat", variant="forward", use_bvh=True, cutoff=1e-4, near=0.2, far=1000.0,
              bg=(0.0, 0.0, 0.0), extra=None, tstart=None, refl=None, ior=None, env=None):
    """tstart: optional per-ray start distance [R] (forward variant, exact chunk mode;
    honoured by forward_f32 and the decision log). refl, ior [P] and env [H][W][3]: the
    bounce variant's per-triangle reflectivity, index of refraction and environment map."""
    o = _Opts()
    o.k = int(k)
    o.chunk_mode = {"compat": 0, "exact": 1}[chunk_mode] if isinstance(chunk_mode, str) else int(chunk_mode)
    o.variant = VARIANTS[variant] if isinstance(variant, str) else int(variant)
    o.use_bvh = 1 if use_bvh else 0
    o.cutoff = float(cutoff)
    o.near_t = float(near)
    o.far_t = float(far)
    for i in range(3):
        o.bg[i] = float(bg[i])
    keep = None
    if extra is not None and len(extra) > 0:
        keep = _f32(extra).reshape(-1)
        o.extra = _f(keep)
        o.extra_len = keep.shape[0]
    else:
        o.extra = C.POINTER(C.c_float)()
        o.extra_len = 0
    o._keep = keep  # keep the extra buffer alive with the struct
    if tstart is not None:
        ts = _f32(tstart).reshape(-1)
        o.tstart = _f(ts)
        o._keep_ts = ts
    else:
        o.tstart = C.POINTER(C.c_float)()
    o._keep_b = []
    for name, arr in (("refl", refl), ("ior", ior), ("env", env)):
        if arr is not None:
            a = _f32(arr)
            setattr(o, name, _f(a.reshape(-1)))
            o._keep_b.append(a)
        else:
            setattr(o, name, C.POINTER(C.c_float)())
    if env is not None:
        o.env_h, o.env_w = int(np.shape(env)[0]), int(np.shape(env)[1])
    return o


def gather(scene: Scene, o, d, tlo=0.0, plo=-1, k=16, far=1000.0, use_bvh=True):
    o, d = _f32(o), _f32(d)
    prim = np.zeros(k, np.int32)
    t = np.zeros(k, np.float32)
    n = lib().or_gather(scene.h, _f(o), _f(d), float(tlo), int(plo), int(k), float(far), 1 if use_bvh else 0,
                        _i(prim), _f(t))
    return prim.view(np.uint32)[:n].astype(np.int64), t[:n]


def env_taps(d, env_w, env_h):
    """Texel indices (y * env_w + x) and bilinear weights of the environment lookup of d."""
    idx = np.zeros(4, np.int32)
    w = np.zeros(4, np.float64)
    lib().or_env_taps(_f(_f32(d).reshape(3)), int(env_w), int(env_h), _i(idx), _d(w))
    return idx, w


def forward_f32(scene: Scene, ray_o, ray_d, opts, scaling=None, density=None):
    """f32 mirror of the HIP forward. Returns dict with color[3,R], others[7,R], max_blending,
    density, scaling, stats[R,2] (consumed, blended)."""
    ro = _f32(ray_o).reshape(-1, 3)
    rd = _f32(ray_d).reshape(-1, 3)
    R = ro.shape[0]
    color = np.zeros((3, R), np.float32)
    others = np.zeros((7, R), np.float32)
    maxb = np.zeros(scene.P, np.float32)
    dens = _f32(density).copy() if density is not None else np.zeros(scene.P, np.float32)
    scal = _f32(scaling).copy() if scaling is not None else np.zeros(scene.P, np.float32)
    stats = np.zeros((R, 2), np.int32)
    lib().or_forward_f32(scene.h, _f(ro), _f(rd), R, C.byref(opts), _f(color), _f(others), _f(maxb), _f(dens),
                         _f(scal), _i(stats))
    return dict(color=color, others=others, max_blending=maxb, density=dens, scaling=scal, stats=stats)


def backward_f32(scene: Scene, ray_o, ray_d, opts, color, others, dL_dpix, dL_dothers, num_vert_rows=None):
    ro = _f32(ray_o).reshape(-1, 3)
    rd = _f32(ray_d).reshape(-1, 3)
    R = ro.shape[0]
    nv = scene.verts.shape[0] if num_vert_rows is None else int(num_vert_rows)
    dverts = np.zeros((nv, 3), np.float32)
    dsig = np.zeros(scene.P, np.float32)
    dop = np.zeros(scene.P, np.float32)
    dsh = np.zeros((scene.P, scene.M, 3), np.float32)
    col, oth = _f32(color).reshape(3, R), _f32(others).reshape(7, R)
    gp, go = _f32(dL_dpix).reshape(3, R), _f32(dL_dothers).reshape(7, R)
    lib().or_backward_f32(scene.h, _f(ro), _f(rd), R, C.byref(opts), _f(col), _f(oth), _f(gp), _f(go), _f(dverts),
                          _f(dsig), _f(dop), _f(dsh))
    return dict(dverts=dverts, dsigma=dsig, dopac=dop, dsh=dsh)


class Log:
    """f32 decision log of a render; the f64 replay and FD surrogate run over it."""

    def __init__(self, scene: Scene, ray_o, ray_d, opts):
        self.scene = scene
        self.ro = _f32(ray_o).reshape(-1, 3)
        self.rd = _f32(ray_d).reshape(-1, 3)
        self.R = self.ro.shape[0]
        self.opts = opts
        self.h = lib().or_record(scene.h, _f(self.ro), _f(self.rd), self.R, C.byref(opts))

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_log_free(self.h)
            self.h = None

    def export(self):
        """Per segment: pixel, and the consumed hits (prim, t, code)."""
        n = int(lib().or_log_size(self.h))
        ns = int(lib().or_log_segments(self.h))
        off = np.zeros(ns + 1, np.int64)
        spix = np.zeros(ns, np.int32)
        prims = np.zeros(max(n, 1), np.int32)
        ts = np.zeros(max(n, 1), np.float32)
        codes = np.zeros(max(n, 1), np.int32)
        lib().or_log_export(self.h, off.ctypes.data_as(C.POINTER(C.c_longlong)), _i(spix), _i(prims), _f(ts),
                            _i(codes))
        return off, spix, prims[:n], ts[:n], codes[:n]

    def segment_rays(self):
        """Per segment: origin [3], direction [3] (float32) and role (0 primary, 1 reflected, 2 refracted)."""
        ns = int(lib().or_log_segments(self.h))
        o = np.zeros((ns, 3), np.float32)
        d = np.zeros((ns, 3), np.float32)
        role = np.zeros(ns, np.int32)
        lib().or_log_segment_rays(self.h, _f(o), _f(d), _i(role))
        return o, d, role

    def params64(self, verts=None, opac=None, sigma=None, shs=None):
        s = self.scene
        v = np.ascontiguousarray(s.verts if verts is None else verts, np.float64).reshape(-1, 3)
        o = np.ascontiguousarray(s.opac if opac is None else opac, np.float64).reshape(-1)
        g = np.ascontiguousarray(s.sigma if sigma is None else sigma, np.float64).reshape(-1)
        h = np.ascontiguousarray(s.shs if shs is None else shs, np.float64).reshape(s.P, s.M, 3)
        return v, o, g, h

    def forward_f64(self, **params):
        v, o, g, h = self.params64(**params)
        color = np.zeros((3, self.R), np.float64)
        others = np.zeros((7, self.R), np.float64)
        lib().or_forward_f64(self.h, _d(v), _d(o), _d(g), _d(h), _d(color), _d(others))
        return color, others

    def backward_f64(self, dL_dpix, dL_dothers, vertex_abs=False, **params):
        """vertex_abs: dverts holds per entry the sum over the blended hits that touch it of each
        hit's largest |vertex-gradient component| (the scale FP32 evaluation error is relative to)."""
        v, o, g, h = self.params64(**params)
        gp = np.ascontiguousarray(dL_dpix, np.float64).reshape(3, self.R)
        go = np.ascontiguousarray(dL_dothers, np.float64).reshape(7, self.R)
        dv = np.zeros_like(v)
        dsig = np.zeros_like(g)
        dop = np.zeros_like(o)
        dsh = np.zeros_like(h)
        lib().or_set_abs_vertex(1 if vertex_abs else 0)
        try:
            rc = lib().or_backward_f64(self.h, _d(v), _d(o), _d(g), _d(h), _d(gp), _d(go), _d(dv), _d(dsig), _d(dop),
                                       _d(dsh))
        finally:
            lib().or_set_abs_vertex(0)
        if rc != 0:
            raise RuntimeError("f64 backward is defined for the forward variant only")
        return dict(dverts=dv, dsigma=dsig, dopac=dop, dsh=dsh)

    def _b64(self, refl, env):
        return (np.ascontiguousarray(refl, np.float64).reshape(-1),
                np.ascontiguousarray(env, np.float64).reshape(-1))

    def forward_f64_bounce(self, refl, env, **params):
        v, o, g, h = self.params64(**params)
        r, e = self._b64(refl, env)
        color = np.zeros((3, self.R), np.float64)
        others = np.zeros((7, self.R), np.float64)
        if lib().or_forward_f64_bounce(self.h, _d(v), _d(o), _d(g), _d(h), _d(r), _d(e), _d(color), _d(others)):
            raise RuntimeError("the log is not of the bounce variant")
        return color, others

    def loss_f64_bounce(self, refl, env, dL_dpix, dL_dothers, **params):
        v, o, g, h = self.params64(**params)
        r, e = self._b64(refl, env)
        gp = np.ascontiguousarray(dL_dpix, np.float64).reshape(3, self.R)
        go = np.ascontiguousarray(dL_dothers, np.float64).reshape(7, self.R)
        return float(lib().or_loss_f64_bounce(self.h, _d(v), _d(o), _d(g), _d(h), _d(r), _d(e), _d(gp), _d(go)))

    def backward_f64_bounce(self, refl, env, dL_dpix, dL_dothers, vertex_abs=False, **params):
        v, o, g, h = self.params64(**params)
        r, e = self._b64(refl, env)
        gp = np.ascontiguousarray(dL_dpix, np.float64).reshape(3, self.R)
        go = np.ascontiguousarray(dL_dothers, np.float64).reshape(7, self.R)
        dv, dsig, dop, dsh = np.zeros_like(v), np.zeros_like(g), np.zeros_like(o), np.zeros_like(h)
        dr, de = np.zeros_like(r), np.zeros_like(e)
        lib().or_set_abs_vertex(1 if vertex_abs else 0)
        try:
            rc = lib().or_backward_f64_bounce(self.h, _d(v), _d(o), _d(g), _d(h), _d(r), _d(e), _d(gp), _d(go), _d(dv),
                                              _d(dsig), _d(dop), _d(dsh), _d(dr), _d(de))
        finally:
            lib().or_set_abs_vertex(0)
        if rc:
            raise RuntimeError("the log is not of the bounce variant")
        return dict(dverts=dv, dsigma=dsig, dopac=dop, dsh=dsh, drefl=dr, denv=de.reshape(np.shape(env)))

    def loss_f64(self, dL_dpix, dL_dothers, **params):
        v, o, g, h = self.params64(**params)
        gp = np.ascontiguousarray(dL_dpix, np.float64).reshape(3, self.R)
        go = np.ascontiguousarray(dL_dothers, np.float64).reshape(7, self.R)
        return float(lib().or_loss_f64(self.h, _d(v), _d(o), _d(g), _d(h), _d(gp), _d(go)))
