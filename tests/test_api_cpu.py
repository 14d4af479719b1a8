"""CPU-side checks of the drop-in boundary (no device work):
  - libtritracer_hip.so loads and exports every entry point include/tritracer_hip.h declares
  - the C ABI rejects bad arguments with status codes and messages before touching a device
  - diff_tritracer._C binds the reference's names (ext.cpp:11-27)
  - the autograd wiring of _Tracer / Tracer follows the reference's __init__.py:54-283
"""
import ctypes as C
import os
import re

import numpy as np
import pytest
import torch

import diff_tritracer as D
from diff_tritracer import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "tritracer_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|float|const char \*)\s*\*?\s*(tt_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    names = _declared()
    assert set(names) == set(capi.EXPORTED)
    lib = C.CDLL(capi.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
    assert b"gfx950" in capi.lib().tt_version()


def test_c_abi_argument_errors_without_device():
    L = capi.lib()
    assert L.tt_build(None, None, 0, None, 0, 0, None) == 1  # TT_ERR_INVALID_ARG
    assert "NULL" in capi.last_error()
    a = capi.FwdArgs()
    a.k = 16
    assert L.tt_trace_fwd(None, C.byref(a), None) == 1
    b = capi.BwdArgs()
    assert L.tt_trace_bwd(None, C.byref(b), None) == 1
    assert L.tt_create(None) == 1
    assert L.tt_destroy(None) == 0
    assert L.tt_num_primitives(None) == -1


def test_densify_argument_errors_without_device():
    L = capi.lib()
    assert L.tt_densify_plan(None, None) == 1
    p = capi.DensifyPlanArgs()
    p.P = 10
    assert L.tt_densify_plan(C.byref(p), None) == 1 and "required" in capi.last_error()
    p.P = -1
    assert L.tt_densify_plan(C.byref(p), None) == 1 and "non-negative" in capi.last_error()
    fake = 256  # never dereferenced: every check below fails before a launch
    checks = [
        (dict(width=0), "width"),
        (dict(width=9, rows=2), "layout"),
        (dict(width=9, rows=0, on_split=3), "policy"),
        (dict(width=8, rows=1, on_split=0), "vertex rows"),
        (dict(width=9, rows=0, on_split=2), "subdivision"),
    ]
    for fields, msg in checks:
        a = capi.DensifyApplyArgs()
        a.P, a.P_out, a.offset_dev, a.cumsum_dev, a.n_arrays = 4, 4, fake, fake, 1
        arr = capi.DensifyArray(fake, fake, 9, 0, 0)
        for k, v in fields.items():
            setattr(arr, k, v)
        a.arrays[0] = arr
        assert L.tt_densify_apply(C.byref(a), None) == 1, fields
        assert msg in capi.last_error(), (fields, capi.last_error())
    a = capi.DensifyApplyArgs()
    a.P, a.P_out, a.n_arrays = 4, 17, 0
    assert L.tt_densify_apply(C.byref(a), None) == 1 and "4 P" in capi.last_error()
    a.P_out, a.n_arrays = 4, 17
    assert L.tt_densify_apply(C.byref(a), None) == 1 and "n_arrays" in capi.last_error()
    with pytest.raises(ValueError):
        capi.densify_apply_args(4, 4, None, None, [(None, None, 1, 0, 0)] * 17)


def test_scene_io_argument_errors_without_device():
    L = capi.lib()
    assert L.tt_scene_unpack(None, None) == 1 and L.tt_scene_pack(None, None) == 1
    fake = 256  # never dereferenced: every check below fails before a launch
    base = dict(P=4, V=12, face_dev=fake, face_stride=213, count_offset=0, index_offset=1, opacity_offset=13,
                sigma_offset=17, M=16, vert_dev=fake, vert_stride=12, vertices_out_dev=fake, opacity_out_dev=fake,
                sigma_out_dev=fake, shs_out_dev=fake)
    cases = [(dict(M=17), "M must"), (dict(face_stride=12), "face_stride"), (dict(face_stride=40000), "face_stride"), (dict(index_offset=205), "outside"),
             (dict(opacity_offset=210), "opacity"), (dict(face_dev=fake + 4), "aligned"),
             (dict(vert_stride=8), "x, y, z"), (dict(shs_out_dev=None), "NULL")]
    for fields, msg in cases:
        a = capi.SceneUnpackArgs()
        for k, v in {**base, **fields}.items():
            setattr(a, k, v)
        for k in range(48):
            a.sh_offset[k] = -1
        a.xyz_offset[0], a.xyz_offset[1], a.xyz_offset[2] = 0, 4, 8
        assert L.tt_scene_unpack(C.byref(a), None) == 1, fields
        assert msg in capi.last_error(), (fields, capi.last_error())
    a = capi.ScenePackArgs(4, 0, fake, fake, fake, fake, fake, fake, fake)
    assert L.tt_scene_pack(C.byref(a), None) == 1 and "M must" in capi.last_error()
    a = capi.ScenePackArgs(4, 16, fake, fake, fake, fake, fake, fake + 4, fake)
    assert L.tt_scene_pack(C.byref(a), None) == 1 and "aligned" in capi.last_error()


def test_extension_binds_reference_names():
    for name in ("OptixTracer", "FwdTraceResult", "BwdTraceResult"):
        assert hasattr(D._C, name)
    for m in ("build_tri_accel", "trace_forward", "trace_backward", "set_options", "get_options"):
        assert hasattr(D._C.OptixTracer, m)
    assert D.TracerSettings._fields == ("bg", "scale_modifier", "sh_degree", "debug", "rebuild_gas", "extra_params")


def test_unknown_program_names_raise():
    with pytest.raises(RuntimeError, match="Unknown forward program"):
        D._C.OptixTracer(D.PACKAGE_DIR, "bogus.ptx", "backward.ptx")
    with pytest.raises(RuntimeError, match="Unknown backward program"):
        D._C.OptixTracer(D.PACKAGE_DIR, "forward.ptx", "forward.ptx")


class _FakeResult:
    def __init__(self, data, hits=None, hit_counts=None, hit_colors=None):
        self.data = data
        self.hits, self.hit_counts, self.hit_colors = hits, hit_counts, hit_colors


class _FakeTracer:
    """Stands in for _C.OptixTracer to check the autograd plumbing on the CPU."""

    def __init__(self, fail=False):
        self.fail = fail
        self.fwd_args = None
        self.bwd_args = None

    store = True

    def set_store_hits(self, store):
        self.store = store

    def trace_forward(self, *args):
        if self.fail:
            raise RuntimeError("forced failure")
        self.fwd_args = args
        self.stored = self.store
        ray_o, P = args[0], args[6]
        H, W = ray_o.shape[:2]
        color = torch.full((3, H, W), 0.5)
        others = torch.zeros(7, H, W)
        return _FakeResult((color, others, torch.zeros(1), args[11], args[12], torch.zeros(P)),
                           torch.zeros(4, H * W, 2, dtype=torch.int32), torch.zeros(H * W, dtype=torch.int32),
                           torch.zeros(4, H * W, 3))

    def trace_backward(self, *args):
        self.bwd_args = args
        verts, sigma, P, shs, opac = args[3], args[4], args[9], args[10], args[12]
        return _FakeResult((torch.ones(verts.numel() // 3, 3), torch.full((P, 1), 2.0), torch.full_like(opac, 3.0),
                            torch.full_like(shs, 4.0)))


def _inputs(P=5, H=2, W=3):
    verts = torch.zeros(3 * P * 3, requires_grad=True)
    sigma = torch.ones(P, 1, requires_grad=True)
    opac = torch.ones(P, 1, requires_grad=True)
    shs = torch.zeros(P, 16, 3, requires_grad=True)
    st = D.TracerSettings(bg=torch.zeros(3), scale_modifier=1.0, sh_degree=3, debug=False, rebuild_gas=0,
                          extra_params=torch.zeros(0))
    return dict(ray_o=torch.zeros(H, W, 3), ray_d=torch.zeros(H, W, 3), verts=verts, sigma=sigma,
                npts=torch.full((P,), 3, dtype=torch.int32), cumsum=torch.arange(P, dtype=torch.int32) * 3, P=P,
                opac=opac, scaling=torch.zeros(P), density=torch.zeros(P), shs=shs, st=st)


def test_autograd_wiring_and_gradient_layouts():
    x = _inputs()
    fake = _FakeTracer()
    out = D._Tracer.apply(fake, x["ray_o"], x["ray_d"], x["verts"], x["sigma"], x["npts"], x["cumsum"], x["P"],
                          x["shs"], torch.empty(0), x["opac"], x["scaling"], x["density"], x["st"])
    color, gpu_time, scaling, density, others, maxb = out
    assert color.shape == (3, 2, 3) and others.shape == (7, 2, 3)
    # forward argument order of trace_forward (reference __init__.py:75-93)
    fa = fake.fwd_args
    assert fa[6] == x["P"] and fa[8] == 3 and fa[15] == 0 and fa[16] is False
    (color.sum() + others.sum()).backward()
    # backward argument order (reference __init__.py:162-180) and gradient layouts (:205-227)
    ba = fake.bwd_args
    assert ba[9] == x["P"] and ba[11] == 3 and ba[13].shape == (3, 2, 3) and ba[14].shape == (7, 2, 3)
    # the forward's stored hits and their colours travel to the backward (trailing optional arguments)
    assert len(ba) == 20 and ba[17].shape == (4, 6, 2) and ba[18].shape == (6,) and ba[19].shape == (4, 6, 3)
    assert torch.all(x["verts"].grad == 1.0) and x["verts"].grad.shape == x["verts"].shape
    assert torch.all(x["sigma"].grad == 2.0) and x["sigma"].grad.shape == (5, 1)
    assert torch.all(x["opac"].grad == 3.0) and torch.all(x["shs"].grad == 4.0)


def test_hit_list_stored_only_when_a_backward_can_need_it():
    tr = D.Tracer.__new__(D.Tracer)
    torch.nn.Module.__init__(tr)
    fake = tr.tracer = _FakeTracer()

    def call(x):
        return tr(x["ray_o"], x["ray_d"], x["verts"], x["sigma"], x["npts"], x["cumsum"], x["P"], x["opac"],
                  x["scaling"], x["density"], x["shs"], None, x["st"])

    call(_inputs())
    assert fake.stored is True
    with torch.no_grad():
        call(_inputs())
    assert fake.stored is False and fake.store is True   # restored after the call
    y = _inputs()
    for k in ("verts", "sigma", "opac", "shs"):
        y[k] = y[k].detach()
    call(y)
    assert fake.stored is False and fake.store is True


def test_tracer_requires_exactly_one_colour_source():
    x = _inputs()
    tr = D.Tracer.__new__(D.Tracer)
    torch.nn.Module.__init__(tr)
    tr.tracer = _FakeTracer()
    args = (x["ray_o"], x["ray_d"], x["verts"], x["sigma"], x["npts"], x["cumsum"], x["P"], x["opac"], x["scaling"],
            x["density"])
    with pytest.raises(Exception, match="exactly|excatly"):
        tr(*args, None, None, x["st"])
    with pytest.raises(Exception, match="exactly|excatly"):
        tr(*args, x["shs"], torch.zeros(5, 3), x["st"])
    tr(*args, x["shs"], None, x["st"])


def test_debug_snapshot_on_failure(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    x = _inputs()
    st = x["st"]._replace(debug=True)
    with pytest.raises(RuntimeError, match="forced failure"):
        D._Tracer.apply(_FakeTracer(fail=True), x["ray_o"], x["ray_d"], x["verts"], x["sigma"], x["npts"],
                        x["cumsum"], x["P"], x["shs"], torch.empty(0), x["opac"], x["scaling"], x["density"], st)
    snap = torch.load(tmp_path / "snapshot_fw.dump", weights_only=True)
    assert len(snap) == 17 and snap[6] == x["P"]


def test_scene_generators_deterministic():
    from diff_tritracer import scenes

    a, ro, rd = scenes.small_config(0, 32, 32)
    b, ro2, rd2 = scenes.small_config(0, 32, 32)
    assert np.array_equal(a.verts, b.verts) and np.array_equal(ro, ro2)
    assert a.verts.shape == (30000, 3) and a.shs.shape == (10000, 16, 3)
    np.testing.assert_allclose(np.linalg.norm(rd.reshape(-1, 3), axis=1), 1.0, atol=1e-6)
    c = scenes.clustered_scene(2000, 8, seed=1)
    assert c.P == 2000 and np.all(np.abs(c.verts) < 2.5)


def test_ctypes_structs_match_the_header(tmp_path):
    """Field offsets and sizes of the ctypes mirrors equal the C header's (gcc on the header)."""
    import shutil
    import subprocess

    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    pairs = [(capi.FwdArgs, "tt_fwd_args"), (capi.BwdArgs, "tt_bwd_args"), (capi.MultiArgs, "tt_multi_args"),
             (capi.ProposalHparams, "tt_proposal_hparams"), (capi.AdamGroup, "tt_adam_group"),
             (capi.DensifyPlanArgs, "tt_densify_plan_args"), (capi.DensifyArray, "tt_densify_array"),
             (capi.DensifyApplyArgs, "tt_densify_apply_args"), (capi.SceneUnpackArgs, "tt_scene_unpack_args"),
             (capi.ScenePackArgs, "tt_scene_pack_args")]
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "tritracer_hip.h"', "int main(void) {"]
    want = []
    for cls, cname in pairs:
        lines.append(f'printf("%zu\\n", sizeof({cname}));')
        want.append(C.sizeof(cls))
        for name, _ in cls._fields_:
            lines.append(f'printf("%zu\\n", offsetof({cname}, {name}));')
            want.append(getattr(cls, name).offset)
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    assert got == want
