"""Pruning / densification on the GPU (tt_densify_plan, tt_densify_apply through the C ABI)
against the oracle (oracle/densify.py), and end to end through the tracer.

Bars: bit-exact (integer work, float copies and exact midpoints):
  - plan: action and offset arrays equal the oracle's, at tile edges (2047..2049) and at
    1M triangles, with and without the density / score gates, with a loose, a binding and
    a below-the-kept-count cap, and with caller actions;
  - apply: every remapped array equals the oracle's (vertex rows through a shuffled cumsum,
    subdivide / copy / zero policies, widths 1, 7, 9, 48), plus npts and cumsum;
  - end to end: the triangles no ray consumed (absent from every ray's hit log) are pruned
    and the scene is rebuilt: the exact-mode render is bit-identical; prune_and_split
    rewrites the parameters and the FusedAdam moments as the oracle does.
"""
import numpy as np
import pytest

import densify as OD
from diff_tritracer import capi, scenes
from helpers import GpuScene, to_dev

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-densify]", *a, flush=True)


def _stats(P, seed):
    rng = np.random.default_rng(seed)
    mb = (rng.random(P) * (rng.random(P) < 0.8)).astype(np.float32)
    op = rng.random(P).astype(np.float32)
    sc = (rng.random(P) * 0.2).astype(np.float32)
    dn = (rng.random(P) < 0.9).astype(np.float32)
    score = rng.random(P).astype(np.float32)
    act = rng.integers(0, 4, P).astype(np.uint8)
    return mb, op, sc, dn, score, act


@pytest.mark.parametrize("P", [0, 1, 2047, 2048, 2049, 100_003, 1_000_000])
@pytest.mark.parametrize("mode", ["basic", "gated", "cap", "tight", "actions"])
def test_plan_matches_oracle(cuda, P, mode):
    import torch

    mb, op, sc, dn, score, act = _stats(P, 11 + P)
    kw = dict(min_blending=0.05, min_opacity=0.1, split_scaling=0.12)
    args = dict(max_blending=mb, opacity=op, scaling=sc)
    if mode == "gated":
        args.update(density=dn, score=score)
        kw.update(split_score=0.3)
    if mode in ("cap", "tight", "actions"):
        _, off0 = OD.plan(**args, **kw)
        keep = int(((mb >= 0.05) & (op >= 0.1)).sum())
        kw["max_out"] = keep + (int(off0[-1]) - keep) // 2 + 1 if mode == "cap" else max(1, keep // 2)
    if mode == "actions":
        args = dict(action_in=act)
        n_keep, n_split = int(((act == 1) | (act == 2)).sum()), int((act == 2).sum())
        kw = dict(max_out=n_keep + 3 * (n_split // 2) + 2)
    _say(f"plan P={P} mode={mode} {kw}")
    ref_a, ref_off = OD.plan(**args, **kw)
    dev = {k: to_dev(v, cuda) for k, v in args.items()}
    a, off = capi.densify_plan(dev.get("max_blending"), dev.get("opacity"), dev.get("density"), dev.get("scaling"),
                               dev.get("score"), dev.get("action_in"), stream=torch.cuda.current_stream().cuda_stream,
                               **kw)
    torch.cuda.synchronize()
    assert np.array_equal(off.cpu().numpy(), ref_off)
    assert np.array_equal(a.cpu().numpy(), ref_a)
    if P >= 2047:
        assert (ref_a == 2).any() == (mode != "tight") and (ref_a == 0).any()


def _apply_case(P, seed):
    rng = np.random.default_rng(seed)
    verts = rng.standard_normal((3 * P + 6, 3)).astype(np.float32)  # a few unused rows
    cumsum = (3 * rng.permutation(P)).astype(np.int32)
    arrays = [(verts, 9, OD.ROWS_VERTEX, OD.SPLIT_SUBDIVIDE), (verts, 9, OD.ROWS_VERTEX, OD.SPLIT_ZERO),
              (verts, 9, OD.ROWS_VERTEX, OD.SPLIT_COPY),
              (rng.random((P, 1)).astype(np.float32), 1, OD.ROWS_TRIANGLE, OD.SPLIT_COPY),
              (rng.standard_normal((P, 16, 3)).astype(np.float32), 48, OD.ROWS_TRIANGLE, OD.SPLIT_COPY),
              (rng.standard_normal((P, 16, 3)).astype(np.float32), 48, OD.ROWS_TRIANGLE, OD.SPLIT_ZERO),
              (rng.standard_normal((P, 7)).astype(np.float32), 7, OD.ROWS_TRIANGLE, OD.SPLIT_ZERO),
              (rng.random((P, 1)).astype(np.float32), 1, OD.ROWS_TRIANGLE, OD.SPLIT_ZERO)]
    act = rng.choice(np.uint8([0, 1, 2]), P, p=[0.3, 0.5, 0.2])
    return verts, cumsum, arrays, act


@pytest.mark.parametrize("P", [1, 255, 256, 257, 5000, 100_003])
def test_apply_matches_oracle(cuda, P):
    import torch

    verts, cumsum, arrays, act = _apply_case(P, 3 + P)
    _say(f"apply P={P}")
    _, off = OD.plan(action_in=act)
    ref, ref_npts, ref_cs = OD.apply(off, cumsum, arrays)
    P_out = int(off[-1])
    doff, dcs = to_dev(off, cuda), to_dev(cumsum, cuda)
    srcs = [to_dev(a[0], cuda) for a in arrays]
    outs = [torch.full((P_out * a[1],), -3.0, device=cuda) for a in arrays]
    npts = torch.full((P_out,), -1, dtype=torch.int32, device=cuda)
    cs = torch.full((P_out,), -1, dtype=torch.int32, device=cuda)
    capi.densify_apply(P, P_out, doff, dcs, [(s, o, a[1], a[2], a[3]) for s, o, a in zip(srcs, outs, arrays)], npts,
                       cs, stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    for k, (o, r) in enumerate(zip(outs, ref)):
        assert np.array_equal(o.cpu().numpy().view(np.uint32), r.view(np.uint32)), k
    assert np.array_equal(npts.cpu().numpy(), ref_npts) and np.array_equal(cs.cpu().numpy(), ref_cs)


def test_plan_then_apply_on_device_offsets(cuda):
    """The GPU plan's offsets feed the GPU apply (no host round trip but the count)."""
    import torch

    P = 30_000
    mb, op, sc, dn, score, _ = _stats(P, 5)
    verts, cumsum, arrays, _ = _apply_case(P, 6)
    kw = dict(min_blending=0.05, min_opacity=0.1, split_scaling=0.12)
    _, ref_off = OD.plan(mb, op, dn, sc, **kw)
    ref, _, _ = OD.apply(ref_off, cumsum, arrays[:1])
    a, off = capi.densify_plan(to_dev(mb, cuda), to_dev(op, cuda), to_dev(dn, cuda), to_dev(sc, cuda), **kw,
                               stream=torch.cuda.current_stream().cuda_stream)
    P_out = int(off[-1].item())
    out = torch.empty(P_out * 9, device=cuda)
    capi.densify_apply(P, P_out, off, to_dev(cumsum, cuda),
                       [(to_dev(verts, cuda), out, 9, capi.ROWS_VERTEX, capi.SPLIT_SUBDIVIDE)],
                       stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert P_out == int(ref_off[-1]) and np.array_equal(out.cpu().numpy().view(np.uint32), ref[0].view(np.uint32))


def test_pruning_unconsumed_triangles_keeps_the_render(cuda):
    """Triangles that no ray consumed (near-skipped, alpha-skipped, blended or terminating
    hits all count as consumed) cannot change any ray's result in exact chunk mode: pruning
    them and rebuilding gives a bit-identical image, normals and alpha."""
    import torch

    scene, ro, rd = scenes.small_config(0, 96, 96)
    g = GpuScene(scene, cuda)
    cap = 400
    a = g.forward(ro, rd, k=16, cutoff=1e-3, chunk_mode="exact", log_cap=cap)
    prim, _, _, count = a["log"]
    assert count.max() <= cap
    used = np.zeros(scene.P, bool)
    used[prim[prim >= 0]] = True
    _say(f"prune: {scene.P - used.sum()} of {scene.P} triangles never consumed")
    assert 0 < used.sum() < scene.P
    verts = to_dev(scene.verts, cuda)
    tri = [to_dev(scene.opac, cuda), to_dev(scene.sigma, cuda), to_dev(scene.shs, cuda)]
    from diff_tritracer import densify

    act = torch.from_numpy(used.astype(np.uint8)).to(cuda)
    action, npts, cs = densify.prune_and_split(verts, to_dev(scene.cumsum, cuda), tri, None, None, action_in=act)
    torch.cuda.synchronize()
    assert np.array_equal(action.cpu().numpy(), used.astype(np.uint8))
    small = scenes.SplatScene(verts.cpu().numpy(), npts.cpu().numpy(), cs.cpu().numpy(), tri[2].cpu().numpy(),
                              tri[0].cpu().numpy(), tri[1].cpu().numpy())
    assert small.P == used.sum()
    b = GpuScene(small, cuda).forward(ro, rd, k=16, cutoff=1e-3, chunk_mode="exact")
    assert np.array_equal(a["color"].view(np.uint32), b["color"].view(np.uint32))
    assert np.array_equal(a["others"].view(np.uint32), b["others"].view(np.uint32))
    assert np.array_equal(a["max_blending"][used].view(np.uint32), b["max_blending"].view(np.uint32))


def test_prune_and_split_rewrites_params_and_adam_moments(cuda):
    import torch
    from diff_tritracer import FusedAdam, densify

    rng = np.random.default_rng(9)
    P = 3000
    verts0, cumsum, _, _ = _apply_case(P, 10)
    verts = torch.nn.Parameter(to_dev(verts0, cuda))
    opac = torch.nn.Parameter(to_dev(rng.random((P, 1)).astype(np.float32), cuda))
    shs = torch.nn.Parameter(to_dev(rng.standard_normal((P, 16, 3)).astype(np.float32), cuda))
    opt = FusedAdam([verts, opac, shs], lr=1e-2)
    for _ in range(2):
        for p in (verts, opac, shs):
            p.grad = torch.randn_like(p)
        opt.step()
    old = {n: t.detach().cpu().numpy() for n, t in (("v", verts), ("o", opac), ("s", shs))}
    mom = {n: [opt.state[p][k].cpu().numpy() for k in ("exp_avg", "exp_avg_sq")]
           for n, p in (("v", verts), ("o", opac), ("s", shs))}
    mb, op, sc, dn, _, _ = _stats(P, 12)
    kw = dict(min_blending=0.05, min_opacity=0.1, split_scaling=0.12, max_out=P + 600)
    _say("prune_and_split with FusedAdam state")
    action, npts, cs = densify.prune_and_split(verts, to_dev(cumsum, cuda), [opac, shs], to_dev(mb, cuda),
                                               to_dev(op, cuda), to_dev(dn, cuda), to_dev(sc, cuda), optimizer=opt,
                                               **kw)
    torch.cuda.synchronize()
    ref_a, off = OD.plan(mb, op, dn, sc, **kw)
    assert np.array_equal(action.cpu().numpy(), ref_a) and (ref_a == 2).sum() >= 200
    rv = OD.apply(off, cumsum, [(old["v"], 9, OD.ROWS_VERTEX, OD.SPLIT_SUBDIVIDE)] +
                  [(m, 9, OD.ROWS_VERTEX, OD.SPLIT_ZERO) for m in mom["v"]])[0]
    ro_ = OD.apply(off, cumsum, [(old["o"], 1, OD.ROWS_TRIANGLE, OD.SPLIT_COPY)] +
                   [(m, 1, OD.ROWS_TRIANGLE, OD.SPLIT_ZERO) for m in mom["o"]])[0]
    rs = OD.apply(off, cumsum, [(old["s"], 48, OD.ROWS_TRIANGLE, OD.SPLIT_COPY)] +
                  [(m, 48, OD.ROWS_TRIANGLE, OD.SPLIT_ZERO) for m in mom["s"]])[0]
    P_out = int(off[-1])
    assert verts.shape == (3 * P_out, 3) and opac.shape == (P_out, 1) and shs.shape == (P_out, 16, 3)
    for p, r in ((verts, rv), (opac, ro_), (shs, rs)):
        got = [p.detach()] + [opt.state[p][k] for k in ("exp_avg", "exp_avg_sq")]
        for x, y in zip(got, r):
            assert np.array_equal(x.cpu().numpy().reshape(-1).view(np.uint32), y.view(np.uint32))
        assert p.grad is None
    # the optimizer keeps working on the new shapes
    for p in (verts, opac, shs):
        p.grad = torch.randn_like(p)
    opt.step()
    torch.cuda.synchronize()
    assert torch.isfinite(verts).all()


def test_apply_unaligned_rows_take_the_scalar_path(cuda):
    """A width-48 array whose buffers are not 16-byte aligned (views at a 1-float offset)
    equals the oracle as well (the float4 path is only taken for aligned buffers)."""
    import torch

    P = 777
    verts, cumsum, arrays, act = _apply_case(P, 21)
    sh = arrays[4][0]
    _, off = OD.plan(action_in=act)
    (ref,), _, _ = OD.apply(off, cumsum, [(sh, 48, OD.ROWS_TRIANGLE, OD.SPLIT_COPY)])
    P_out = int(off[-1])
    src = torch.zeros(sh.size + 1, device=cuda)
    src[1:] = to_dev(sh.reshape(-1), cuda)
    dst = torch.full((P_out * 48 + 1,), -3.0, device=cuda)
    _say(f"apply unaligned P={P}")
    capi.densify_apply(P, P_out, to_dev(off, cuda), to_dev(cumsum, cuda),
                       [(src[1:], dst[1:], 48, capi.ROWS_TRIANGLE, capi.SPLIT_COPY)],
                       stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert np.array_equal(dst[1:].cpu().numpy().view(np.uint32), ref.view(np.uint32))
    assert dst[0].item() == -3.0


def test_prune_and_split_keeps_a_flat_vertex_tensor_flat(cuda):
    """The reference passes triangles_points flat ([9 P]); prune_and_split keeps that layout
    and writes the same values as for a [3 P, 3] tensor."""
    import torch
    from diff_tritracer import densify

    P = 2000
    verts0, cumsum, _, act = _apply_case(P, 31)
    rows = to_dev(verts0, cuda)
    flat = to_dev(verts0.reshape(-1), cuda)
    opac = [to_dev(np.ones((P, 1), np.float32), cuda) for _ in range(2)]
    a = torch.from_numpy(act).to(cuda)
    _say("prune_and_split, flat vs [V, 3] vertices")
    densify.prune_and_split(rows, to_dev(cumsum, cuda), [opac[0]], None, None, action_in=a)
    densify.prune_and_split(flat, to_dev(cumsum, cuda), [opac[1]], None, None, action_in=a)
    torch.cuda.synchronize()
    assert flat.dim() == 1 and rows.dim() == 2 and flat.numel() == rows.numel()
    assert np.array_equal(flat.cpu().numpy(), rows.cpu().numpy().reshape(-1))
