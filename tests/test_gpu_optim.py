"""The fused Adam step on the GPU against torch.optim.Adam on the CPU (the reference's
caller's optimizer), over three consecutive steps on four parameter groups of uneven
sizes (the vectorised and the scalar path), with and without gradient zeroing, and the
refit after a vertex update against a fresh build.

Bars: parameters and both moments within 2e-6 relative (+1e-7 absolute) of torch after
each step (the formulas are torch's; its CPU kernels may fuse multiply-adds); zeroed
gradients exactly 0; images after the refit bit-identical to a rebuild."""
import numpy as np
import pytest

from diff_tritracer import capi, scenes
from helpers import GpuScene, to_dev

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-optim]", *a, flush=True)


@pytest.mark.parametrize("sizes", [(4096, 1024, 1024, 48 * 1024), (3001, 37, 1000, 48 * 1000 + 3)])
@pytest.mark.parametrize("zero", [False, True])
def test_adam_matches_torch(cuda, sizes, zero):
    import torch

    rng = np.random.Generator(np.random.PCG64(17))
    lrs = (1e-3, 5e-2, 1e-2, 2.5e-3)
    cpu = [torch.tensor(rng.standard_normal(n).astype(np.float32), requires_grad=True) for n in sizes]
    opt = torch.optim.Adam([{"params": [p], "lr": lr} for p, lr in zip(cpu, lrs)], betas=(0.9, 0.999), eps=1e-8)
    dev = [p.detach().clone().to(cuda) for p in cpu]
    m = [torch.zeros_like(p) for p in dev]
    v = [torch.zeros_like(p) for p in dev]
    _say(f"sizes {sizes}, zero_grads {zero}")
    for step in range(1, 4):
        grads = [rng.standard_normal(n).astype(np.float32) * (0.1 if i == 3 else 1.0) for i, n in enumerate(sizes)]
        for p, g in zip(cpu, grads):
            p.grad = torch.from_numpy(g.copy())
        opt.step()
        gdev = [to_dev(g, cuda) for g in grads]
        capi.adam_step([(p, g, mm, vv, lr) for p, g, mm, vv, lr in zip(dev, gdev, m, v, lrs)], step, 0.9, 0.999,
                       1e-8, zero_grads=zero, stream=torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        for i, p in enumerate(cpu):
            st = opt.state[p]
            for name, a, b in (("param", dev[i], p.detach()), ("exp_avg", m[i], st["exp_avg"]),
                               ("exp_avg_sq", v[i], st["exp_avg_sq"])):
                a, b = a.cpu().numpy().astype(np.float64), b.numpy().astype(np.float64)
                err = np.abs(a - b) - (2e-6 * np.abs(b) + 1e-7)
                assert err.max() <= 0.0, (step, i, name, float(np.abs(a - b).max()))
            if zero:
                assert not gdev[i].any().item()
            else:
                assert np.array_equal(gdev[i].cpu().numpy(), grads[i])


def test_vertex_step_then_refit_equals_rebuild(cuda):
    import torch

    sc = scenes.uniform_scene(8_000, 4)
    ro, rd = scenes.look_at_rays(64, 48, scenes.orbit_eye(3.0, 4), fov_deg=45.0)
    g = GpuScene(sc, cuda)
    rng = np.random.Generator(np.random.PCG64(3))
    grad = to_dev((1e-2 * rng.standard_normal(sc.verts.shape)).astype(np.float32), cuda)
    m, v = torch.zeros_like(g.verts), torch.zeros_like(g.verts)
    s = torch.cuda.current_stream().cuda_stream
    capi.adam_step([(g.verts, grad, m, v, 1e-2)], 1, stream=s)
    g.h.build(g.verts, g.cumsum, g.P, capi.REFIT, s)
    a = g.forward(ro, rd, k=16, cutoff=1e-3, chunk_mode="exact")
    g.h.build(g.verts, g.cumsum, g.P, capi.BUILD, s)
    b = g.forward(ro, rd, k=16, cutoff=1e-3, chunk_mode="exact")
    _say("refit after an Adam step on the vertices vs rebuild")
    assert np.array_equal(a["color"].view(np.uint32), b["color"].view(np.uint32))
    assert np.array_equal(a["stats"], b["stats"])


def test_fused_adam_optimizer_matches_torch_adam(cuda):
    import torch
    from diff_tritracer import FusedAdam

    rng = np.random.Generator(np.random.PCG64(23))
    shapes = [(300, 3), (100, 1), (100, 16, 3)]
    init = [rng.standard_normal(s).astype(np.float32) for s in shapes]
    a = [torch.tensor(x, device=cuda, requires_grad=True) for x in init]
    b = [torch.tensor(x, requires_grad=True) for x in init]
    fa = FusedAdam([{"params": [a[0]], "lr": 1e-3}, {"params": a[1:], "lr": 1e-2}], betas=(0.9, 0.99), eps=1e-6)
    ta = torch.optim.Adam([{"params": [b[0]], "lr": 1e-3}, {"params": b[1:], "lr": 1e-2}], betas=(0.9, 0.99), eps=1e-6)
    _say("FusedAdam vs torch.optim.Adam, 4 steps")
    for _ in range(4):
        grads = [rng.standard_normal(s).astype(np.float32) for s in shapes]
        for p, q, g in zip(a, b, grads):
            p.grad = torch.tensor(g, device=cuda)
            q.grad = torch.tensor(g)
        fa.step()
        ta.step()
    torch.cuda.synchronize()
    for p, q in zip(a, b):
        x, y = p.detach().cpu().numpy().astype(np.float64), q.detach().numpy().astype(np.float64)
        assert np.all(np.abs(x - y) <= 2e-6 * np.abs(y) + 1e-7)
