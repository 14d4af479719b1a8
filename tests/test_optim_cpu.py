"""The Adam update the GPU kernel applies (k_adam in csrc/tt_optim.hip), restated in numpy
float32 with the kernel's operation order and its host-side scalars, against
torch.optim.Adam on the CPU over five steps: pins the formula (bias corrections, lerp
weight, eps placement, sign) without a GPU. Bar: 2e-6 relative + 1e-7 absolute."""
import numpy as np
import torch


def _kernel_step(p, g, m, v, lr, step, b1=0.9, b2=0.999, eps=1e-8):
    f = np.float32
    w1 = f(1.0 - b1)
    m = (m + w1 * (g - m)) if w1 < f(0.5) else (g - (g - m) * (f(1) - w1))
    v = v * f(b2) + (f(1.0 - b2) * g) * g
    ss = f(lr / (1.0 - b1 ** step))
    denom = np.sqrt(v) / f(np.sqrt(1.0 - b2 ** step)) + f(eps)
    p = p + (-ss) * (m / denom)
    return p.astype(np.float32), m.astype(np.float32), v.astype(np.float32)


def test_adam_formula_matches_torch():
    rng = np.random.Generator(np.random.PCG64(4))
    n, lr = 2000, 3e-3
    p0 = rng.standard_normal(n).astype(np.float32)
    t = torch.tensor(p0.copy(), requires_grad=True)
    opt = torch.optim.Adam([t], lr=lr)
    p, m, v = p0.copy(), np.zeros(n, np.float32), np.zeros(n, np.float32)
    for step in range(1, 6):
        g = rng.standard_normal(n).astype(np.float32)
        t.grad = torch.from_numpy(g.copy())
        opt.step()
        p, m, v = _kernel_step(p, g, m, v, lr, step)
        st = opt.state[t]
        for a, b in ((p, t.detach().numpy()), (m, st["exp_avg"].numpy()), (v, st["exp_avg_sq"].numpy())):
            assert np.all(np.abs(a.astype(np.float64) - b) <= 2e-6 * np.abs(b) + 1e-7)
