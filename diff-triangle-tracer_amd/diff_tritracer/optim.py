"""FusedAdam: a torch.optim.Optimizer with torch.optim.Adam's update (no weight decay, no
amsgrad) that runs every parameter group of a step in one launch of the library's Adam
kernel (tt_adam_step). For the training loop around the tracer: the reference's callers
build torch.optim.Adam over the splat tensors its autograd Function differentiates
(diff_tritracer/__init__.py:205-227); this class takes the same constructor arguments.
Parameters must be contiguous float32 tensors on the GPU; up to 8 groups per launch
(more are split over launches)."""
from __future__ import annotations

import torch

from . import capi


class FusedAdam(torch.optim.Optimizer):
    def __init__(self, params, lr=1e-3, betas=(0.9, 0.999), eps=1e-8, zero_grads=False):
        if not lr >= 0.0:
            raise ValueError(f"invalid learning rate {lr}")
        if not (0.0 <= betas[0] < 1.0 and 0.0 <= betas[1] < 1.0):
            raise ValueError(f"invalid betas {betas}")
        if not eps >= 0.0:
            raise ValueError(f"invalid eps {eps}")
        super().__init__(params, dict(lr=lr, betas=betas, eps=eps))
        self.zero_grads = zero_grads  # zero each gradient after using it (as optimizer.zero_grad(set_to_none=False))

    @torch.no_grad()
    def step(self, closure=None):
        loss = None
        if closure is not None:
            with torch.enable_grad():
                loss = closure()
        # one launch per (betas, eps, step) class of parameter tensors, at most 8 tensors per launch
        batches = {}
        for group in self.param_groups:
            b1, b2 = group["betas"]
            for p in group["params"]:
                if p.grad is None:
                    continue
                if p.dtype != torch.float32 or not p.is_cuda or not p.is_contiguous() or not p.grad.is_contiguous():
                    raise RuntimeError("FusedAdam needs contiguous float32 GPU parameters and gradients")
                st = self.state[p]
                if not st:
                    st["step"] = 0
                    st["exp_avg"] = torch.zeros_like(p)
                    st["exp_avg_sq"] = torch.zeros_like(p)
                st["step"] += 1
                key = (float(b1), float(b2), float(group["eps"]), st["step"], p.device)
                batches.setdefault(key, []).append((p, p.grad, st["exp_avg"], st["exp_avg_sq"], float(group["lr"])))
        for (b1, b2, eps, step, dev), items in batches.items():
            stream = torch.cuda.current_stream(dev).cuda_stream
            for i in range(0, len(items), 8):
                capi.adam_step(items[i:i + 8], step, b1, b2, eps, zero_grads=self.zero_grads, stream=stream)
        return loss
