"""torch.autograd wrapper of a DifferentiableRun.

PyTorch is used only here, for the autograd plumbing: the forward histograms
come from the fused HIP kernel and the backward pass is the hand-written
adjoint kernel (mcv_instr_adjoint) evaluated on the same random streams.
"""
import numpy as np
import torch


class _Histograms(torch.autograd.Function):

    @staticmethod
    def forward(ctx, theta, run):
        th = theta.detach().to("cpu", torch.float64).numpy().copy()
        h = run.forward(th)
        ctx.run, ctx.theta = run, th
        return torch.from_numpy(h).to(theta.device)

    @staticmethod
    def backward(ctx, g):
        grad = ctx.run.vjp(g.detach().to("cpu", torch.float64).numpy(), theta=ctx.theta)
        return torch.from_numpy(np.asarray(grad, dtype=np.float64)).to(g.device), None


def histograms(theta, run):
    """The selected monitors' sum-p histograms (concatenated) of `run` at
    `theta`, differentiable with respect to `theta` (float64 tensor)."""
    return _Histograms.apply(theta, run)
