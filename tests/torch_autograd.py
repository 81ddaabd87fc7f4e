"""torch.autograd wrapper of the engine's adjoint for the test harness.

``network_loglikelihood(ann, brlen, probs)`` returns the network lnL as a
0-d float64 tensor whose backward pass is the engine's one-pass adjoint
(nrx_gradient): d lnL / d branch length for every edge and d lnL / d p_r
for every reticulation's first-parent probability.  Branch lengths below
or above the engine's clamp range are clamped inside the engine, as
RaxmlWrapper.cpp:652-663 clamps them, so the gradient there is the one of
the clamped length.
"""
import numpy as np
import torch


class NetworkLogLikelihood(torch.autograd.Function):
    @staticmethod
    def forward(ctx, brlen, probs, ann):
        ann.set_branch_lengths(brlen.detach().cpu().numpy().astype(np.float64))
        if probs.numel():
            ann.set_reticulation_probs(probs.detach().cpu().numpy().astype(np.float64))
        lnl, g, gp = ann.gradient()
        ctx.save_for_backward(torch.as_tensor(g, dtype=torch.float64),
                              torch.as_tensor(np.asarray(gp), dtype=torch.float64))
        return torch.tensor(lnl, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_out):
        g, gp = ctx.saved_tensors
        return grad_out * g, grad_out * gp, None


def network_loglikelihood(ann, brlen: torch.Tensor, probs: torch.Tensor) -> torch.Tensor:
    return NetworkLogLikelihood.apply(brlen, probs, ann)
