"""Differentiable SPECT projector for torch tensors on the GPU.

    from niftyrec_amd.api import Plan
    from niftyrec_amd.torch_ops import project, poisson_nll

    plan = Plan(activity.shape, cameras, psf)
    p = project(activity, attenuation, plan)      # (V, ny, nx)
    loss = poisson_nll(p, counts)
    loss.backward()                               # activity.grad, attenuation.grad

Backward runs two hand-written kernels: the exact adjoint backprojection
(transpose of the forward trilinear pull, PSF correlation) for d/d activity,
and the exact attenuation gradient (reverse depth sweep) for d/d mu.  Both
recompute the rotated volumes per view instead of storing them.  The forward
does not clamp negative projections (the reference's truncation has no
derivative); for non-negative activity nothing is clamped anyway.
"""
import torch

from .api import ADJOINT


class _Project(torch.autograd.Function):
    @staticmethod
    def forward(ctx, activity, attenuation, plan):
        act = activity.detach().contiguous()
        mu = attenuation.detach().contiguous() if attenuation is not None else None
        out = plan.project(act, mu, truncate=False)
        ctx.plan = plan
        ctx.has_mu = mu is not None
        ctx.save_for_backward(act, mu if mu is not None else act)
        return out

    @staticmethod
    def backward(ctx, grad_out):
        act, mu = ctx.saved_tensors
        if not ctx.has_mu:
            mu = None
        g = grad_out.contiguous()
        plan = ctx.plan
        grad_act = grad_mu = None
        if ctx.has_mu and ctx.needs_input_grad[0] and ctx.needs_input_grad[1]:
            # both in one pass (after this projection, the plan reuses its depth-chunk sums)
            grad_act, grad_mu = plan.project_backward(g, act, mu, truncate=False)
            return grad_act, grad_mu, None
        if ctx.needs_input_grad[0]:
            grad_act = plan.backproject(g, mu, mode=ADJOINT, truncate=False)
        if ctx.has_mu and ctx.needs_input_grad[1]:
            grad_mu = plan.gradient_attenuation(g, act, mu, exact=True, truncate=False)
        return grad_act, grad_mu, None


def project(activity, attenuation, plan):
    """Attenuated, PSF-blurred projection (V, ny, nx), differentiable in
    activity and attenuation."""
    return _Project.apply(activity, attenuation, plan)


def poisson_nll(projection, counts, eps=1e-8):
    """sum(p - y log p) (the Poisson negative log-likelihood up to a constant)."""
    return (projection - counts * torch.log(projection.clamp_min(eps))).sum()
