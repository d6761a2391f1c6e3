"""torch.autograd binding of the forward modelling and its hand-written adjoint.

`SurveyModeling.apply(vp, eps, delta, plan)` returns the survey gathers (the reference's shot2 records,
src/parallelcalls.cu:113-116, for every shot) of the model (vp, epsilon, delta).  Its backward runs the
exact discrete adjoint of the library (fwi_gradient_src: forward with checkpoints, adjoint driven by
dL/dgathers, imaging) and the chain rule to (vp, epsilon, delta) (fwi_chain_rule).  Any torch loss of the
gathers therefore back-propagates through the HIP kernels; the misfit 1/2||syn - obs||^2 reproduces
fwi_gradient exactly.
"""
from __future__ import annotations

import torch

from .survey import Plan, chain_rule, pre_variables


class SurveyModeling(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vp: torch.Tensor, eps: torch.Tensor, delta: torch.Tensor, plan: Plan):
        if plan.grad_mode != 1:
            raise RuntimeError("autograd needs the exact adjoint: build the Plan with grad_mode=GRAD_EXACT")
        model = pre_variables(vp, eps, delta)
        syn = plan.forward(model)
        ctx.save_for_backward(vp, eps, delta)
        ctx.plan = plan
        return syn

    @staticmethod
    def backward(ctx, gsyn: torch.Tensor):
        vp, eps, delta = ctx.saved_tensors
        plan: Plan = ctx.plan
        model = pre_variables(vp, eps, delta)
        g_sq, _ = plan.gradient_src(model, gsyn.contiguous().float())
        g = chain_rule(vp, eps, delta, g_sq)
        return g[0].to(vp.dtype), g[1].to(eps.dtype), g[2].to(delta.dtype), None


def modeling(vp, eps, delta, plan: Plan):
    """Gathers of (vp, epsilon, delta) with the exact adjoint as backward."""
    return SurveyModeling.apply(vp, eps, delta, plan)


def misfit(vp, eps, delta, plan: Plan, obs: torch.Tensor):
    """1/2 ||gathers - obs||^2 (the reference's misfit, main.c:448-454), differentiable."""
    r = modeling(vp, eps, delta, plan) - obs
    return 0.5 * torch.sum(r.double() * r.double())
