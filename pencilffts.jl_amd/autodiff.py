"""Reverse-mode differentiation through a plan.

PencilFFTs.jl has no AD rules of its own.  The Julia side of this build exposes
ChainRulesCore `rrule`s for `*`, `ldiv!`/`\\` and `bmul!`.  Their pullbacks
call the adjoint directions of the C ABI (PFFT_ADJOINT_FORWARD /
PFFT_ADJOINT_BACKWARD: the transposes of mul! and bmul! for the real inner
product Re sum conj(a) b, with the Hermitian weights of the r2c/c2r stage
fused into the kernels).  This module is the torch counterpart:
torch.autograd.Function subclasses.

torch's gradient of a real loss with respect to a complex tensor is
dL/dRe + i dL/dIm, which is exactly the transpose for that inner product.
"""
import torch

from . import _lib as L


def _run(plan, direction, src):
    """Execute `direction` on a parent tensor; returns a new parent tensor in
    the destination pencil's layout."""
    from .plans import PencilArray
    if direction in (L.FORWARD, L.ADJOINT_BACKWARD):   # input-shaped -> output-shaped
        p_in, e_in = plan.pencil_input, plan.eltype_input
        p_out, e_out = plan.pencil_output, plan.eltype_output
    else:
        p_in, e_in = plan.pencil_output, plan.eltype_output
        p_out, e_out = plan.pencil_input, plan.eltype_input
    src_a = PencilArray(p_in, e_in, src.contiguous())
    dst = PencilArray.empty(p_out, e_out, src.device)
    plan._execute(direction, src_a, dst)
    return dst.parent


class PlanForward(torch.autograd.Function):
    """v = plan * u  (mul!); pullback u_bar = (mul!)^T v_bar."""

    @staticmethod
    def forward(ctx, u, plan):
        ctx.plan = plan
        return _run(plan, L.FORWARD, u)

    @staticmethod
    def backward(ctx, v_bar):
        return _run(ctx.plan, L.ADJOINT_FORWARD, v_bar), None


class PlanBackward(torch.autograd.Function):
    """u = plan \\ v (ldiv!, scaled) or bmul! (unscaled); pullback
    v_bar = (bmul!)^T u_bar, times 1/scale_factor for ldiv!."""

    @staticmethod
    def forward(ctx, v, plan, scaled):
        ctx.plan = plan
        ctx.scaled = scaled
        return _run(plan, L.BACKWARD_SCALED if scaled else L.BACKWARD, v)

    @staticmethod
    def backward(ctx, u_bar):
        v_bar = _run(ctx.plan, L.ADJOINT_BACKWARD, u_bar)
        if ctx.scaled:
            v_bar = v_bar * (1.0 / ctx.plan.scale_factor)
        return v_bar, None, None


def forward(plan, u):
    """Differentiable plan * u on parent tensors (memory order of the input pencil)."""
    return PlanForward.apply(u, plan)


def ldiv(plan, v):
    """Differentiable plan \\ v (normalised backward) on parent tensors."""
    return PlanBackward.apply(v, plan, True)


def bmul(plan, v):
    """Differentiable unnormalised backward transform."""
    return PlanBackward.apply(v, plan, False)
