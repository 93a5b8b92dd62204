"""torch.autograd wrapper around plfft plans.

Every transform of the library is (real-)linear, so its backward is the
adjoint map, executed by the plan's own adjoint kernels
(plfft_plan_execute_adjoint_async); nothing is saved for backward but the
plan. PyTorch's complex gradient convention (dL/dRe + i dL/dIm) makes the
adjoint with respect to the real inner product sum Re(a conj b) exactly the
vector-Jacobian product.

    plan = Plan("r2c", 4096, dtype="f32", howmany=B, istride=2, idist=2*4096+64)
    X = PlanFunction.apply(x_flat, plan)   # x_flat: the plan's input buffer
    loss = (w * X.abs() ** 2).sum(); loss.backward()
"""
from __future__ import annotations

import torch

from . import Plan

_REAL = {"f32": torch.float32, "f64": torch.float64}
_CPLX = {"f32": torch.complex64, "f64": torch.complex128}


def _span(count, stride, howmany, dist):
    return (howmany - 1) * dist + (count - 1) * stride + 1


def _side_dtype(plan: Plan, side: str):
    real_side = (plan.kind == "r2r") or (plan.kind == "r2c" and side == "in") or \
        (plan.kind == "c2r" and side == "out")
    if plan.dtype not in _REAL:
        raise TypeError("the autograd wrapper supports f32 and f64 plans")
    return _REAL[plan.dtype] if real_side else _CPLX[plan.dtype]


def output_size(plan: Plan) -> int:
    L = plan.layout
    return _span(plan.out_count, L.ostride, L.howmany, L.odist)


def input_size(plan: Plan) -> int:
    L = plan.layout
    return _span(plan.in_count, L.istride, L.howmany, L.idist)


class PlanFunction(torch.autograd.Function):
    """y = plan(x) on flat device buffers laid out as the plan describes."""

    @staticmethod
    def forward(ctx, x: torch.Tensor, plan: Plan) -> torch.Tensor:
        if x.dtype != _side_dtype(plan, "in") or x.numel() < input_size(plan):
            raise ValueError("input buffer does not match the plan")
        x = x.contiguous()
        y = torch.zeros(output_size(plan), dtype=_side_dtype(plan, "out"), device=x.device)
        plan.execute_async(x, y, stream=torch.cuda.current_stream())
        ctx.plan = plan
        ctx.in_numel = x.numel()
        return y

    @staticmethod
    def backward(ctx, gy: torch.Tensor):
        plan = ctx.plan
        gy = gy.contiguous()
        gx = torch.zeros(ctx.in_numel, dtype=_side_dtype(plan, "in"), device=gy.device)
        plan.adjoint_async(gy, gx, stream=torch.cuda.current_stream())
        return gx, None


def apply(plan: Plan, x: torch.Tensor) -> torch.Tensor:
    return PlanFunction.apply(x, plan)
