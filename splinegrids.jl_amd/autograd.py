"""Differentiation through the grid evaluation.

The reference differentiates with Enzyme through its KernelAbstractions
kernel (test/test_EnzymeExt.jl:24-49, ext/SplineGridsEnzymeExt.jl).  The
map control_points -> eval is linear, so the rules are exact:

  reverse:  dP = M^T d(eval)   (evaluate_adjoint_, accumulate into dP)
  forward:  d(eval) = M dP     (evaluate_ on the tangent)

Basis tables, knots and NURBS weights are constants.  In reverse mode a
NURBS grid (derivative order 0) takes the rational adjoint,
evaluate_rational_adjoint_ (the reference's evaluate_adjoint! rejects NURBS,
src/adjoint.jl:57); it runs the native tier.
"""
from __future__ import annotations

import torch

from .spline_grid import SplineGrid, evaluate_, evaluate_adjoint_, evaluate_rational_adjoint_
from .utils import colmajor, colmajor_empty, colmajor_zeros


class SplineEval(torch.autograd.Function):
    """eval = M(grid, derivative_order) @ control_points, differentiable."""

    @staticmethod
    def forward(ctx, control_points: torch.Tensor, grid: SplineGrid, derivative_order,
                tier: str):
        out = colmajor_empty(grid.eval.shape, grid.dtype, grid.device)
        evaluate_(grid, derivative_order=derivative_order,
                  control_points=colmajor(control_points.detach()), eval=out, tier=tier)
        ctx.grid, ctx.order, ctx.tier = grid, derivative_order, tier
        return out

    @staticmethod
    def backward(ctx, grad_eval: torch.Tensor):
        grid = ctx.grid
        grad_cp = colmajor_zeros(tuple(grid.control_points.shape), grid.dtype, grid.device)
        if grid.is_nurbs:
            assert not any(ctx.order), "Computing derivatives of NURBS is currently not supported."
            evaluate_rational_adjoint_(grid, control_points=grad_cp, eval=colmajor(grad_eval),
                                       accumulate=True)
        else:
            evaluate_adjoint_(grid, derivative_order=ctx.order, control_points=grad_cp,
                              eval=colmajor(grad_eval), tier=ctx.tier, accumulate=True)
        return grad_cp, None, None, None

    @staticmethod
    def jvp(ctx, d_control_points: torch.Tensor, *_):
        grid = ctx.grid
        out = colmajor_empty(grid.eval.shape, grid.dtype, grid.device)
        evaluate_(grid, derivative_order=ctx.order,
                  control_points=colmajor(d_control_points), eval=out, tier=ctx.tier)
        return out


def spline_eval(grid: SplineGrid, control_points: torch.Tensor, derivative_order=None,
                tier: str = "native") -> torch.Tensor:
    """Out-of-place, differentiable evaluate!: returns a new column-major eval."""
    order = tuple(derivative_order) if derivative_order is not None else (0,) * grid.Nin
    return SplineEval.apply(control_points, grid, order, tier)
