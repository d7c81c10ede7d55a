"""LinearMap(grid): the matrix-free operator control points -> eval.

Mirrors ext/SplineGridsLinearMapsExt.jl:5-52: the forward closure copies the
flat vector into grid.control_points, evaluates the control-point container
(the THB chain for locally refined control points) and then the grid into
the output buffer (:16-31); the adjoint closure runs the grid's adjoint, the
container's adjoint and copies the control points out (:33-48).  Vectors are
flattened in Julia's column-major order.  The result is a scipy
LinearOperator (host vectors, one device round trip per product) for use
with scipy.sparse.linalg solvers such as lsqr.
"""
from __future__ import annotations

import numpy as np
import torch

from .control_points import control_point_matrix, get_n_control_points
from .spline_grid import (SplineGrid, evaluate_, evaluate_adjoint_, reshape_colmajor)
from .validation import validate_partial_derivatives


def LinearMap(grid: SplineGrid, derivative_order=None):
    from scipy.sparse.linalg import LinearOperator

    order = tuple(derivative_order) if derivative_order is not None else (0,) * grid.Nin
    validate_partial_derivatives(grid, order)
    cps = grid.control_points
    n_cp, nout = get_n_control_points(grid), grid.Nout
    shape_eval = tuple(grid.eval.shape)
    n_eval = grid.eval.numel()
    np_dtype = np.float32 if grid.dtype == torch.float32 else np.float64
    ev = torch.empty(n_eval, dtype=grid.dtype, device=grid.device)

    def matvec(x):
        x = torch.as_tensor(np.asarray(x, dtype=np_dtype).reshape(-1)).to(grid.device)
        cps.copy_(x.view(nout, n_cp).t())                     # copyto!(control_points, ...)
        evaluate_(cps)                                        # evaluate!(control_points)
        evaluate_(grid, derivative_order=order, eval=reshape_colmajor(ev, shape_eval))
        return ev.cpu().numpy().copy()

    def rmatvec(y):
        y = torch.as_tensor(np.asarray(y, dtype=np_dtype).reshape(-1)).to(grid.device)
        evaluate_adjoint_(grid, derivative_order=order, eval=reshape_colmajor(y, shape_eval))
        evaluate_adjoint_(cps)                                # evaluate_adjoint!(control_points)
        M = control_point_matrix(cps)                         # copyto!(flat, control_points)
        return M.t().reshape(-1).cpu().numpy().copy()

    return LinearOperator((n_eval, n_cp * nout), matvec=matvec, rmatvec=rmatvec,
                          dtype=np_dtype)
