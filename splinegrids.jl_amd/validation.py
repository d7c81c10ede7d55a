"""Validation with the reference's messages: src/validation.jl:1-88.

The reference logs one ``@error`` line per problem and then throws
``error(...)``; here each line goes to the ``splinegrids_amd`` logger at
ERROR level and the throw is a ValueError with the same text.
"""
from __future__ import annotations

import logging

logger = logging.getLogger("splinegrids_amd")


def _jtuple(t) -> str:
    """Julia's printing of an integer tuple."""
    t = tuple(int(x) for x in t)
    return f"({t[0]},)" if len(t) == 1 else "(" + ", ".join(str(x) for x in t) + ")"


def validate_spline_grid(spline_dimensions, control_points, weights, eval) -> None:
    """src/validation.jl:1-50."""
    from .control_points import obtain

    cp = obtain(control_points)
    errors = False
    nin_sd, nin_cp, nin_eval = len(spline_dimensions), cp.ndim - 1, eval.ndim - 1
    if weights is None:
        nins = (nin_sd, nin_cp, nin_eval)
        if len(set(nins)) != 1:
            logger.error("The number of input dimensions from the spline dimensions, "
                         f"control_points, and eval must agree, got {_jtuple(nins)} respectively.")
            errors = True
    else:
        nins = (nin_sd, nin_cp, nin_eval, weights.ndim)
        if len(set(nins)) != 1:
            logger.error("The number of input dimensions from the spline dimensions, "
                         f"control_points, eval and weights must agree, got {_jtuple(nins)} "
                         "respectively.")
            errors = True
    nout_cp, nout_eval = cp.shape[-1], eval.shape[-1]
    if nout_cp != nout_eval:
        logger.error(f"The number of output dimensions from the control points {nout_cp} "
                     f"and eval {nout_eval} must agree.")
        errors = True
    size_sd = tuple(s.n_basis_functions for s in spline_dimensions)
    size_cp = tuple(cp.shape[:-1])
    if size_sd != size_cp:
        logger.error(f"The control point grid sizes from the spline dimensions {_jtuple(size_sd)} "
                     f"and the control points {_jtuple(size_cp)} must agree.")
        errors = True
    if errors:
        raise ValueError("Errors encountered when validating spline grid constructor inputs.")


def validate_partial_derivatives(target, derivative_order) -> None:
    """src/validation.jl:52-88.  ``target`` is a grid or a tuple of
    spline dimensions."""
    from .spline_grid import SplineGrid

    derivative_order = tuple(int(o) for o in derivative_order)
    if isinstance(target, SplineGrid):
        if target.is_nurbs:  # :81-88
            if any(o != 0 for o in derivative_order):
                raise ValueError("Computing derivatives of NURBS is currently not supported.")
            return
        dims = target.spline_dimensions
    else:
        dims = tuple(target)
    assert len(dims) == len(derivative_order), \
        "derivative_order needs one entry per input dimension"
    errors = False
    for i, (s, o) in enumerate(zip(dims, derivative_order), start=1):  # :65-73
        if not (0 <= o <= s.max_derivative_order):
            errors = True
            logger.error(f"The maximum derivative order available for spline dimension {i} "
                         f"is {s.max_derivative_order}, got {o}.")
    if errors:
        raise ValueError("Invalid derivative order(s) supplied. If you want to evaluate "
                         "(higher order) derivatives, specify this at construction as "
                         "SplineDimension(...; max_derivative_order).")
