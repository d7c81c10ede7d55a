"""Knot insertion, refinement and local (THB) refinement of spline grids:
src/refinement.jl and src/control_points.jl:358-680.

Knot vectors are host data (as everywhere in this package); the refinement
matrices, the refined control points and the THB bookkeeping (overwrites,
gathers, the structural adjoint that finds overwritten control points) run
in libsplinegrids_hip.so.  Index sets that the reference forms on the CPU
(the ``unique`` of activate_local_refinement!, src/control_points.jl:
481-493) are formed on the host here too.
"""
from __future__ import annotations

import math
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib
from .control_points import (DefaultControlPoints, LocalRefinement,
                             LocallyRefinedControlPoints, base_local_refinement,
                             local_refinement_gather_, local_refinement_overwrite_, obtain)
from .knot_vector import KnotVector
from .refinement_matrix import RefinementMatrix, mult_, mult_adjoint_, rmeye
from .spline_dimension import SplineDimension
from .spline_grid import SplineGrid, evaluate_adjoint_
from .utils import colmajor, colmajor_empty, colmajor_zeros, current_stream_handle, dtype_code


# --------------------------------------------------------------------------
# knot insertion
# --------------------------------------------------------------------------

def insert_knot_vector(knot_vector: KnotVector, knot_new: float):
    """insert_knot(::KnotVector, knot_new) (src/refinement.jl:107-133): the
    knot vector with knot_new of multiplicity 1, and the 1-based span index
    k with knots_all[k] <= knot_new < knots_all[k + 1]."""
    values = np.asarray(knot_vector.knot_values)
    mults = np.asarray(knot_vector.multiplicities)
    knot_new = values.dtype.type(knot_new)
    assert not np.any(knot_new == values)
    pos = int(np.searchsorted(values, knot_new, side="left"))  # searchsortedfirst - 1
    new = KnotVector(np.insert(values, pos, knot_new), np.insert(mults, pos, 1).astype(mults.dtype))
    return new, int(mults[:pos].sum())


def _dimension_with_knots(s: SplineDimension, knot_vector: KnotVector) -> SplineDimension:
    """A dimension with the same degree, derivative orders and sample points
    on a new knot vector, its tables evaluated (set_sample_indices! and
    evaluate!, src/refinement.jl:180-183).  Unlike setproperties in the
    reference, the new dimension owns its tables: the old one is unchanged."""
    n = len(knot_vector.knots_all) - s.degree - 1
    return SplineDimension(n, s.degree, s.n_sample_points,
                           max_derivative_order=s.max_derivative_order,
                           knot_vector=knot_vector, device=s.device,
                           float_type=s._host_t.dtype.type, sample_points=s._host_t)


def _insertion_matrix(s: SplineDimension, knot_span_index: int, knot_new: float) -> RefinementMatrix:
    """RefinementMatrix(spline_dimension, knot_span_index, knot_new)
    (src/refinement.jl:53-88) on the GPU."""
    n, p = s.n_basis_functions, s.degree
    dev = s.device
    row_pointer = torch.empty(n + 1, dtype=torch.int32, device=dev)
    column_start = torch.empty(n + 1, dtype=torch.int32, device=dev)
    nzval = torch.empty(n + p + 1, dtype=s.float_type, device=dev)
    with torch.cuda.device(dev):
        _lib.check(_lib.lib().sg_refinement_insert_knot(
            s._knots_dev.data_ptr(), n, p, int(knot_span_index), float(knot_new),
            row_pointer.data_ptr(), column_start.data_ptr(), nzval.data_ptr(),
            dtype_code(s.float_type), current_stream_handle(dev)))
    return RefinementMatrix(n + 1, n, row_pointer, column_start, nzval)


def insert_knot(obj, *args, **kwargs):
    """insert_knot, as the reference dispatches it (src/refinement.jl:107-235):

    - ``insert_knot(knot_vector, knot_new)`` -> (KnotVector, knot_span_index)
    - ``insert_knot(spline_dimension, knot_new)`` -> (SplineDimension,
      RefinementMatrix), the matrix (n+1) x n expressing the old basis in the
      new one
    - ``insert_knot(spline_grid, dim_refinement, knot_new)`` -> (SplineGrid,
      RefinementMatrix), the grid's control points refined along the
      1-based dimension so that its geometry is unchanged (B-splines only).
    The reference's ``recompute_sample_indices`` / ``evaluate`` keywords are
    accepted; the new dimension is always evaluated."""
    if isinstance(obj, KnotVector):
        return insert_knot_vector(obj, *args)
    if isinstance(obj, SplineDimension):
        (knot_new,) = args
        kv, span = insert_knot_vector(obj.knot_vector, knot_new)
        R = _insertion_matrix(obj, span, kv.knot_values.dtype.type(knot_new))
        return _dimension_with_knots(obj, kv), R
    if isinstance(obj, SplineGrid):
        dim_refinement, knot_new = args
        assert not obj.is_nurbs, "Knot insertion not supported for NURBS"
        s = obj.spline_dimensions[int(dim_refinement) - 1]
        s_new, R = insert_knot(s, knot_new)
        return refine_grid(obj, s_new, int(dim_refinement), R)
    raise TypeError(f"insert_knot: unsupported {type(obj).__name__}")


def refine(obj, *args, knots_new: Optional[Sequence[float]] = None):
    """refine, as the reference dispatches it (src/refinement.jl:256-330):

    - ``refine(spline_dimension; knots_new)`` -> (SplineDimension,
      RefinementMatrix): every knot of knots_new inserted (default: the
      midpoint of every knot span), the matrix the product of the single
      insertions, newest on the left
    - ``refine(spline_grid, dim_refinement; knots_new)`` -> (SplineGrid,
      RefinementMatrix) with the grid's control points refined."""
    if isinstance(obj, SplineDimension):
        assert not args
        vals = np.asarray(obj.knot_vector.knot_values)
        if knots_new is None:
            knots_new = vals[:-1] + np.diff(vals) / vals.dtype.type(2)
        R = rmeye(obj.n_basis_functions, device=obj.device, float_type=vals.dtype.type)
        kv = obj.knot_vector
        for k in np.asarray(knots_new, dtype=vals.dtype):
            kv_new, span = insert_knot_vector(kv, k)
            R = _insertion_matrix(_KnotsOnly(obj, kv), span, k) @ R
            kv = kv_new
        return _dimension_with_knots(obj, kv), R
    if isinstance(obj, SplineGrid):
        (dim_refinement,) = args
        assert not obj.is_nurbs, "Knot insertion not supported for NURBS"
        s = obj.spline_dimensions[int(dim_refinement) - 1]
        s_new, R = refine(s, knots_new=knots_new)
        return refine_grid(obj, s_new, int(dim_refinement), R)
    raise TypeError(f"refine: unsupported {type(obj).__name__}")


class _KnotsOnly:
    """A dimension's degree with the knots of one step of refine(): an
    insertion needs only these and the expanded knots on the device."""

    def __init__(self, s, kv: KnotVector):
        self.degree = s.degree
        self.device = s.device
        self.float_type = s.float_type
        self.n_basis_functions = len(kv.knots_all) - s.degree - 1
        self._knots_dev = torch.from_numpy(
            np.asarray(kv.knots_all, dtype=np.asarray(kv.knot_values).dtype).copy()).to(s.device)


def refine_grid(grid: SplineGrid, spline_dimension_new: SplineDimension, dim_refinement: int,
                refinement_matrix: RefinementMatrix):
    """refine(spline_grid, spline_dimension_new, dim_refinement,
    refinement_matrix) (src/refinement.jl:341-386): the control points (the
    last level's array for THB control points) times the matrix along the
    dimension, in a new grid that shares eval and weights."""
    assert not grid.is_nurbs, "Knot insertion not supported for NURBS"
    cp = obtain(grid.control_points)
    shape = list(cp.shape)
    shape[dim_refinement - 1] = refinement_matrix.m
    cp_new = colmajor_empty(shape, cp.dtype, cp.device)
    mult_(cp_new, [refinement_matrix], cp, [dim_refinement])
    dims = tuple(spline_dimension_new if d == dim_refinement - 1 else s
                 for d, s in enumerate(grid.spline_dimensions))
    return SplineGrid(dims, DefaultControlPoints(cp_new), grid.weights, grid.eval), refinement_matrix


# --------------------------------------------------------------------------
# local (THB) refinement
# --------------------------------------------------------------------------

def add_default_local_refinement(grid: SplineGrid) -> SplineGrid:
    """add_default_local_refinement (src/control_points.jl:364-425): every
    dimension refined at the midpoints of its knot spans, a new THB level
    with no active control points (the first call also makes the dense
    control points the base level).  Returns the new grid."""
    dims_new, mats = [], []
    for s in grid.spline_dimensions:
        s_new, R = refine(s)
        dims_new.append(s_new)
        mats.append(R)
    cp = grid.control_points
    base = obtain(cp)
    nout = int(base.shape[-1])
    shape = tuple(s.n_basis_functions for s in dims_new) + (nout,)
    cp_new = colmajor_zeros(shape, base.dtype, base.device)
    dims_refinement = list(range(1, grid.Nin + 1))
    mult_(cp_new, mats, base, dims_refinement)
    level = LocalRefinement(dims_refinement, mats,
                            colmajor_empty((0, grid.Nin), torch.int32, base.device),
                            colmajor_empty((0, nout), base.dtype, base.device))
    if isinstance(cp, LocallyRefinedControlPoints):
        cp.control_points_refined.append(cp_new)
        cp.local_refinements.append(level)
        new_cp = cp
    else:
        new_cp = LocallyRefinedControlPoints([base, cp_new], [base_local_refinement(base), level])
    return SplineGrid(tuple(dims_new), new_cp, grid.weights, grid.eval)


def _thb(obj) -> LocallyRefinedControlPoints:
    cp = obj.control_points if isinstance(obj, SplineGrid) else obj
    assert isinstance(cp, LocallyRefinedControlPoints), "needs locally refined control points"
    return cp


def _unique_rows(rows: np.ndarray) -> np.ndarray:
    """unique(A; dims = 1): the distinct rows in order of first occurrence."""
    if len(rows) == 0:
        return rows
    _, first = np.unique(rows, axis=0, return_index=True)
    return rows[np.sort(first)]


def activate_local_refinement_(obj, refinement_indices, *,
                               refinement_index: Optional[int] = None) -> None:
    """activate_local_refinement! (src/control_points.jl:470-516): add the
    control points at ``refinement_indices`` ([k, Nin], 1-based) to the
    active ones of level ``refinement_index`` (1-based, default the last),
    each taking the value the level's array holds now (so the geometry is
    unchanged); rows already active keep their values."""
    cp = _thb(obj)
    level = len(cp.local_refinements) if refinement_index is None else int(refinement_index)
    lr = cp.local_refinements[level - 1]
    arr = cp.control_points_refined[level - 1]
    nin = arr.ndim - 1
    idx = np.asarray(refinement_indices.cpu() if isinstance(refinement_indices, torch.Tensor)
                     else refinement_indices, dtype=np.int64)
    idx = idx.reshape(-1, nin) if idx.size else np.zeros((0, nin), np.int64)
    # @assert size(refinement_indices)[2] == Nin (src/control_points.jl:479)
    assert idx.shape[1] == nin, \
        "Number of indices per control point must match the number of input dimensions."
    shape = np.asarray(arr.shape[:-1])
    if len(idx) and (np.any(idx < 1) or np.any(idx > shape)):
        raise IndexError("refinement indices outside the control point grid")
    old = lr.refinement_indices.cpu().numpy().astype(np.int64)
    new_idx = _unique_rows(np.concatenate([old.reshape(-1, nin), idx], axis=0))
    n_old, n_new = len(old), len(new_idx)
    dev = arr.device
    idx_t = colmajor_empty((n_new, nin), torch.int32, dev)
    idx_t.copy_(torch.from_numpy(new_idx.astype(np.int32)))
    vals = colmajor_empty((n_new, arr.shape[-1]), arr.dtype, dev)
    # refinement_values_new_kernel (src/control_points.jl:427-456)
    vals[:n_old].copy_(lr.refinement_values)
    if n_new > n_old:
        tail_idx = colmajor(idx_t[n_old:])
        tail = colmajor_empty((n_new - n_old, arr.shape[-1]), arr.dtype, dev)
        local_refinement_gather_(tail, arr, tail_idx, clear=False)
        vals[n_old:].copy_(tail)
    cp.local_refinements[level - 1] = LocalRefinement(lr.dims_refinement,
                                                      lr.refinement_matrices, idx_t, vals)


def _as_range(r) -> range:
    if isinstance(r, range):
        return r
    lo, hi = r
    return range(int(lo), int(hi) + 1)


def activate_local_control_point_range_(grid: SplineGrid, *ranges) -> None:
    """activate_local_control_point_range! (src/control_points.jl:531-539):
    activate the Cartesian product of one 1-based index range per dimension,
    each a ``range`` or an inclusive ``(first, last)`` pair (Julia's
    ``first:last``), first dimension fastest."""
    assert len(ranges) == grid.Nin
    activate_local_refinement_(grid, _range_product(ranges))


def _range_product(ranges) -> np.ndarray:
    """The rows of Iterators.product(ranges...) (the first range fastest)."""
    axes = [np.asarray(list(_as_range(r)), dtype=np.int64) for r in ranges]
    mesh = np.meshgrid(*axes, indexing="ij")
    return np.stack([m.reshape(-1, order="F") for m in mesh], axis=1)


def error_informed_local_refinement_(grid: SplineGrid, error: torch.Tensor, *,
                                     threshold_factor: float = 1.0) -> None:
    """error_informed_local_refinement! (src/control_points.jl:555-578): the
    error field (shaped like eval) mapped onto the last level's control
    points by the grid's adjoint, summed over the output channels; the
    control points above threshold_factor times the mean become active."""
    assert tuple(error.shape) == tuple(grid.eval.shape), \
        "The error array must have the same size as the eval array."
    cp_arr = obtain(grid.control_points)
    cp_err = colmajor_zeros(cp_arr.shape, cp_arr.dtype, cp_arr.device)
    evaluate_adjoint_(grid, eval=error, control_points=cp_err)
    grid_err = cp_err.sum(dim=-1)
    threshold = threshold_factor * grid_err.sum() / grid_err.numel()
    above = (grid_err > threshold).cpu().numpy()
    # findall on a column-major array: first index fastest
    hits = np.argwhere(above.transpose())[:, ::-1] + 1
    activate_local_refinement_(grid, hits)


def deactivate_overwritten_control_points_(cp, local_refinement_level: Optional[int] = None) -> None:
    """deactivate_overwritten_control_points! (src/control_points.jl:
    584-680): drop the active control points of a level (1-based; default
    every level but the last, in reverse) whose values no longer reach the
    next level's array because every entry they feed is overwritten there.
    Run on flags: the next level's array all 1 with its own active entries
    0, mapped back through its refinement matrices structurally, then read
    at this level's active entries."""
    cp = _thb(cp)
    if local_refinement_level is None:
        for level in range(len(cp.local_refinements) - 1, 0, -1):
            deactivate_overwritten_control_points_(cp, level)
        return
    level = int(local_refinement_level)
    assert 1 <= level <= len(cp.local_refinements) - 1
    lr, nxt = cp.local_refinements[level - 1], cp.local_refinements[level]
    arr, arr_next = cp.control_points_refined[level - 1], cp.control_points_refined[level]
    dev = arr.device
    flags_next = colmajor_empty(tuple(arr_next.shape[:-1]) + (1,), torch.uint8, dev)
    flags_next.fill_(1)
    n_next = int(nxt.refinement_indices.shape[0])
    zeros_next = colmajor_zeros((n_next, 1), torch.uint8, dev)
    local_refinement_overwrite_(flags_next, nxt.refinement_indices, zeros_next)
    flags = colmajor_zeros(tuple(arr.shape[:-1]) + (1,), torch.uint8, dev)
    mult_adjoint_(flags, nxt.refinement_matrices, flags_next, nxt.dims_refinement)
    n = int(lr.refinement_indices.shape[0])
    alive = colmajor_zeros((n, 1), torch.uint8, dev)
    local_refinement_gather_(alive, flags, lr.refinement_indices, clear=False)
    keep = alive[:, 0].bool()
    cp.local_refinements[level - 1] = LocalRefinement(
        lr.dims_refinement, lr.refinement_matrices,
        colmajor(lr.refinement_indices[keep]), colmajor(lr.refinement_values[keep]))


__all__ = ["insert_knot", "refine", "refine_grid", "add_default_local_refinement",
           "activate_local_refinement_", "activate_local_control_point_range_",
           "error_informed_local_refinement_", "deactivate_overwritten_control_points_"]
