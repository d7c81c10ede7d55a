"""Control-point containers: src/control_points.jl.

DefaultControlPoints (:1-57) wraps the dense array the grid kernels read.
LocalRefinement and LocallyRefinedControlPoints (:59-349) hold a truncated
hierarchical (THB) basis: a chain of dense arrays, each the previous one
times refinement matrices with some control points then overwritten by
their own values; ``evaluate_control_points_`` runs that chain and
``evaluate_adjoint_control_points_`` its transpose (src/adjoint.jl:154-205),
both in libsplinegrids_hip.so.  Refinement indices are 1-based int32, as the
reference stores them.
"""
from __future__ import annotations

import ctypes
import math
from typing import List, Sequence

import numpy as np
import torch

from . import _lib
from .refinement_matrix import RefinementMatrix, mult_, mult_adjoint_
from .utils import (colmajor, colmajor_empty, current_stream_handle, dtype_code,
                    from_numpy_colmajor, is_colmajor)

_SUPERSCRIPT = str.maketrans("0123456789", "⁰¹²³⁴⁵⁶⁷⁸⁹")


def _jl_type(dtype) -> str:
    return {torch.float32: "Float32", torch.float64: "Float64"}.get(dtype, str(dtype))


def _jl_size(shape) -> str:
    t = tuple(int(x) for x in shape)
    return f"({t[0]},)" if len(t) == 1 else "(" + ", ".join(str(x) for x in t) + ")"


class DefaultControlPoints:
    """A thin wrapper around a column-major array (n_1, ..., n_Nin, Nout)."""

    def __init__(self, control_points: torch.Tensor):
        assert is_colmajor(control_points), "control points must be column-major"
        self.control_points = control_points

    # Base.size / ndims / length / eltype (src/control_points.jl:269-274)
    @property
    def shape(self):
        return tuple(self.control_points.shape)

    @property
    def ndim(self) -> int:
        return self.control_points.ndim

    @property
    def dtype(self):
        return self.control_points.dtype

    @property
    def device(self):
        return self.control_points.device

    def __len__(self) -> int:
        return self.control_points.numel()

    def __getitem__(self, idx):
        return self.control_points[idx]

    def __setitem__(self, idx, val):
        self.control_points[idx] = val

    def copy_(self, vals) -> "DefaultControlPoints":
        """copyto!(control_points, vals): an array of the same shape, or an
        (n_control_points, Nout) matrix in column-major order
        (src/control_points.jl:45-57)."""
        cp = self.control_points
        if isinstance(vals, np.ndarray):
            vals = from_numpy_colmajor(vals, cp.dtype, cp.device)
        vals = vals.to(device=cp.device, dtype=cp.dtype)
        if tuple(vals.shape) != tuple(cp.shape):
            n, nout = get_n_control_points(self), cp.shape[-1]
            assert vals.numel() == n * nout, "size mismatch in copyto!"
            flat = _colmajor_flat(vals)
            vals = flat.view(tuple(cp.shape)[::-1]).permute(*range(cp.ndim - 1, -1, -1))
        cp.copy_(vals)
        return self

    def __repr__(self):
        """Base.show (src/control_points.jl:25-33)."""
        nin = self.ndim - 1
        return (f"DefaultControlPoints for grid of size {_jl_size(self.shape[:nin])} in "
                f"\u211d{str(self.shape[-1]).translate(_SUPERSCRIPT)} ({_jl_type(self.dtype)}).")


def _colmajor_flat(t: torch.Tensor) -> torch.Tensor:
    return t.permute(*range(t.ndim - 1, -1, -1)).reshape(-1)


def obtain(control_points):
    """The dense array behind a control-point container: for THB control
    points the last level (src/control_points.jl:351-356)."""
    if isinstance(control_points, DefaultControlPoints):
        return control_points.control_points
    if isinstance(control_points, LocallyRefinedControlPoints):
        return control_points.control_points_refined[-1]
    return control_points


def get_n_control_points(obj) -> int:
    """src/control_points.jl:276-294: the grid's control points, or for THB
    control points the active ones of every level."""
    from .spline_grid import SplineGrid

    if isinstance(obj, SplineGrid):
        obj = obj.control_points
    if isinstance(obj, LocallyRefinedControlPoints):
        return sum(int(lr.refinement_indices.shape[0]) for lr in obj.local_refinements)
    shape = obtain(obj).shape
    return int(math.prod(shape[:-1]))


def set_unit_cp_grid_(control_points: torch.Tensor) -> None:
    """Default control points: a unit grid (src/utils.jl:1-15)."""
    nin = control_points.ndim - 1
    nout = control_points.shape[-1]
    for n in range(min(nin, nout)):
        k = control_points.shape[n]
        coords = torch.linspace(0, 1, k, dtype=torch.float64) if k > 1 else torch.zeros(1, dtype=torch.float64)
        shape = [1] * nin
        shape[n] = k
        control_points[..., n] = coords.view(shape).to(control_points)


# --------------------------------------------------------------------------
# locally refined (THB) control points
# --------------------------------------------------------------------------

class LocalRefinement:
    """One level of a THB hierarchy (src/control_points.jl:59-110): the
    refinement matrices that map the previous level's array to this one's
    along the 1-based dimensions ``dims_refinement``, then the control points
    at ``refinement_indices`` ([k, Nin] int32, 1-based) overwritten by
    ``refinement_values`` ([k, Nout]); both column-major device tensors."""

    def __init__(self, dims_refinement: Sequence[int],
                 refinement_matrices: Sequence[RefinementMatrix],
                 refinement_indices: torch.Tensor, refinement_values: torch.Tensor):
        assert refinement_indices.dtype == torch.int32, "refinement indices are int32"
        assert refinement_indices.ndim == 2 and refinement_values.ndim == 2
        assert refinement_indices.shape[0] == refinement_values.shape[0]
        self.dims_refinement: List[int] = [int(d) for d in dims_refinement]
        self.refinement_matrices: List[RefinementMatrix] = list(refinement_matrices)
        self.refinement_indices = colmajor(refinement_indices)
        self.refinement_values = colmajor(refinement_values)

    @property
    def Nin(self) -> int:
        return int(self.refinement_indices.shape[1])

    @property
    def Nout(self) -> int:
        return int(self.refinement_values.shape[1])


def base_local_refinement(control_points_base: torch.Tensor) -> LocalRefinement:
    """The base level (src/control_points.jl:125-154): every control point of
    the dense base array active, in column-major order, with its value."""
    shape = tuple(control_points_base.shape[:-1])
    nout = int(control_points_base.shape[-1])
    dev = control_points_base.device
    n = int(math.prod(shape))
    idx = colmajor_empty((n, len(shape)), torch.int32, dev)
    lin = torch.arange(n, device=dev, dtype=torch.int64)
    for d, k in enumerate(shape):  # CartesianIndices, first index fastest
        idx[:, d] = (lin % k + 1).to(torch.int32)
        lin = lin // k
    vals = colmajor_empty((n, nout), control_points_base.dtype, dev)
    vals.copy_(_flat_rows(control_points_base))
    return LocalRefinement([], [], idx, vals)


def control_point_matrix(control_points) -> torch.Tensor:
    """The control points as an (n_control_points, Nout) matrix: the dense
    array reshaped (a view), or the active values of every level stacked
    (a copy) -- the matrix copyto!(M, control_points) fills."""
    if isinstance(control_points, LocallyRefinedControlPoints):
        return control_points.get_control_point_view()
    return _flat_rows(obtain(control_points))


def _flat_rows(a: torch.Tensor) -> torch.Tensor:
    """reshape(a, n_control_points, Nout) of a column-major array."""
    n = int(math.prod(a.shape[:-1]))
    return a.permute(*range(a.ndim - 1, -1, -1)).reshape(a.shape[-1], n).t()


class LocallyRefinedControlPoints:
    """``LocallyRefinedControlPoints(control_points_refined,
    local_refinements)`` (src/control_points.jl:156-198): the dense arrays
    of every level (the first the base) and one LocalRefinement per level.
    ``obtain`` gives the last level's array, which the grid kernels read."""

    def __init__(self, control_points_refined: Sequence[torch.Tensor],
                 local_refinements: Sequence[LocalRefinement]):
        assert len(control_points_refined) == len(local_refinements) >= 1
        for a in control_points_refined:
            assert is_colmajor(a), "control point arrays must be column-major"
        self.control_points_refined: List[torch.Tensor] = list(control_points_refined)
        self.local_refinements: List[LocalRefinement] = list(local_refinements)

    # Base.size / ndims / length / eltype (:269-274)
    @property
    def shape(self):
        return tuple(obtain(self).shape)

    @property
    def ndim(self) -> int:
        return obtain(self).ndim

    @property
    def dtype(self):
        return obtain(self).dtype

    @property
    def device(self):
        return obtain(self).device

    def __len__(self) -> int:
        return obtain(self).numel()

    @property
    def Nout(self) -> int:
        return int(self.shape[-1])

    def get_control_point_view(self) -> torch.Tensor:
        """The active control points of every level stacked ([n, Nout],
        get_control_point_view, :200-208), as a copy."""
        return torch.cat([lr.refinement_values for lr in self.local_refinements], dim=0)

    def __getitem__(self, idx):
        return self.get_control_point_view()[idx]

    def __setitem__(self, idx, vals):
        M = self.get_control_point_view()
        M[idx] = vals
        self.copy_(M)

    def copy_(self, vals) -> "LocallyRefinedControlPoints":
        """copyto!(control_points, M) (:222-229): an (n_control_points, Nout)
        matrix, split over the levels in order."""
        n, nout = get_n_control_points(self), self.Nout
        if isinstance(vals, np.ndarray):
            vals = torch.from_numpy(np.asarray(vals))
        vals = vals.to(device=self.device, dtype=self.dtype)
        if tuple(vals.shape) != (n, nout):
            assert vals.numel() == n * nout, "size mismatch in copyto!"
            vals = vals.reshape(-1).view(nout, n).t()  # column-major flat vector
        k = 0
        for lr in self.local_refinements:
            r = int(lr.refinement_values.shape[0])
            lr.refinement_values.copy_(vals[k:k + r])
            k += r
        return self

    def copy_to(self, M: torch.Tensor) -> torch.Tensor:
        """copyto!(M, control_points) (:231-233)."""
        M.copy_(self.get_control_point_view().reshape(M.shape) if M.ndim != 2
                else self.get_control_point_view())
        return M

    def __repr__(self):
        """Base.show (:235-267): the header line and one row per refinement
        matrix (input dimension, control points before and after, activated
        control points)."""
        nin = self.ndim - 1
        head = (f"LocallyRefinedControlPoints for final grid of size "
                f"{_jl_size(self.shape[:nin])} in "
                f"\u211d{str(self.Nout).translate(_SUPERSCRIPT)} ({_jl_type(self.dtype)}). "
                "Local refinements:")
        rows = [("NaN", "NaN", "NaN", str(int(self.local_refinements[0].refinement_indices.shape[0])))]
        for lr in self.local_refinements[1:]:
            k = len(lr.dims_refinement)
            for i, (d, A) in enumerate(zip(lr.dims_refinement, lr.refinement_matrices)):
                act = int(lr.refinement_indices.shape[0]) if i == k - 1 else 0
                rows.append((str(d), str(A.n), str(A.m), str(act)))
        header = ("input dim.", "# c.p. before", "# c.p. after", "# activated c.p.")
        widths = [max(len(header[c]), *(len(r[c]) for r in rows)) for c in range(4)]
        line = "+" + "+".join("-" * (w + 2) for w in widths) + "+"
        fmt = lambda r: "| " + " | ".join(v.rjust(w) for v, w in zip(r, widths)) + " |"
        return "\n".join([head, line, fmt(header), line, *map(fmt, rows), line])


def _lr_call(cp: torch.Tensor, indices: torch.Tensor):
    nin, nout = cp.ndim - 1, int(cp.shape[-1])
    return nin, nout, _lib.i32_array(cp.shape[:-1]), int(indices.shape[0])


def _dtype_of(t: torch.Tensor) -> int:
    return _lib.SG_U8 if t.dtype == torch.uint8 else dtype_code(t.dtype)


def local_refinement_overwrite_(cp: torch.Tensor, indices: torch.Tensor,
                                values: torch.Tensor) -> None:
    """local_refinement_kernel (src/control_points.jl:296-311):
    cp[indices[i, :]..., c] = values[i, c]."""
    assert is_colmajor(cp) and is_colmajor(indices) and is_colmajor(values)
    assert values.dtype == cp.dtype and indices.dtype == torch.int32
    assert values.shape[0] == indices.shape[0] and indices.shape[1] == cp.ndim - 1
    nin, nout, shape, n = _lr_call(cp, indices)
    with torch.cuda.device(cp.device):
        _lib.check(_lib.lib().sg_local_refinement_overwrite(
            nin, nout, shape, n, indices.data_ptr(), values.data_ptr(), cp.data_ptr(),
            _dtype_of(cp), current_stream_handle(cp.device)))


def local_refinement_gather_(values: torch.Tensor, cp: torch.Tensor, indices: torch.Tensor,
                             *, clear: bool) -> None:
    """values[i, c] = cp[indices[i, :]..., c]; with ``clear`` those entries of
    cp are then zeroed (local_refinement_adjoint_kernel, src/adjoint.jl:
    154-170)."""
    assert is_colmajor(cp) and is_colmajor(indices) and is_colmajor(values)
    assert values.dtype == cp.dtype and indices.dtype == torch.int32
    assert values.shape[0] == indices.shape[0] and indices.shape[1] == cp.ndim - 1
    nin, nout, shape, n = _lr_call(cp, indices)
    with torch.cuda.device(cp.device):
        _lib.check(_lib.lib().sg_local_refinement_gather(
            nin, nout, shape, n, indices.data_ptr(), values.data_ptr(), cp.data_ptr(),
            int(bool(clear)), _dtype_of(cp), current_stream_handle(cp.device)))


def evaluate_control_points_(cp: LocallyRefinedControlPoints) -> None:
    """evaluate!(::LocallyRefinedControlPoints) (src/control_points.jl:
    319-349): level by level, the previous array times the level's
    refinement matrices, then the level's control points overwritten."""
    for i, lr in enumerate(cp.local_refinements):
        new = cp.control_points_refined[i]
        if i > 0:
            mult_(new, lr.refinement_matrices, cp.control_points_refined[i - 1],
                  lr.dims_refinement)
        if lr.refinement_indices.shape[0] > 0:
            local_refinement_overwrite_(new, lr.refinement_indices, lr.refinement_values)


def evaluate_adjoint_control_points_(cp: LocallyRefinedControlPoints) -> None:
    """evaluate_adjoint!(::LocallyRefinedControlPoints) (src/adjoint.jl:
    174-205): from the last level's array (holding a cotangent), levels in
    reverse: the level's control points take their entries (which are then
    zeroed), and the previous array becomes the transposed refinement
    product of this one (overwritten)."""
    for i in range(len(cp.local_refinements) - 1, -1, -1):
        lr = cp.local_refinements[i]
        new = cp.control_points_refined[i]
        if lr.refinement_indices.shape[0] > 0:
            local_refinement_gather_(lr.refinement_values, new, lr.refinement_indices, clear=True)
        if i > 0:
            mult_adjoint_(cp.control_points_refined[i - 1], lr.refinement_matrices, new,
                          lr.dims_refinement)
