"""SplineGrid / NURBSGrid and the evaluate entry points.

src/spline_grid.jl (grid types, constructors, evaluate!) and
src/adjoint.jl:52-83 (evaluate_adjoint!).  The launches go through the
C-ABI of libsplinegrids_hip.so; a grid builds its plan (FP64 basis tables,
tiling, adjoint workspace) on first use and keeps it.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib
from .control_points import (DefaultControlPoints, LocallyRefinedControlPoints,
                             evaluate_adjoint_control_points_, evaluate_control_points_,
                             obtain, set_unit_cp_grid_)
from .spline_dimension import SplineDimension, evaluate_dimension_
from .utils import (colmajor_zeros, current_stream_handle, dtype_code,
                    is_colmajor)
from .validation import validate_partial_derivatives, validate_spline_grid


class Plan:
    """Owner of one sg_plan (C-ABI handle)."""

    def __init__(self, dims: Sequence[SplineDimension], nout: int, nurbs: bool, device):
        L = _lib.lib()
        self._lib = L
        self.device = torch.device(device)
        self._keep = []
        descs = (_lib.DimDesc * len(dims))()
        for i, s in enumerate(dims):
            knots = np.ascontiguousarray(s.host_knots(), dtype=np.float64)
            t = np.ascontiguousarray(s.host_sample_points(), dtype=np.float64)
            self._keep += [knots, t]
            descs[i] = _lib.DimDesc(
                s.degree, s.n_basis_functions, s.n_sample_points, s.max_derivative_order,
                knots.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                t.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        index = self.device.index if self.device.index is not None else torch.cuda.current_device()
        desc = _lib.GridDesc(len(dims), int(nout), int(bool(nurbs)), index, descs)
        handle = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(L.sg_plan_create(ctypes.byref(desc),
                                        current_stream_handle(self.device),
                                        ctypes.byref(handle)))
        self.handle = handle
        self.dims = [(s.degree, s.max_derivative_order, s.n_sample_points) for s in dims]

    def basis(self, dim: int):
        """(FP64 table (m, p+1, D+1) Fortran order, 1-based sample indices)."""
        p, D, m = self.dims[dim]
        table = np.zeros((m, p + 1, D + 1), dtype=np.float64, order="F")
        idx = np.zeros(m, dtype=np.int32)
        _lib.check(self._lib.sg_plan_basis(
            self.handle, dim, table.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
            idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return table, idx

    def device_bytes(self) -> int:
        out = ctypes.c_int64()
        _lib.check(self._lib.sg_plan_device_bytes(self.handle, ctypes.byref(out)))
        return out.value

    def describe(self) -> str:
        """The kernels and tile shapes this plan uses, as one line of text."""
        buf = ctypes.create_string_buffer(1024)
        _lib.check(self._lib.sg_plan_describe(self.handle, buf, len(buf)))
        return buf.value.decode()

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                self._lib.sg_plan_destroy(h)
            except Exception:
                pass
            self.handle = None


class SplineGrid:
    """``SplineGrid(spline_dimensions, Nout)`` (src/spline_grid.jl:70-94) or
    ``SplineGrid(spline_dimensions, control_points, weights, eval)`` (the
    inner constructor, :34-57).  A single SplineDimension may be passed
    instead of a tuple (:111-113)."""

    def __init__(self, spline_dimensions, *args):
        if isinstance(spline_dimensions, SplineDimension):
            spline_dimensions = (spline_dimensions,)
        spline_dimensions = tuple(spline_dimensions)
        if len(args) == 1:
            nout = int(args[0])
            s0 = spline_dimensions[0]
            dev, tv = s0.device, s0.float_type
            size_cp = tuple(s.n_basis_functions for s in spline_dimensions)
            size_eval = tuple(s.n_sample_points for s in spline_dimensions)
            cp = colmajor_zeros(size_cp + (nout,), tv, dev)
            set_unit_cp_grid_(cp)
            control_points = DefaultControlPoints(cp)
            weights = None
            eval = colmajor_zeros(size_eval + (nout,), tv, dev)
        elif len(args) == 3:
            control_points, weights, eval = args
            if not isinstance(control_points, (DefaultControlPoints,
                                               LocallyRefinedControlPoints)):
                control_points = DefaultControlPoints(control_points)
        else:
            raise TypeError("SplineGrid(dims, Nout) or SplineGrid(dims, control_points, weights, eval)")
        validate_spline_grid(spline_dimensions, control_points, weights, eval)
        self.spline_dimensions = spline_dimensions
        self.control_points = control_points
        self.weights = weights
        self.eval = eval
        self._plan: Optional[Plan] = None
        self._plan_key = None

    @property
    def Nin(self) -> int:
        return len(self.spline_dimensions)

    @property
    def Nout(self) -> int:
        return int(obtain(self.control_points).shape[-1])

    @property
    def is_nurbs(self) -> bool:
        return self.weights is not None

    @property
    def device(self):
        return obtain(self.control_points).device

    @property
    def dtype(self):
        return obtain(self.control_points).dtype

    def _dims_key(self):
        return tuple((id(s), s.version) for s in self.spline_dimensions)

    def plan(self) -> Plan:
        """The C-ABI plan of this grid, rebuilt when a dimension has been
        re-evaluated (new sample points or knots) since it was built."""
        key = self._dims_key()
        if self._plan is None or self._plan_key != key:
            self._plan = None
            self._plan = Plan(self.spline_dimensions, self.Nout, self.is_nurbs, self.device)
            self._plan_key = key
        return self._plan

    def __repr__(self):
        kind = "NURBSGrid" if self.is_nurbs else "SplineGrid"
        return (f"{kind}(Nin={self.Nin}, Nout={self.Nout}, "
                f"control points {tuple(obtain(self.control_points).shape[:-1])}, "
                f"samples {tuple(self.eval.shape[:-1])}, {self.dtype})")


def NURBSGrid(spline_dimensions, Nout: int) -> SplineGrid:
    """src/spline_grid.jl:100-108: a SplineGrid with unit weights."""
    grid = SplineGrid(spline_dimensions, Nout)
    size_cp = tuple(s.n_basis_functions for s in grid.spline_dimensions)
    w = colmajor_zeros(size_cp, grid.dtype, grid.device)
    w.fill_(1)
    grid.weights = w
    grid._plan = None
    return grid


def _orders(grid: SplineGrid, derivative_order):
    if derivative_order is None:
        return (0,) * grid.Nin
    return tuple(int(o) for o in derivative_order)


def _check_array(name: str, t: torch.Tensor, grid: SplineGrid) -> None:
    if t.dtype != grid.dtype:
        raise TypeError(f"{name} has value type {t.dtype}, the grid {grid.dtype}")
    if t.device != grid.device:
        raise ValueError(f"{name} is on {t.device}, the grid on {grid.device}")
    if not is_colmajor(t):
        raise ValueError(f"{name} must be column-major (Julia layout); see reshape_colmajor")


def evaluate_(obj, *, derivative_order=None, control_points=None, eval=None,
              tier: str = "native", synchronize: bool = False) -> None:
    """evaluate!: dispatches like the reference.

    - SplineDimension: recompute its basis table (src/spline_dimension.jl:231)
    - DefaultControlPoints: nothing to do (src/control_points.jl:313)
    - LocallyRefinedControlPoints: run the THB chain of refinement products
      and overwrites (src/control_points.jl:319-349)
    - SplineGrid: eval[J, c] = sum_I prod_d B_d[J_d, I_d, o_d] (w) P[...]
      (src/spline_grid.jl:200-230); eval is overwritten.  ``tier`` picks the
      arithmetic of the contraction ("native", "split_f16", "bf16").
    """
    if isinstance(obj, SplineDimension):
        evaluate_dimension_(obj)
        return None
    if isinstance(obj, DefaultControlPoints) or isinstance(obj, torch.Tensor):
        return None
    if isinstance(obj, LocallyRefinedControlPoints):
        evaluate_control_points_(obj)
        return None
    grid: SplineGrid = obj
    orders = _orders(grid, derivative_order)
    validate_partial_derivatives(grid, orders)
    P = obtain(grid.control_points if control_points is None else control_points)
    ev = grid.eval if eval is None else eval
    assert tuple(P.shape) == tuple(obtain(grid.control_points).shape), \
        "size(control_points) == size(spline_grid.control_points)"
    assert tuple(ev.shape) == tuple(grid.eval.shape), "size(eval) == size(spline_grid.eval)"
    _check_array("control_points", P, grid)
    _check_array("eval", ev, grid)
    W = grid.weights
    if W is not None:
        _check_array("weights", W, grid)
    plan = grid.plan()
    L = _lib.lib()
    with torch.cuda.device(grid.device):
        stream = current_stream_handle(grid.device)
        _lib.check(L.sg_evaluate(plan.handle, _lib.i32_array(orders), P.data_ptr(),
                                 None if W is None else W.data_ptr(), ev.data_ptr(),
                                 dtype_code(grid.dtype), _lib.TIERS[tier], stream))
        if synchronize:
            _lib.check(L.sg_stream_synchronize(stream))
    return None


def evaluate_fields_(grid: SplineGrid, derivative_orders, evals, *,
                     control_points=None, tier: str = "native") -> None:
    """Several derivative fields of a B-spline grid in one pass (as many
    evaluate! calls with different derivative_order and eval buffers)."""
    orders = [tuple(int(o) for o in d) for d in derivative_orders]
    assert len(orders) == len(evals) and 1 <= len(orders) <= _lib.MAX_FIELDS
    for o in orders:
        validate_partial_derivatives(grid, o)
    assert not grid.is_nurbs, "Computing derivatives of NURBS is currently not supported."
    P = obtain(grid.control_points if control_points is None else control_points)
    _check_array("control_points", P, grid)
    for ev in evals:
        assert tuple(ev.shape) == tuple(grid.eval.shape)
        _check_array("eval", ev, grid)
    plan = grid.plan()
    L = _lib.lib()
    flat = [o for d in orders for o in d]
    ptrs = (ctypes.c_void_p * len(evals))(*[e.data_ptr() for e in evals])
    with torch.cuda.device(grid.device):
        _lib.check(L.sg_evaluate_fields(plan.handle, len(orders), _lib.i32_array(flat),
                                        P.data_ptr(), ptrs, dtype_code(grid.dtype),
                                        _lib.TIERS[tier], current_stream_handle(grid.device)))


def evaluate_rational_fields_(grid: SplineGrid, derivative_orders, evals, *,
                              control_points=None) -> None:
    """Derivative fields of a NURBS grid's rational map R = sum N w P / sum N w.

    The reference rejects NURBS derivatives (src/validation.jl:81-88), and so
    does evaluate_; this extension applies the quotient rule to the
    homogeneous B-spline fields of (w P, w) and combines them in FP64.  The
    device workspace is allocated on the first call and kept on the grid."""
    assert grid.is_nurbs, "rational derivatives need a NURBS grid"
    orders = [tuple(int(o) for o in d) for d in derivative_orders]
    assert len(orders) == len(evals) and 1 <= len(orders) <= _lib.MAX_FIELDS
    for o in orders:
        assert len(o) == grid.Nin, "one derivative order per input dimension"
    P = obtain(grid.control_points if control_points is None else control_points)
    _check_array("control_points", P, grid)
    _check_array("weights", grid.weights, grid)
    for ev in evals:
        assert tuple(ev.shape) == tuple(grid.eval.shape)
        _check_array("eval", ev, grid)
    plan = grid.plan()
    L = _lib.lib()
    flat = _lib.i32_array([o for d in orders for o in d])
    dt = dtype_code(grid.dtype)
    nbytes = ctypes.c_int64(0)
    _lib.check(L.sg_rational_workspace_bytes(plan.handle, len(orders), flat, dt,
                                             ctypes.byref(nbytes)))
    ws = getattr(grid, "_rational_workspace", None)
    if ws is None or ws.numel() < nbytes.value:
        ws = torch.empty(max(nbytes.value, 1), dtype=torch.uint8, device=grid.device)
        grid._rational_workspace = ws
    ptrs = (ctypes.c_void_p * len(evals))(*[e.data_ptr() for e in evals])
    with torch.cuda.device(grid.device):
        _lib.check(L.sg_evaluate_rational_fields(plan.handle, len(orders), flat, P.data_ptr(),
                                                 grid.weights.data_ptr(), ptrs, ws.data_ptr(),
                                                 ws.numel(), dt,
                                                 current_stream_handle(grid.device)))


def evaluate_rational_adjoint_(grid: SplineGrid, *, control_points=None, eval=None,
                               accumulate: bool = False) -> None:
    """The adjoint of a NURBS grid's evaluate_ (derivative order 0) in its
    control points, the weights fixed: control_points = w_k [B^T (eval /
    w(x))]_k, overwritten (or added to with ``accumulate=True``).

    The reference's evaluate_adjoint! rejects NURBS (src/adjoint.jl:57), and
    so does evaluate_adjoint_; this is the extension.  The device workspace is
    allocated on the first call and kept on the grid."""
    assert grid.is_nurbs, "the rational adjoint needs a NURBS grid"
    P = obtain(grid.control_points if control_points is None else control_points)
    ev = grid.eval if eval is None else eval
    assert tuple(P.shape) == tuple(obtain(grid.control_points).shape)
    assert tuple(ev.shape) == tuple(grid.eval.shape)
    _check_array("control_points", P, grid)
    _check_array("eval", ev, grid)
    _check_array("weights", grid.weights, grid)
    plan = grid.plan()
    L = _lib.lib()
    dt = dtype_code(grid.dtype)
    nbytes = ctypes.c_int64(0)
    _lib.check(L.sg_rational_adjoint_workspace_bytes(plan.handle, dt, ctypes.byref(nbytes)))
    ws = getattr(grid, "_rational_adjoint_workspace", None)
    if ws is None or ws.numel() < nbytes.value:
        ws = torch.empty(max(nbytes.value, 1), dtype=torch.uint8, device=grid.device)
        grid._rational_adjoint_workspace = ws
    with torch.cuda.device(grid.device):
        _lib.check(L.sg_evaluate_rational_adjoint(plan.handle, grid.weights.data_ptr(),
                                                  ev.data_ptr(), P.data_ptr(), ws.data_ptr(),
                                                  ws.numel(), dt, int(bool(accumulate)),
                                                  current_stream_handle(grid.device)))


def evaluate_adjoint_(grid, *, derivative_order=None, control_points=None,
                      eval=None, tier: str = "native", accumulate: bool = False,
                      synchronize: bool = False) -> None:
    """evaluate_adjoint! (src/adjoint.jl:52-83): control_points = M^T eval,
    overwritten (or added to with ``accumulate=True``).  B-splines only.

    On a control-point container it is the adjoint of that container's
    evaluate_: nothing for DefaultControlPoints, the reversed THB chain for
    LocallyRefinedControlPoints (src/adjoint.jl:172-205)."""
    if isinstance(grid, (DefaultControlPoints, torch.Tensor)):
        return None
    if isinstance(grid, LocallyRefinedControlPoints):
        evaluate_adjoint_control_points_(grid)
        return None
    assert not grid.is_nurbs, "Adjoint evaluation not supported for NURBS."
    orders = _orders(grid, derivative_order)
    validate_partial_derivatives(grid.spline_dimensions, orders)
    P = obtain(grid.control_points if control_points is None else control_points)
    ev = grid.eval if eval is None else eval
    assert tuple(P.shape) == tuple(obtain(grid.control_points).shape), \
        "size(control_points) == size(spline_grid.control_points)"
    assert tuple(ev.shape) == tuple(grid.eval.shape), "size(eval) == size(spline_grid.eval)"
    _check_array("control_points", P, grid)
    _check_array("eval", ev, grid)
    plan = grid.plan()
    L = _lib.lib()
    with torch.cuda.device(grid.device):
        stream = current_stream_handle(grid.device)
        _lib.check(L.sg_evaluate_adjoint(plan.handle, _lib.i32_array(orders), ev.data_ptr(),
                                         P.data_ptr(), dtype_code(grid.dtype),
                                         _lib.TIERS[tier], int(bool(accumulate)), stream))
        if synchronize:
            _lib.check(L.sg_stream_synchronize(stream))
    return None


def reshape_colmajor(flat: torch.Tensor, shape) -> torch.Tensor:
    """Julia's reshape: a column-major view of a flat vector."""
    shape = tuple(int(s) for s in shape)
    return flat.view(shape[::-1]).permute(*range(len(shape) - 1, -1, -1))


def vec_colmajor(t: torch.Tensor) -> torch.Tensor:
    """Julia's vec: the column-major flattening (a view if t is column-major)."""
    return t.permute(*range(t.ndim - 1, -1, -1)).reshape(-1)
