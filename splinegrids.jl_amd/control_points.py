"""DefaultControlPoints and obtain: src/control_points.jl:1-57, :269-294,
:313, :351-356.  Only the dense container is on the hot path; the locally
refined (THB) containers of the reference are outside this library."""
from __future__ import annotations

import math

import numpy as np
import torch

from .utils import from_numpy_colmajor, is_colmajor


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
        nin = self.ndim - 1
        return (f"DefaultControlPoints for grid of size {tuple(self.shape[:nin])} in "
                f"R^{self.shape[-1]} ({self.dtype}).")


def _colmajor_flat(t: torch.Tensor) -> torch.Tensor:
    return t.permute(*range(t.ndim - 1, -1, -1)).reshape(-1)


def obtain(control_points):
    """The dense array behind a control-point container."""
    if isinstance(control_points, DefaultControlPoints):
        return control_points.control_points
    return control_points


def get_n_control_points(obj) -> int:
    from .spline_grid import SplineGrid

    if isinstance(obj, SplineGrid):
        obj = obj.control_points
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
