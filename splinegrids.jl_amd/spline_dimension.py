"""SplineDimension: src/spline_dimension.jl.

The basis table of a dimension (span search + Cox-de Boor with derivative
orders up to ``max_derivative_order``) is computed on the GPU in FP64 by
``sg_dimension_evaluate`` and stored in the value type Tv, in the reference
layout ``eval[m, p+1, D+1]`` (sample fastest).
"""
from __future__ import annotations

from typing import Optional

import numpy as np
import torch

from . import _lib
from .knot_vector import KnotVector, uniform_range
from .utils import (colmajor_empty, colmajor_zeros, current_stream_handle,
                    dtype_code, np_float, np_int, torch_float)


class SplineDimension:
    """Basis functions of one dimension and how they are sampled.

    ``SplineDimension(n_basis_functions, degree, n_sample_points, *,
    max_derivative_order=0, knot_vector=None, device="cuda",
    float_type=np.float32, int_type=np.int32, sample_points=None, **kwargs)``
    follows src/spline_dimension.jl:97-156; extra keyword arguments go to
    the default KnotVector.  ``sample_points`` (host values) replaces the
    default uniform sampling of the knot extent, as the reference's inner
    constructor (src/spline_dimension.jl:40-65) allows.
    """

    def __init__(self, n_basis_functions: int, degree: int, n_sample_points: int, *,
                 max_derivative_order: int = 0,
                 knot_vector: Optional[KnotVector] = None,
                 device="cuda", float_type=np.float32, int_type=np.int32,
                 sample_points=None, **kwargs):
        assert 0 <= max_derivative_order <= degree, (
            "The max_degree must be positive and derivatives order higher than "
            "`degree` are all 0.")
        _lib.lib()  # the tables are computed on the GPU: fail here without one
        ft = np_float(float_type)
        if knot_vector is None:
            knot_vector = KnotVector(n_basis_functions, degree, float_type=ft,
                                     int_type=np_int(int_type), **kwargs)
        else:
            assert len(knot_vector.knots_all) == n_basis_functions + degree + 1, \
                "Incompatible knot vector supplied."
            ft = knot_vector.knot_values.dtype.type
        self.degree = int(degree)
        self.max_derivative_order = int(max_derivative_order)
        self.knot_vector = knot_vector
        self.device = torch.device(device)
        self.float_type = torch_float(ft)
        if sample_points is None:  # src/spline_dimension.jl:121-128
            lo, hi = knot_vector.extent
            host_t = uniform_range(float(lo), float(hi), int(n_sample_points)).astype(ft)
        else:
            host_t = np.asarray(sample_points, dtype=ft)
            assert len(host_t) == n_sample_points
        self._host_t = host_t
        self._host_knots = np.asarray(knot_vector.knots_all, dtype=np.float64).copy()
        # bumped by every evaluate_dimension_: grids rebuild their plans when
        # a dimension's sample points or knots have been re-evaluated
        self.version = 0
        self.sample_points = torch.from_numpy(host_t.copy()).to(self.device)
        self._knots_dev = torch.from_numpy(
            np.asarray(knot_vector.knots_all, dtype=ft).copy()).to(self.device)
        m, w, nd = int(n_sample_points), self.degree + 1, self.max_derivative_order + 1
        self.sample_indices = torch.zeros(m, dtype=torch.int32, device=self.device)
        self.eval = colmajor_zeros((m, w, nd), self.float_type, self.device)
        evaluate_dimension_(self, _refresh=False)

    # ---- size helpers, src/utils.jl:31-56 ----
    @property
    def n_basis_functions(self) -> int:
        return len(self.knot_vector.knots_all) - self.degree - 1

    @property
    def n_sample_points(self) -> int:
        return len(self._host_t)

    def host_knots(self) -> np.ndarray:
        """knots_all as of the last evaluation of this dimension (FP64)."""
        return self._host_knots.copy()

    def host_sample_points(self) -> np.ndarray:
        return self._host_t.astype(np.float64)

    def __repr__(self):
        return (f"SplineDimension(n_basis_functions={self.n_basis_functions}, "
                f"degree={self.degree}, n_sample_points={self.n_sample_points}, "
                f"max_derivative_order={self.max_derivative_order}, "
                f"float_type={self.float_type})")


def evaluate_dimension_(s: SplineDimension, _refresh: bool = True) -> None:
    """set_sample_indices! + evaluate!(::SplineDimension)
    (src/utils.jl:19-29, src/spline_dimension.jl:231-242) on the GPU.

    As in the reference, a caller may edit ``sample_points`` (the device
    vector) or the knot vector's ``knots_all`` and re-evaluate: the call
    takes the current values of both, and grids built on this dimension
    rebuild their plans (tables, tiling) on their next evaluation."""
    L = _lib.lib()
    if _refresh:
        knots = np.asarray(s.knot_vector.knots_all)
        if len(knots) != len(s._knots_dev):
            raise ValueError("the number of knots of a dimension cannot change; "
                             "build a new SplineDimension")
        s._knots_dev.copy_(torch.from_numpy(np.asarray(knots, dtype=s._host_t.dtype).copy()))
        s._host_t = s.sample_points.detach().cpu().numpy().astype(s._host_t.dtype)
        s._host_knots = np.asarray(knots, dtype=np.float64).copy()
        s.version += 1
    with torch.cuda.device(s.device):
        _lib.check(L.sg_dimension_evaluate(
            s._knots_dev.data_ptr(), len(s._knots_dev), s.sample_points.data_ptr(),
            s.n_sample_points, s.degree, s.max_derivative_order,
            s.sample_indices.data_ptr(), s.eval.data_ptr(),
            dtype_code(s.float_type), current_stream_handle(s.device)))


def decompress(s: SplineDimension, derivative_order: int = 0) -> torch.Tensor:
    """Dense (m, n) basis matrix for one derivative order
    (src/spline_dimension.jl:251-272); test support."""
    assert derivative_order <= s.max_derivative_order
    m, n, p = s.n_sample_points, s.n_basis_functions, s.degree
    out = colmajor_zeros((m, n), s.float_type, s.device)
    rows = torch.arange(m, device=s.device).unsqueeze(1).expand(m, p + 1)
    cols = (s.sample_indices.long() - p - 1).unsqueeze(1) + torch.arange(p + 1, device=s.device)
    out[rows, cols] = s.eval[:, :, derivative_order]
    return out


__all__ = ["SplineDimension", "evaluate_dimension_", "decompress", "colmajor_empty"]
