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
        self.sample_points = torch.from_numpy(host_t.copy()).to(self.device)
        self._knots_dev = torch.from_numpy(
            np.asarray(knot_vector.knots_all, dtype=ft).copy()).to(self.device)
        m, w, nd = int(n_sample_points), self.degree + 1, self.max_derivative_order + 1
        self.sample_indices = torch.zeros(m, dtype=torch.int32, device=self.device)
        self.eval = colmajor_zeros((m, w, nd), self.float_type, self.device)
        evaluate_dimension_(self)

    # ---- size helpers, src/utils.jl:31-56 ----
    @property
    def n_basis_functions(self) -> int:
        return len(self.knot_vector.knots_all) - self.degree - 1

    @property
    def n_sample_points(self) -> int:
        return len(self._host_t)

    def host_knots(self) -> np.ndarray:
        return np.asarray(self.knot_vector.knots_all, dtype=np.float64)

    def host_sample_points(self) -> np.ndarray:
        return self._host_t.astype(np.float64)

    def __repr__(self):
        return (f"SplineDimension(n_basis_functions={self.n_basis_functions}, "
                f"degree={self.degree}, n_sample_points={self.n_sample_points}, "
                f"max_derivative_order={self.max_derivative_order}, "
                f"float_type={self.float_type})")


def evaluate_dimension_(s: SplineDimension) -> None:
    """set_sample_indices! + evaluate!(::SplineDimension)
    (src/utils.jl:19-29, src/spline_dimension.jl:231-242) on the GPU."""
    L = _lib.lib()
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
