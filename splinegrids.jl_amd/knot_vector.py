"""KnotVector: src/knot_vector.jl.

Knot vectors are set-up data, held on the host as numpy arrays of the value
type Tv; a plan uploads the expanded knots to the GPU once.
"""
from __future__ import annotations

from typing import Optional

import numpy as np

from .utils import np_float, np_int


class KnotVector:
    """Knots with multiplicities.

    ``KnotVector(knot_values, multiplicities)`` validates and expands
    (src/knot_vector.jl:21-49); ``KnotVector(n_basis_functions, degree, *,
    extent=(0, 1), distribution="equispaced", float_type, int_type, rng)``
    builds the clamped default (src/knot_vector.jl:75-113), mirroring the
    reference's two constructor methods.
    """

    def __init__(self, *args, **kwargs):
        if args and isinstance(args[0], (int, np.integer)):
            values, mults = _default_knots(*args, **kwargs)
        else:
            values, mults = args
            assert not kwargs, f"unexpected keyword arguments {sorted(kwargs)}"
        values = np.asarray(values)
        mults = np.asarray(mults)
        # src/knot_vector.jl:22-26
        assert len(values) == len(mults), \
            "knot_values and multiplicities must be of the same length."
        assert np.array_equal(values, np.sort(values)), "knot_values must be sorted."
        assert len(np.unique(values)) == len(values), "knot_values must be unique."
        if not np.issubdtype(values.dtype, np.floating):
            values = values.astype(np.float64)
        self.knot_values = values
        self.multiplicities = mults
        # expand_knot_vector_kernel, src/util_kernels.jl:1-20
        self.knots_all = np.repeat(values, mults.astype(np.int64))
        self.extent = (values[0], values[-1])

    def __repr__(self):
        return (f"KnotVector(knot_values={self.knot_values!r}, "
                f"multiplicities={self.multiplicities!r})")


def _default_knots(n_basis_functions: int, degree: int, *, extent=(0, 1),
                   distribution: str = "equispaced", float_type=np.float32,
                   int_type=np.int32, rng: Optional[np.random.Generator] = None):
    ft, it = np_float(float_type), np_int(int_type)
    assert n_basis_functions - degree >= 1
    k = n_basis_functions - degree + 1
    lo, hi = float(extent[0]), float(extent[1])
    if distribution == "random":  # src/knot_vector.jl:87-92
        rng = np.random.default_rng() if rng is None else rng
        v = np.cumsum(rng.random(k).astype(ft))
        v = v - v[0]
        v = v / (v[-1] - v[0])
        v = v * ft(hi - lo) + ft(lo)
        values = v.astype(ft)
    elif distribution == "equispaced":  # src/knot_vector.jl:94-100
        values = uniform_range(lo, hi, k).astype(ft)
    else:
        raise ValueError(f"Unsupported knot distribution type {distribution}.")
    mults = np.ones(k, dtype=it)  # src/knot_vector.jl:106-108
    mults[0] = degree + 1
    mults[-1] = degree + 1
    return values, mults


def uniform_range(lo: float, hi: float, length: int) -> np.ndarray:
    """``range(lo, hi; length)`` in FP64: uniform, both endpoints exact."""
    if length == 0:
        return np.zeros(0)
    if length == 1:
        return np.array([lo], dtype=np.float64)
    i = np.arange(length, dtype=np.float64)
    t = lo + (hi - lo) * (i / (length - 1))
    t[0], t[-1] = lo, hi
    return t
