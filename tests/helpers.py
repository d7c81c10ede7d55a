"""Shared helpers: build the same grid for the HIP path and the oracle."""
from __future__ import annotations

import numpy as np
import torch

import oracle as O

TOL = {np.float64: 1e-12, np.float32: 1e-6}  # relative to max|oracle|


def rel_err(got, want) -> float:
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    scale = max(float(np.max(np.abs(want))), 1e-300) if want.size else 1.0
    return float(np.max(np.abs(got - want))) / scale if want.size else 0.0


def clamped_knot_vector(sg, values, p, dtype=np.float32, doubled=()):
    """KnotVector with end multiplicity p+1 and the given interior knots
    repeated twice."""
    values = np.asarray(values, dtype=dtype)
    mults = np.ones(len(values), dtype=np.int32)
    mults[0] = mults[-1] = p + 1
    for i in doubled:
        mults[i] = 2
    return sg.KnotVector(values, mults)


def make_dims(sg, n, p, m, D=None, dtype=np.float32, knot_vectors=None,
              sample_points=None):
    nin = len(n)
    D = D or [0] * nin
    dims = []
    for i in range(nin):
        kw = {}
        if knot_vectors is not None and knot_vectors[i] is not None:
            kw["knot_vector"] = knot_vectors[i]
        if sample_points is not None and sample_points[i] is not None:
            kw["sample_points"] = sample_points[i]
        dims.append(sg.SplineDimension(n[i], p[i], m[i], max_derivative_order=D[i],
                                       float_type=dtype, device="cuda", **kw))
    return dims


def oracle_dims(dims):
    return [O.Dim(s.host_knots(), s.host_sample_points(), s.degree,
                  s.max_derivative_order) for s in dims]


def random_cp(rng, n, nout, dtype):
    return rng.uniform(-1, 1, tuple(n) + (nout,)).astype(dtype)


def eval_host(sg, t: torch.Tensor) -> np.ndarray:
    torch.cuda.synchronize()
    return sg.to_numpy_colmajor(t)
