"""Derivative fields of NURBS grids (evaluate_rational_fields_) against the
oracle's quotient-rule restatement (oracle.rational_derivs, itself pinned by
the unit-weight, circle and finite-difference checks of tests/test_oracle.py).

The reference rejects NURBS derivatives (src/validation.jl:81-88); evaluate_
keeps doing so, and this extension is checked against mathematics only.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle as O
from helpers import eval_host, make_dims, oracle_dims, rel_err

pytestmark = pytest.mark.gpu

# (n, p, m, D, nout, field sets)
CASES = [
    ((9, 8), (3, 2), (37, 29), (1, 1), 3, [[(0, 0), (1, 0), (0, 1), (1, 1)]]),
    ((7, 6), (3, 3), (40, 33), (2, 2), 2, [[(2, 0), (0, 2)], [(2, 1), (1, 2)]]),
    ((6, 7, 5), (2, 3, 2), (21, 18, 13), (1, 1, 1), 3,
     [[(1, 0, 0), (0, 1, 0), (0, 0, 1)], [(1, 1, 1)]]),
    ((12,), (3,), (60,), (2,), 2, [[(1,), (2,)]]),  # 1-D (direct kernel)
    # value and first derivatives only: the fused single pass
    ((9, 8), (3, 2), (37, 29), (1, 1), 3, [[(0, 0), (1, 0), (0, 1)], [(0, 1)]]),
    ((6, 7, 5), (2, 3, 2), (21, 18, 13), (1, 1, 1), 2,
     [[(0, 0, 0), (0, 0, 1)], [(1, 0, 0), (0, 1, 0), (0, 0, 1), (0, 0, 0)]]),
    ((7, 6, 5), (3, 2, 1), (70, 20, 9), (1, 1, 1), 4, [[(1, 0, 0), (0, 0, 1)]]),
]


@pytest.mark.parametrize("dtype,bar", [(np.float64, 1e-12), (np.float32, 1e-5)])
@pytest.mark.parametrize("case", CASES, ids=lambda c: f"n{c[0]}")
def test_rational_fields_vs_oracle(gpu, case, dtype, bar):
    sg = gpu
    n, p, m, D, nout, field_sets = case
    dims = make_dims(sg, n, p, m, D=D, dtype=dtype)
    grid = sg.NURBSGrid(tuple(dims), nout)
    rng = np.random.default_rng(sum(m) + nout)
    P = rng.uniform(-1, 1, tuple(n) + (nout,)).astype(dtype)
    W = rng.uniform(0.5, 2.0, n).astype(dtype)
    grid.control_points.copy_(sg.from_numpy_colmajor(P, grid.dtype, "cuda"))
    grid.weights.copy_(sg.from_numpy_colmajor(W, grid.dtype, "cuda"))
    od = oracle_dims(dims)
    for fields in field_sets:
        outs = [torch.empty_like(grid.eval) for _ in fields]
        sg.evaluate_rational_fields_(grid, fields, outs)
        for f, out in zip(fields, outs):
            want = O.rational_derivs(od, P.astype(np.float64), W.astype(np.float64), f)
            assert rel_err(eval_host(sg, out), want) < bar, (f, rel_err(eval_host(sg, out), want))


@pytest.mark.parametrize("dtype,bar", [(np.float64, 1e-12), (np.float32, 1e-5)])
def test_rational_fused_first_order_matches_two_pass(gpu, dtype, bar):
    """d/dx1 and d/dx3 from the fused pass against the same fields from the
    two-pass path (forced by adding a mixed second order to the request)."""
    sg = gpu
    n, p, m = (8, 9, 6), (3, 2, 2), (50, 31, 11)
    dims = make_dims(sg, n, p, m, D=(1, 1, 1), dtype=dtype)
    grid = sg.NURBSGrid(tuple(dims), 3)
    rng = np.random.default_rng(11)
    P = rng.uniform(-1, 1, tuple(n) + (3,)).astype(dtype)
    W = rng.uniform(0.5, 2.0, n).astype(dtype)
    grid.control_points.copy_(sg.from_numpy_colmajor(P, grid.dtype, "cuda"))
    grid.weights.copy_(sg.from_numpy_colmajor(W, grid.dtype, "cuda"))
    fused = [torch.empty_like(grid.eval) for _ in range(2)]
    sg.evaluate_rational_fields_(grid, [(1, 0, 0), (0, 0, 1)], fused)
    general = [torch.empty_like(grid.eval) for _ in range(3)]
    sg.evaluate_rational_fields_(grid, [(1, 0, 0), (0, 0, 1), (1, 0, 1)], general)
    for a, b in zip(fused, general[:2]):
        assert rel_err(eval_host(sg, a), eval_host(sg, b)) < bar


def test_rational_order_zero_is_evaluate(gpu):
    sg = gpu
    dims = make_dims(sg, (8, 7), (2, 3), (30, 25), D=(1, 1), dtype=np.float64)
    grid = sg.NURBSGrid(tuple(dims), 2)
    rng = np.random.default_rng(7)
    grid.control_points.copy_(sg.from_numpy_colmajor(rng.uniform(-1, 1, (8, 7, 2)),
                                                     grid.dtype, "cuda"))
    grid.weights.copy_(sg.from_numpy_colmajor(rng.uniform(0.5, 2, (8, 7)), grid.dtype, "cuda"))
    sg.evaluate_(grid)
    out = torch.empty_like(grid.eval)
    sg.evaluate_rational_fields_(grid, [(0, 0)], [out])
    assert float((out - grid.eval).abs().max()) < 1e-13


def test_rational_rejections(gpu):
    sg = gpu
    dims = make_dims(sg, (8, 7), (2, 3), (30, 25), D=(1, 0))
    grid = sg.NURBSGrid(tuple(dims), 1)
    out = torch.empty_like(grid.eval)
    # evaluate_ keeps the reference's behaviour
    with pytest.raises(ValueError, match="Computing derivatives of NURBS"):
        sg.evaluate_(grid, derivative_order=(1, 0))
    with pytest.raises(sg.SplineGridsError, match="maximum derivative order"):
        sg.evaluate_rational_fields_(grid, [(0, 1)], [out])
    bs = sg.SplineGrid(tuple(make_dims(sg, (8, 7), (2, 3), (30, 25))), 1)
    with pytest.raises(AssertionError, match="NURBS grid"):
        sg.evaluate_rational_fields_(bs, [(0, 0)], [torch.empty_like(bs.eval)])


# ---- the adjoint in the control points (weights fixed) ----

ADJ_CASES = [
    ((40, 40, 20), (3, 3, 3), (400, 96, 40), 3),   # matrix-core B-spline adjoint inside
    ((60, 8, 5), (3, 2, 2), (64, 30, 9), 2),       # streaming adjoint inside
    ((9, 8), (3, 2), (37, 29), 3),
    ((12,), (3,), (60,), 2),                       # 1-D: the general adjoint inside
]


@pytest.mark.parametrize("dtype,bar", [(np.float64, 1e-12), (np.float32, 1e-5)])
@pytest.mark.parametrize("case", ADJ_CASES, ids=lambda c: f"n{c[0]}")
def test_rational_adjoint_vs_oracle(gpu, case, dtype, bar):
    sg = gpu
    n, p, m, nout = case
    dims = make_dims(sg, n, p, m, dtype=dtype)
    grid = sg.NURBSGrid(tuple(dims), nout)
    rng = np.random.default_rng(sum(n) + nout)
    W = rng.uniform(0.5, 2.0, n).astype(dtype)
    E = rng.standard_normal(tuple(m) + (nout,)).astype(dtype)
    grid.weights.copy_(sg.from_numpy_colmajor(W, grid.dtype, "cuda"))
    ebar = sg.from_numpy_colmajor(E, grid.dtype, "cuda")
    pbar = sg.colmajor_zeros(tuple(n) + (nout,), grid.dtype, "cuda")
    sg.evaluate_rational_adjoint_(grid, eval=ebar, control_points=pbar)
    want = O.adjoint_rational(oracle_dims(dims), W.astype(np.float64), E.astype(np.float64),
                              nthreads=8)
    assert rel_err(eval_host(sg, pbar), want) < bar
    sg.evaluate_rational_adjoint_(grid, eval=ebar, control_points=pbar, accumulate=True)
    assert rel_err(eval_host(sg, pbar), 2 * want) < bar


def test_nurbs_reverse_mode(gpu):
    """spline_eval on a NURBS grid differentiates through the rational
    adjoint: d/dP sum(R(P) * E) = R^T E."""
    sg = gpu
    n, p, m = (8, 7, 5), (3, 2, 2), (30, 22, 13)
    dims = make_dims(sg, n, p, m, dtype=np.float64)
    grid = sg.NURBSGrid(tuple(dims), 2)
    rng = np.random.default_rng(9)
    W = rng.uniform(0.5, 2.0, n)
    grid.weights.copy_(sg.from_numpy_colmajor(W, grid.dtype, "cuda"))
    P = sg.from_numpy_colmajor(rng.standard_normal(n + (2,)), grid.dtype, "cuda")
    P.requires_grad_(True)
    E = rng.standard_normal(m + (2,))
    out = sg.spline_eval(grid, P)
    (out * sg.from_numpy_colmajor(E, grid.dtype, "cuda")).sum().backward()
    want = O.adjoint_rational(oracle_dims(dims), W, E)
    assert rel_err(sg.to_numpy_colmajor(P.grad), want) < 1e-12
