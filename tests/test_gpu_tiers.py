"""MFMA precision tiers against the FP64 oracle.

split-FP16 (hi+lo operands, three MFMA passes, FP32 accumulation) must meet
the FP32 bar, 1e-6 relative to max|oracle|; BF16 (one pass) is measured and
held to 2e-2 here.  The oracle consumes the same FP32-rounded inputs.
"""
from __future__ import annotations

import itertools

import numpy as np
import pytest
import torch

import oracle as O
from helpers import clamped_knot_vector, eval_host, make_dims, oracle_dims, random_cp, rel_err

pytestmark = pytest.mark.gpu

BAR = {"split_f16": 1e-6, "bf16": 2e-2}

CASES = [
    ((12, 10), (3, 4), (26, 73), 3, False),
    ((40, 37), (2, 3), (257, 191), 1, False),
    ((13, 11), (5, 5), (129, 64), 4, False),
    ((8, 7), (0, 1), (65, 70), 2, False),
    ((10, 9, 8), (2, 3, 2), (70, 33, 21), 3, False),
    ((14, 6, 9), (3, 3, 4), (130, 150, 7), 5, False),
    ((9, 8), (3, 2), (90, 77), 3, True),
    ((9, 8, 7), (2, 3, 2), (40, 33, 29), 3, True),
]


@pytest.mark.parametrize("tier", ["split_f16", "bf16"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: f"n{c[0]}p{c[1]}{'w' if c[4] else ''}")
def test_tier_vs_oracle(gpu, case, tier):
    sg = gpu
    n, p, m, nout, nurbs = case
    D = [0 if nurbs else min(1, pi) for pi in p]
    dims = make_dims(sg, n, p, m, D=D, dtype=np.float32)
    grid = (sg.NURBSGrid if nurbs else sg.SplineGrid)(tuple(dims), nout)
    rng = np.random.default_rng(sum(m))
    P = random_cp(rng, n, nout, np.float32) * 37.0  # off unit scale on purpose
    grid.control_points.copy_(P)
    W = None
    if nurbs:
        W = rng.uniform(0.5, 2.0, n).astype(np.float32)
        grid.weights.copy_(sg.from_numpy_colmajor(W, torch.float32, "cuda"))
    od = oracle_dims(dims)
    for deriv in itertools.product(*[range(d + 1) for d in D]):
        sg.evaluate_(grid, derivative_order=deriv, tier=tier)
        want = O.eval_direct(od, P.astype(np.float64), deriv=deriv,
                             W=None if W is None else W.astype(np.float64))
        err = rel_err(eval_host(sg, grid.eval), want)
        assert err < BAR[tier], (deriv, err)


@pytest.mark.parametrize("tier", ["split_f16", "bf16"])
def test_tier_doubled_knots_derivatives(gpu, tier):
    """Doubled knots make derivative tables large and uneven; the per-order
    and per-window power-of-two scaling keeps FP16 in range."""
    sg = gpu
    rng = np.random.default_rng(5)
    kvs = []
    for _ in range(2):
        vals = np.concatenate([[0.0], np.sort(rng.uniform(0, 1, 30)), [1.0]])
        kvs.append(clamped_knot_vector(sg, vals, 3, np.float32, doubled=(4, 11, 20)))
    n = tuple(len(k.knots_all) - 4 for k in kvs)
    dims = make_dims(sg, n, (3, 3), (400, 300), D=(1, 1), knot_vectors=kvs)
    grid = sg.SplineGrid(tuple(dims), 1)
    P = rng.standard_normal(n + (1,)).astype(np.float32)
    grid.control_points.copy_(P)
    od = oracle_dims(dims)
    fields = [(0, 0), (1, 0), (0, 1), (1, 1)]
    outs = [torch.empty_like(grid.eval) for _ in fields]
    sg.evaluate_fields_(grid, fields, outs, tier=tier)
    for f, out in zip(fields, outs):
        want = O.eval_direct(od, P.astype(np.float64), deriv=f)
        assert rel_err(eval_host(sg, out), want) < BAR[tier], f


def test_tier_rejections(gpu):
    sg = gpu
    dims = make_dims(sg, (6, 5), (2, 2), (20, 20), dtype=np.float64)
    grid = sg.SplineGrid(tuple(dims), 1)
    with pytest.raises(sg.SplineGridsError, match="FP32 arrays"):
        sg.evaluate_(grid, tier="split_f16")
    # 64 samples of dimension 1 over 60 control points: windows wider than
    # 16, so no matrix-core adjoint and no adjoint tiers
    dims = make_dims(sg, (60, 8), (3, 2), (64, 30))
    grid = sg.SplineGrid(tuple(dims), 1)
    with pytest.raises(sg.SplineGridsError, match="matrix-core adjoint"):
        sg.evaluate_adjoint_(grid, tier="split_f16")
    d1 = make_dims(sg, (6,), (2,), (20,))
    with pytest.raises(sg.SplineGridsError, match="Nin 2-3"):
        sg.evaluate_(sg.SplineGrid(tuple(d1), 1), tier="bf16")
