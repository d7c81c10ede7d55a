"""The FP32 adjoint paths and the matrix-core adjoint's tiers against the oracle.

FP32 adjoints run on matrix cores (adj_mma) whenever every 64-sample tile of
dimension 1 and every chunk of 16..64 rows of dimension 2 has a band window
of at most 16 control points; otherwise, and with SG_ADJ_MMA=0 set when the
plan is built, on the streaming kernel.  The cases cover every chunk height,
rows whose length is not a multiple of 4 samples (no 16-byte loads), ragged
last tiles and chunks, 1..6 channels (channel blocks of 4), p3 = 0..5, 2-D
grids, derivative orders, accumulation, and a grid too sparse along
dimension 1 for the matrix-core tiles.  The split-FP16 and BF16 tiers run
the first product of the matrix-core adjoint (ebar times B1) on FP16 / BF16
MFMA and are held to the FP32 bar of 1e-6 relative to max|oracle| (split
FP16) or 1e-2 (BF16, stated).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from helpers import eval_host, make_dims, oracle_dims, rel_err

pytestmark = pytest.mark.gpu

# (n, p, m, nout, matrix-core path expected)
CASES = [
    ((40, 40, 20), (3, 3, 3), (400, 96, 40), 3, True),     # chunks 64 + 32 rows
    ((30, 50, 12), (2, 4, 5), (133, 70, 33), 2, True),     # m1 % 4 != 0, p3 = 5
    ((20, 100, 9), (3, 3, 0), (128, 300, 17), 6, True),    # short chunks, 4 + 2 channels
    ((12, 10), (1, 2), (256, 90), 1, True),                # 2-D
    ((16, 16, 16), (2, 2, 1), (64, 64, 64), 4, True),      # p3 = 1
    ((9, 9, 30), (3, 3, 4), (70, 50, 200), 5, True),       # p3 = 4, 4 + 1 channels
    ((8, 8, 6), (2, 1, 2), (19, 17, 11), 1, True),         # one small ragged tile
    ((60, 8, 5), (3, 2, 2), (64, 30, 9), 2, False),        # W1 > 16: streaming kernel
]


def _run(sg, case, deriv, mma, monkeypatch, accumulate=False, tier="native"):
    n, p, m, nout, _ = case
    if not mma:
        monkeypatch.setenv("SG_ADJ_MMA", "0")
    D = [min(1, pi) for pi in p]
    dims = make_dims(sg, n, p, m, D=D, dtype=np.float32)
    grid = sg.SplineGrid(tuple(dims), nout)
    rng = np.random.default_rng(sum(m) + nout)
    E = rng.standard_normal(tuple(m) + (nout,)).astype(np.float32)
    ebar = sg.from_numpy_colmajor(E, grid.dtype, "cuda")
    Pbar = sg.colmajor_zeros(tuple(n) + (nout,), grid.dtype, "cuda")
    sg.evaluate_adjoint_(grid, derivative_order=deriv(D), eval=ebar, control_points=Pbar, tier=tier)
    desc = grid.plan().describe()
    monkeypatch.delenv("SG_ADJ_MMA", raising=False)
    if accumulate:
        sg.evaluate_adjoint_(grid, derivative_order=deriv(D), eval=ebar, control_points=Pbar,
                             accumulate=True, tier=tier)
    want = O.adjoint_direct(oracle_dims(dims), E.astype(np.float64), deriv=deriv(D), nthreads=8)
    return eval_host(sg, Pbar), want * (2 if accumulate else 1), desc


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"n{c[0]}m{c[2]}c{c[3]}")
@pytest.mark.parametrize("order", ["zero", "deriv"])
def test_adjoint_paths_vs_oracle(gpu, case, order, monkeypatch):
    sg = gpu
    deriv = (lambda D: tuple(0 for _ in D)) if order == "zero" else (lambda D: tuple(D))
    got_m, want, desc_m = _run(sg, case, deriv, True, monkeypatch)
    assert ("adjoint=mma_f32" in desc_m) == case[4], desc_m
    assert rel_err(got_m, want) < 1e-6, desc_m
    got_s, want_s, desc_s = _run(sg, case, deriv, False, monkeypatch)
    assert "adjm=none" in desc_s and "adjoint=stream " in desc_s, desc_s
    assert rel_err(got_s, want_s) < 1e-6, desc_s


TIER_BAR = {"split_f16": 1e-6, "bf16": 1e-2}


@pytest.mark.parametrize("case", [c for c in CASES if c[4]], ids=lambda c: f"n{c[0]}m{c[2]}c{c[3]}")
@pytest.mark.parametrize("tier", ["split_f16", "bf16"])
def test_adjoint_tiers_vs_oracle(gpu, case, tier, monkeypatch):
    deriv = lambda D: tuple(D)  # noqa: E731
    got, want, desc = _run(gpu, case, deriv, True, monkeypatch, tier=tier)
    assert "adjoint=mma_f32" in desc
    assert rel_err(got, want) < TIER_BAR[tier], (tier, rel_err(got, want))


def test_adjoint_split_f16_tiny_and_huge_ebar(gpu, monkeypatch):
    """ebar of 1e-30 and 1e30 (far outside FP16's range) keeps the split
    FP16 tier at the FP32 bar: each unit is scaled by a power of two."""
    sg = gpu
    n, p, m, nout, _ = CASES[0]
    dims = make_dims(sg, n, p, m, dtype=np.float32)
    grid = sg.SplineGrid(tuple(dims), nout)
    rng = np.random.default_rng(5)
    for scale in (1e-30, 1e30):
        E = (rng.standard_normal(tuple(m) + (nout,)) * scale).astype(np.float32)
        ebar = sg.from_numpy_colmajor(E, grid.dtype, "cuda")
        Pbar = sg.colmajor_zeros(tuple(n) + (nout,), grid.dtype, "cuda")
        sg.evaluate_adjoint_(grid, eval=ebar, control_points=Pbar, tier="split_f16")
        want = O.adjoint_direct(oracle_dims(dims), E.astype(np.float64), nthreads=8)
        assert rel_err(eval_host(sg, Pbar), want) < 1e-6, scale


def test_adjoint_mma_accumulates(gpu, monkeypatch):
    got, want, desc = _run(gpu, CASES[0], lambda D: (0, 0, 0), True, monkeypatch,
                           accumulate=True)
    assert "adjoint=mma_f32" in desc
    assert rel_err(got, want) < 1e-6


def test_adjoint_mma_is_deterministic(gpu):
    sg = gpu
    n, p, m, nout, _ = CASES[1]
    dims = make_dims(sg, n, p, m, dtype=np.float32)
    grid = sg.SplineGrid(tuple(dims), nout)
    E = np.random.default_rng(3).standard_normal(tuple(m) + (nout,)).astype(np.float32)
    ebar = sg.from_numpy_colmajor(E, grid.dtype, "cuda")
    outs = []
    for _ in range(3):
        Pbar = sg.colmajor_zeros(tuple(n) + (nout,), grid.dtype, "cuda")
        sg.evaluate_adjoint_(grid, eval=ebar, control_points=Pbar)
        outs.append(eval_host(sg, Pbar))
    assert all(np.array_equal(outs[0], o) for o in outs[1:])
