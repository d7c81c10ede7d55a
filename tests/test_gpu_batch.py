"""Batched persistent launch: many uneven grids, one launch.

Every grid of the batch is checked against the oracle and against its own
single-grid launch; the tuned tiling must not change results; a captured
hipGraph (torch.cuda.CUDAGraph) is replayed several times with eager work
in between and every replay is checked.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import bench
import oracle as O
from helpers import eval_host, oracle_dims, rel_err

pytestmark = pytest.mark.gpu


def _batch_grids(sg, rng, n=40):
    grids = bench.c5_specs(rng, n, small=True)
    # edge cases: degree 0 / 5 windows, a degree > 5 grid and unsorted samples
    # (both evaluated with their own launches), many channels
    grids += [
        dict(n=(9, 7), p=(0, 5), m=(70, 50), nout=2, nurbs=False),
        dict(n=(12, 9), p=(7, 2), m=(40, 30), nout=1, nurbs=False),
        dict(n=(10, 8, 6), p=(2, 3, 1), m=(33, 20, 9), nout=6, nurbs=False),
        # 3-D plane lists: p3 = 0 and 5, fewer samples than control points
        # along dimension 3 (disjoint bands), ragged groups of four planes,
        # NURBS over two channel blocks
        dict(n=(9, 12, 7), p=(1, 2, 0), m=(20, 17, 7), nout=3, nurbs=True),
        dict(n=(8, 10, 40), p=(4, 1, 5), m=(30, 22, 13), nout=5, nurbs=True),
        dict(n=(8, 8), p=(3, 3), m=(65, 40), nout=3, nurbs=True, unsorted=True),
        # rows of 65..128 samples: one tile per row, two samples per lane
        # (odd rows end on a single sample; NURBS in blocks of three channels;
        # five channels in blocks of four); degree 5: one sample per lane
        dict(n=(20, 9, 6), p=(3, 2, 2), m=(65, 17, 9), nout=3, nurbs=False),
        dict(n=(30, 8, 7), p=(4, 4, 3), m=(128, 12, 6), nout=4, nurbs=True),
        dict(n=(25, 7, 5), p=(2, 1, 1), m=(101, 10, 5), nout=1, nurbs=True),
        dict(n=(16, 6, 5), p=(5, 2, 2), m=(90, 9, 5), nout=2, nurbs=False),
        dict(n=(40, 10, 6), p=(1, 3, 0), m=(77, 19, 7), nout=5, nurbs=False),
        # 2-D rows of 301 samples: three tiles of 128 (the last 45 wide, odd)
        dict(n=(30, 20), p=(3, 4), m=(301, 45), nout=2, nurbs=True),
    ]
    return [bench.build_spec_grid(sg, spec, rng) for spec in grids]


def _check_against_oracle(sg, grid, tol=1e-6):
    P = sg.to_numpy_colmajor(sg.obtain(grid.control_points)).astype(np.float64)
    W = None if grid.weights is None else sg.to_numpy_colmajor(grid.weights).astype(np.float64)
    want = O.eval_direct(oracle_dims(grid.spline_dimensions), P, W=W, nthreads=8)
    return rel_err(eval_host(sg, grid.eval), want)


def test_batch_matches_oracle_and_single_launches(gpu):
    sg = gpu
    rng = np.random.default_rng(21)
    grids = _batch_grids(sg, rng)
    batch = sg.SplineGridBatch(grids)
    info = batch.info()
    assert info["grids"] == len(grids) and info["persistent_grids"] == len(grids) - 2
    batch.evaluate_()
    torch.cuda.synchronize()
    got = [g.eval.clone() for g in grids]
    for i, g in enumerate(grids):
        assert _check_against_oracle(sg, g) < 1e-6, i
    for g in grids:
        g.eval.zero_()
    batch.evaluate_each_()
    for g, ref in zip(grids, got):
        assert float((g.eval - ref).abs().max()) <= 1e-6 * max(1.0, float(ref.abs().max()))


def test_batch_autotune_keeps_results(gpu):
    sg = gpu
    rng = np.random.default_rng(22)
    grids = _batch_grids(sg, rng, n=24)
    batch = sg.SplineGridBatch(grids)
    batch.evaluate_()
    torch.cuda.synchronize()
    ref = [g.eval.clone() for g in grids]
    info = batch.autotune(reps=2)
    assert info["rows_per_wave"] in (8, 16, 32) and info["blocks_per_cu"] in (2, 3, 4, 6)
    for g in grids:
        g.eval.zero_()
    batch.evaluate_()
    torch.cuda.synchronize()
    for g, r in zip(grids, ref):
        assert float((g.eval - r).abs().max()) <= 1e-6 * max(1.0, float(r.abs().max()))


def test_batch_graph_replays(gpu):
    sg = gpu
    rng = np.random.default_rng(23)
    grids = [g for g in _batch_grids(sg, rng, n=24)]
    batch = sg.SplineGridBatch(grids)
    batch.evaluate_()
    torch.cuda.synchronize()
    ref = [g.eval.clone() for g in grids]
    stream = torch.cuda.Stream()
    stream.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(stream):
        batch.evaluate_()  # warm-up on the capture stream
        torch.cuda.synchronize()
        with torch.cuda.graph(graph, stream=stream):
            batch.evaluate_()
    torch.cuda.synchronize()
    for rep in range(3):
        for g in grids:
            g.eval.fill_(float(rep + 7))  # eager work between replays
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        for i, (g, r) in enumerate(zip(grids, ref)):
            assert torch.equal(g.eval, r), (rep, i)
    # the captured launch keeps its tiling: retuning is refused from now on
    with pytest.raises(sg.SplineGridsError, match="autotune before capturing"):
        batch.autotune(reps=1)


def test_batch_rejects_bad_input(gpu):
    sg = gpu
    rng = np.random.default_rng(24)
    g64 = bench.build_spec_grid(sg, dict(n=(6, 5), p=(2, 2), m=(20, 20), nout=1, nurbs=False),
                                rng, dtype=np.float64)
    with pytest.raises(TypeError, match="FP32"):
        sg.SplineGridBatch([g64])
    # the edge-case grids (degree 7, unsorted samples) cannot take the MFMA tiers
    grids = _batch_grids(sg, rng, n=4)
    batch = sg.SplineGridBatch(grids)
    assert "mfma_tiers=0" in batch.describe()
    with pytest.raises(sg.SplineGridsError, match="every grid of the batch"):
        batch.evaluate_(tier="split_f16")


@pytest.mark.parametrize("tier", ["split_f16", "bf16"])
def test_batch_mfma_tiers(gpu, tier):
    """The split-FP16 and BF16 tiers over a batch: the persistent launch gives
    every grid exactly its own tier launch's result (same tile code), and the
    oracle bound of the tier."""
    sg = gpu
    rng = np.random.default_rng(25)
    specs = bench.c5_specs(rng, 40, small=True) + [
        dict(n=(9, 12, 7), p=(1, 2, 0), m=(20, 17, 7), nout=3, nurbs=True),
        dict(n=(8, 10, 12), p=(4, 1, 5), m=(30, 22, 13), nout=5, nurbs=True),
        dict(n=(10, 8, 6), p=(2, 3, 1), m=(33, 20, 9), nout=6, nurbs=False),
        dict(n=(9, 7), p=(0, 5), m=(70, 50), nout=2, nurbs=False),
    ]
    grids = [bench.build_spec_grid(sg, spec, rng) for spec in specs]
    batch = sg.SplineGridBatch(grids)
    assert "mfma_tiers=1" in batch.describe(), batch.describe()
    batch.evaluate_(tier=tier)
    torch.cuda.synchronize()
    got = [g.eval.clone() for g in grids]
    for g in grids:
        g.eval.zero_()
        sg.evaluate_(g, tier=tier)
    torch.cuda.synchronize()
    for i, (g, r) in enumerate(zip(grids, got)):
        assert torch.equal(g.eval, r), i
    bar = 1e-6 if tier == "split_f16" else 2e-2
    for i, g in enumerate(grids):
        assert _check_against_oracle(sg, g) < bar, i
