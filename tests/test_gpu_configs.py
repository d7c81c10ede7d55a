"""The benchmark workloads at full size (bench.py's c1..c5).

The oracle checks random samples of every field (the full direct sum would
take minutes on the host); size-independent properties cover the rest:
partition of unity, unit weights = B-spline, the transpose identity, central
finite differences of the 400 x 400 x 200 least-squares loss, determinism.
"""
from __future__ import annotations

import math

import numpy as np
import pytest
import torch

import bench
import oracle as O
from helpers import eval_host, oracle_dims, rel_err

pytestmark = pytest.mark.gpu


def _sampled_err(sg, grid, wl, out, deriv, n_points=30000, seed=0):
    dims = oracle_dims(grid.spline_dimensions)
    rng = np.random.default_rng(seed)
    lin = rng.integers(0, math.prod(wl["m"]), n_points)
    P = wl["P"].astype(wl["dtype"]).astype(np.float64)
    W = None if wl["W"] is None else wl["W"].astype(wl["dtype"]).astype(np.float64)
    want = O.eval_points(dims, P, lin, deriv=deriv, W=W, nthreads=8)
    flat = out.permute(*range(out.ndim - 1, -1, -1)).reshape(wl["nout"], -1).t()
    got = flat[torch.as_tensor(lin, device=out.device)].double().cpu().numpy()
    return float(np.abs(got - want).max() / out.abs().max().item())


def test_c1_readme_shape_fp64(gpu):
    sg = gpu
    wl = bench.workload("c1")
    grid = bench.build_grid(sg, wl)
    sg.evaluate_(grid)
    want = O.eval_direct(oracle_dims(grid.spline_dimensions), wl["P"], nthreads=8)
    assert rel_err(eval_host(sg, grid.eval), want) < 1e-12


def test_c2_freeform_surface_fields(gpu):
    sg = gpu
    wl = bench.workload("c2")
    grid = bench.build_grid(sg, wl)
    outs = [torch.empty_like(grid.eval) for _ in wl["fields"]]
    sg.evaluate_fields_(grid, wl["fields"], outs)
    for i, (f, out) in enumerate(zip(wl["fields"], outs)):
        assert _sampled_err(sg, grid, wl, out, f, seed=i) < 1e-6, f
    again = torch.empty_like(grid.eval)
    sg.evaluate_(grid, derivative_order=(1, 1), eval=again)
    assert torch.equal(again, outs[3])  # fused and single launches agree bit for bit
    grid.control_points[...] = 1  # partition of unity at full size
    sg.evaluate_(grid)
    assert float((grid.eval - 1).abs().max()) < 1e-6


@pytest.mark.parametrize("tier,bar", [("native", 1e-6), ("split_f16", 1e-6), ("bf16", 2e-2)])
def test_c3_nurbs_tiers(gpu, tier, bar):
    sg = gpu
    wl = bench.workload("c3")
    grid = bench.build_grid(sg, wl)
    sg.evaluate_(grid, tier=tier)
    assert _sampled_err(sg, grid, wl, grid.eval, (0, 0, 0)) < bar


def test_c3_unit_weights_equal_bspline(gpu):
    sg = gpu
    wl = bench.workload("c3")
    grid = bench.build_grid(sg, wl)
    grid.weights.fill_(1)
    sg.evaluate_(grid)
    wl_b = bench.workload("c3n")
    plain = bench.build_grid(sg, wl_b)
    sg.evaluate_(plain)
    assert float((grid.eval - plain.eval).abs().max()) < 1e-6


def test_c4_adjoint_finite_differences(gpu):
    """Loss L = 1/2 sum (eval - target)^2, target = a second seeded spline plus
    N(0, 0.01) noise.  dL/dP from the FP32 adjoint is checked against central
    differences (h = 1e-3, 64 random entries) with the losses evaluated on the
    FP64 path, and the FP64 adjoint against the transpose identity."""
    sg = gpu
    wl = bench.workload("c4")
    rng = np.random.default_rng(bench.SEED0 + 40)
    grid = bench.build_grid(sg, wl)
    P2 = rng.uniform(-1, 1, wl["P"].shape).astype(np.float32)
    target = sg.colmajor_empty(grid.eval.shape, torch.float32, "cuda")
    sg.evaluate_(grid, control_points=sg.from_numpy_colmajor(P2, torch.float32, "cuda"),
                 eval=target)
    target += 0.01 * torch.randn(target.shape, device="cuda", generator=torch.Generator(
        device="cuda").manual_seed(4))
    sg.evaluate_(grid)
    ebar = sg.colmajor(grid.eval - target)
    grad = sg.colmajor_zeros(grid.control_points.shape, torch.float32, "cuda")
    sg.evaluate_adjoint_(grid, eval=ebar, control_points=grad)

    # the same map in FP64: FP32 knots and sample points promoted exactly
    dims64 = tuple(sg.SplineDimension(
        s.n_basis_functions, s.degree, s.n_sample_points, float_type=np.float64,
        knot_vector=sg.KnotVector(s.knot_vector.knot_values.astype(np.float64),
                                  s.knot_vector.multiplicities),
        sample_points=s.host_sample_points(), device="cuda") for s in grid.spline_dimensions)
    g64 = sg.SplineGrid(dims64, wl["nout"])
    t64 = target.double()
    P0 = sg.from_numpy_colmajor(wl["P"].astype(np.float64), torch.float64, "cuda")
    ev = sg.colmajor_empty(g64.eval.shape, torch.float64, "cuda")

    def loss(P):
        sg.evaluate_(g64, control_points=P, eval=ev)
        return 0.5 * float(((ev - t64) ** 2).sum())

    h = 1e-3
    flat0 = sg.vec_colmajor(P0)
    gflat = sg.vec_colmajor(grad).double().cpu().numpy()
    scale = np.abs(gflat).max()
    for q in rng.integers(0, flat0.numel(), 64):
        e = torch.zeros_like(flat0)
        e[q] = h
        fd = (loss(sg.reshape_colmajor(flat0 + e, P0.shape)) -
              loss(sg.reshape_colmajor(flat0 - e, P0.shape))) / (2 * h)
        assert abs(fd - gflat[q]) <= 2e-5 * scale, (q, fd, gflat[q])

    x = sg.from_numpy_colmajor(rng.standard_normal(wl["P"].shape), torch.float64, "cuda")
    y = sg.colmajor(torch.randn(g64.eval.shape, dtype=torch.float64, device="cuda"))
    Mx = sg.colmajor_empty(g64.eval.shape, torch.float64, "cuda")
    MTy = sg.colmajor_zeros(x.shape, torch.float64, "cuda")
    sg.evaluate_(g64, control_points=x, eval=Mx)
    sg.evaluate_adjoint_(g64, eval=y, control_points=MTy)
    lhs, rhs = float((Mx * y).sum()), float((x * MTy).sum())
    assert abs(lhs - rhs) <= 1e-11 * abs(lhs)
