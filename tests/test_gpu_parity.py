"""HIP path against the oracle and the reference's golden vectors.

Mirrors the reference's hot-path tests (test/test_spline_dimension.jl,
test/test_spline_grid.jl, test/test_nurbs_grid.jl,
test/test_local_refinement.jl:48-67) and adds oracle parity over degrees,
derivative orders, dimensions, NURBS, value types and sample orders.
Bars: FP64 <= 1e-12, FP32 <= 1e-6, relative to max|oracle| of the field.
"""
from __future__ import annotations

import itertools
import logging

import numpy as np
import pytest
import torch

import oracle as O
from helpers import (TOL, clamped_knot_vector, eval_host, make_dims, oracle_dims,
                     random_cp, rel_err)

pytestmark = pytest.mark.gpu


# --------------------------------------------------------------------------
# basis tables (a4, a5)
# --------------------------------------------------------------------------

@pytest.mark.parametrize("p,D", [(0, 0), (1, 1), (2, 2), (3, 1), (3, 3), (5, 2), (7, 3)])
@pytest.mark.parametrize("random", [False, True])
def test_basis_table_bit_exact(gpu, p, D, random):
    sg = gpu
    rng = np.random.default_rng(p * 10 + D)
    kw = {"distribution": "random", "rng": rng} if random else {}
    s = sg.SplineDimension(25, p, 301, max_derivative_order=D, float_type=np.float64,
                           device="cuda", **kw)
    od = O.Dim(s.host_knots(), s.host_sample_points(), p, D)
    idx = s.sample_indices.cpu().numpy()
    assert np.array_equal(idx, od.idx1)
    table = eval_host(sg, s.eval)
    assert np.array_equal(table, od.table)  # same recursion, same roundings


def test_basis_doubled_knots_f32(gpu):
    sg = gpu
    rng = np.random.default_rng(7)
    vals = np.concatenate([[0.0], np.sort(rng.uniform(0, 1, 20)), [1.0]])
    kv = clamped_knot_vector(sg, vals, 3, np.float32, doubled=(3, 9, 15))
    n = len(kv.knots_all) - 4
    s = sg.SplineDimension(n, 3, 777, max_derivative_order=2, knot_vector=kv, device="cuda")
    od = O.Dim(s.host_knots(), s.host_sample_points(), 3, 2)
    assert np.array_equal(s.sample_indices.cpu().numpy(), od.idx1)
    table = eval_host(sg, s.eval)
    assert np.array_equal(table, od.table.astype(np.float32))  # FP64 then one rounding


@pytest.mark.parametrize("degree", range(6))
def test_partition_of_unity(gpu, degree):
    # test/test_spline_dimension.jl:10-29
    sg = gpu
    for kw in ({}, {"distribution": "random", "rng": np.random.default_rng(degree)}):
        s = sg.SplineDimension(25, degree, 500, device="cuda", **kw)
        B = s.eval[:, :, 0]
        assert tuple(B.shape) == (500, degree + 1)
        assert bool((B >= 0).all())
        torch.testing.assert_close(B.sum(dim=1), torch.ones(500, device="cuda"),
                                   rtol=1e-6, atol=1e-6)


def test_fd_derivatives(gpu):
    # test/test_spline_dimension.jl:31-66
    sg = gpu
    s = sg.SplineDimension(10, 3, 5000, max_derivative_order=2, float_type=np.float64,
                           device="cuda")
    t = s.sample_points.double()
    dt = torch.diff(t)
    B0, B1, B2 = (sg.decompress(s, derivative_order=o) for o in range(3))
    fd1 = torch.diff(B0, dim=0) / dt[:, None]
    fd2 = torch.diff(fd1, dim=0) / dt[1:, None]
    torch.testing.assert_close(B1[:-1], fd1, rtol=1e-2, atol=1e-2 * fd1.abs().max().item())
    torch.testing.assert_close(B2[2:], fd2, rtol=1e-2, atol=1e-2 * fd2.abs().max().item())


@pytest.mark.parametrize("p", [0, 1, 2, 3, 5])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_basis_samples_on_knots_and_outside_extent(gpu, p, dtype):
    """Span-search ties and the clamp (src/util_kernels.jl:22-49): samples
    placed exactly on every knot (interior, repeated and the end knots, so
    ties go right) and samples left and right of the knot extent (clamped
    to the first / last non-empty span).  Indices and FP64 tables bit-exact
    against the oracle."""
    sg = gpu
    rng = np.random.default_rng(100 + p)
    inner = np.sort(rng.uniform(0.05, 0.95, 9))
    vals = np.concatenate([[0.0], inner, [1.0]])
    mults = np.ones(len(vals), dtype=np.int32)
    mults[0] = mults[-1] = p + 1
    if p > 0:
        mults[4] = min(p, 2)  # a repeated interior knot
        mults[7] = p          # a knot of full interior multiplicity
    kv = sg.KnotVector(vals.astype(dtype), mults)
    knots = np.asarray(kv.knots_all, dtype=np.float64)
    t = np.concatenate([knots, [-0.5, -1e-3, 1.0 + 1e-3, 1.7], rng.uniform(-0.2, 1.2, 40)])
    t = t.astype(dtype)
    n = len(knots) - p - 1
    D = min(p, 2)
    s = sg.SplineDimension(n, p, len(t), max_derivative_order=D, knot_vector=kv,
                           sample_points=t, device="cuda")
    od = O.Dim(s.host_knots(), s.host_sample_points(), p, D)
    want_idx = np.clip(np.searchsorted(knots, t.astype(np.float64), side="right"), p + 1,
                       len(knots) - p - 1)
    assert np.array_equal(od.idx1, want_idx)  # the oracle's span rule by brute force
    assert np.array_equal(s.sample_indices.cpu().numpy(), od.idx1)
    assert np.array_equal(eval_host(sg, s.eval), od.table.astype(dtype))
    grid = sg.SplineGrid(s, 1)
    table, idx = grid.plan().basis(0)  # the plan's tables take the same path
    assert np.array_equal(idx, od.idx1)
    assert np.array_equal(table, od.table)


def test_edited_sample_points_rebuild_the_plan(gpu):
    """src/utils.jl:19-29, src/spline_dimension.jl:231-242: a caller may
    overwrite a dimension's sample points and re-evaluate it; the next grid
    evaluation must use the new samples (indices, tables and plan)."""
    sg = gpu
    rng = np.random.default_rng(31)
    dims = make_dims(sg, (11, 9), (3, 2), (60, 45), dtype=np.float64)
    grid = sg.SplineGrid(tuple(dims), 2)
    P = random_cp(rng, (11, 9), 2, np.float64)
    grid.control_points.copy_(P)
    sg.evaluate_(grid)
    before = grid.plan()
    new_t = np.sort(rng.uniform(0, 1, 60)) ** 2
    dims[0].sample_points.copy_(torch.from_numpy(new_t))
    sg.evaluate_(dims[0])
    od = oracle_dims(dims)
    assert np.array_equal(od[0].t, new_t)
    assert np.array_equal(dims[0].sample_indices.cpu().numpy(), od[0].idx1)
    sg.evaluate_(grid)
    assert grid.plan() is not before
    want = O.eval_direct(od, P)
    assert rel_err(eval_host(sg, grid.eval), want) < 1e-12


def test_plan_tables_match_dimension(gpu):
    sg = gpu
    dims = make_dims(sg, (9, 8), (3, 2), (50, 40), D=(2, 1), dtype=np.float64)
    grid = sg.SplineGrid(tuple(dims), 1)
    for i, s in enumerate(dims):
        table, idx = grid.plan().basis(i)
        assert np.array_equal(idx, s.sample_indices.cpu().numpy())
        assert np.array_equal(table, eval_host(sg, s.eval))


# --------------------------------------------------------------------------
# forward (a10, a11)
# --------------------------------------------------------------------------

def test_g1_golden(gpu, goldens):
    # test/test_spline_grid.jl:29-53
    sg = gpu
    g = goldens["G1"]
    n, p, m = g["n_control_points"], g["degree"], g["n_sample_points"]
    P = np.array(g["control_points_colmajor"]).reshape(n + [2], order="F")
    E = np.array(g["eval_colmajor"]).reshape(m + [2], order="F")
    for dtype, tol in ((np.float64, 1e-14), (np.float32, 1e-6)):
        dims = [sg.SplineDimension(n[i], p[i], m[i], float_type=dtype, device="cuda")
                for i in range(2)]
        grid = sg.SplineGrid(tuple(dims), 2)
        grid.control_points.copy_(P.astype(dtype))
        sg.evaluate_(grid)
        assert rel_err(eval_host(sg, grid.eval), E) < tol, dtype


def test_constant_control_points(gpu):
    # test/test_spline_grid.jl:11-27
    sg = gpu
    for n in range(2, 11):
        for p in range(1, n):
            s = sg.SplineDimension(n, p, 100, device="cuda")
            grid = sg.SplineGrid(s, 1)
            grid.control_points[...] = 1
            sg.evaluate_(grid)
            torch.testing.assert_close(grid.eval, torch.ones_like(grid.eval),
                                       rtol=1e-6, atol=1e-6)


CASES = [
    # (n, p, m, nout)
    ((17,), (3,), (301,), 2),
    ((10,), (9,), (50,), 1),
    ((12, 10), (3, 4), (26, 73), 3),
    ((40, 37), (2, 3), (257, 191), 1),
    ((8, 7), (0, 1), (65, 70), 2),
    ((13, 11), (5, 5), (129, 64), 4),
    ((11, 9), (7, 2), (80, 60), 1),       # degree > 5 in dim 1: direct kernel
    ((10, 9, 8), (2, 3, 2), (70, 33, 21), 3),
    ((14, 6, 9), (3, 3, 4), (130, 150, 7), 5),
    ((6, 5, 5, 4), (2, 1, 2, 1), (9, 8, 7, 6), 2),  # Nin = 4: direct kernel
]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("case", CASES, ids=lambda c: f"n{c[0]}p{c[1]}")
def test_forward_vs_oracle(gpu, case, dtype):
    sg = gpu
    n, p, m, nout = case
    D = [min(1, pi) for pi in p]
    dims = make_dims(sg, n, p, m, D=D, dtype=dtype)
    grid = sg.SplineGrid(tuple(dims), nout)
    rng = np.random.default_rng(sum(n))
    P = random_cp(rng, n, nout, dtype)
    grid.control_points.copy_(P)
    od = oracle_dims(dims)
    for deriv in itertools.product(*[range(d + 1) for d in D]):
        sg.evaluate_(grid, derivative_order=deriv)
        want = O.eval_direct(od, P.astype(np.float64), deriv=deriv)
        err = rel_err(eval_host(sg, grid.eval), want)
        assert err < TOL[dtype], (deriv, err)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("nin", [1, 2, 3])
def test_nurbs_vs_oracle(gpu, nin, dtype):
    sg = gpu
    n, p, m, nout = [(13,), (9, 8), (9, 8, 7)][nin - 1], [(2,), (3, 2), (2, 3, 2)][nin - 1], \
        [(301,), (90, 77), (40, 33, 29)][nin - 1], 3
    dims = make_dims(sg, n, p, m, dtype=dtype)
    grid = sg.NURBSGrid(tuple(dims), nout)
    rng = np.random.default_rng(nin)
    P = random_cp(rng, n, nout, dtype)
    W = rng.uniform(0.5, 2.0, n).astype(dtype)
    grid.control_points.copy_(P)
    grid.weights.copy_(sg.from_numpy_colmajor(W, grid.dtype, "cuda"))
    sg.evaluate_(grid)
    want = O.eval_direct(oracle_dims(dims), P.astype(np.float64), W=W.astype(np.float64))
    assert rel_err(eval_host(sg, grid.eval), want) < TOL[dtype]


def test_g4_nurbs_circle(gpu, goldens):
    # test/test_nurbs_grid.jl:25-52
    sg = gpu
    g = goldens["G4"]
    kv = sg.KnotVector(np.array(g["knot_values"], dtype=np.float32),
                       np.array(g["multiplicities"], dtype=np.int32))
    s = sg.SplineDimension(g["n_control_points"], g["degree"], g["n_sample_points"],
                           knot_vector=kv, device="cuda")
    grid = sg.NURBSGrid(s, 2)
    grid.weights[1::2] = 1 / np.sqrt(2)
    grid.control_points.copy_(np.array(g["control_points"], dtype=np.float32))
    sg.evaluate_(grid)
    pts = eval_host(sg, grid.eval).astype(np.float64)
    assert len({tuple(x) for x in pts[1:]}) == len(pts) - 1
    np.testing.assert_allclose(pts[:, 0] ** 2 + pts[:, 1] ** 2, 1.0, rtol=1e-6)


def test_unsorted_samples_use_general_kernel(gpu):
    sg = gpu
    rng = np.random.default_rng(11)
    n, p, m = (12, 9), (3, 2), (57, 41)
    samples = [rng.permutation(np.linspace(0, 1, m[0])), rng.uniform(0, 1, m[1])]
    dims = make_dims(sg, n, p, m, D=(1, 1), dtype=np.float64, sample_points=samples)
    grid = sg.SplineGrid(tuple(dims), 2)
    P = random_cp(rng, n, 2, np.float64)
    grid.control_points.copy_(P)
    sg.evaluate_(grid, derivative_order=(1, 1))
    want = O.eval_direct(oracle_dims(dims), P, deriv=(1, 1))
    assert rel_err(eval_host(sg, grid.eval), want) < 1e-12


@pytest.mark.parametrize("nin", [2, 3])
def test_fields_fused_equal_separate(gpu, nin):
    sg = gpu
    n, p, m = [(20, 18), (10, 9, 8)][nin - 2], [(3, 3), (3, 2, 2)][nin - 2], \
        [(300, 257), (70, 50, 20)][nin - 2]
    dims = make_dims(sg, n, p, m, D=[1] * nin, dtype=np.float32)
    grid = sg.SplineGrid(tuple(dims), 1)
    rng = np.random.default_rng(3)
    grid.control_points.copy_(random_cp(rng, n, 1, np.float32))
    orders = list(itertools.product(*[range(2)] * nin))[:4]
    outs = [torch.empty_like(grid.eval) for _ in orders]
    sg.evaluate_fields_(grid, orders, outs)
    for o, out in zip(orders, outs):
        sg.evaluate_(grid, derivative_order=o)
        assert torch.equal(out, grid.eval), o


def test_empty_sample_grid(gpu):
    sg = gpu
    dims = make_dims(sg, (6, 5), (2, 2), (0, 7))
    grid = sg.SplineGrid(tuple(dims), 2)
    sg.evaluate_(grid)
    assert grid.eval.numel() == 0
    out = sg.colmajor(torch.full((6, 5, 2), 7.0, device="cuda"))
    sg.evaluate_adjoint_(grid, control_points=out)  # no samples: M^T e = 0
    assert bool((out == 0).all())


# --------------------------------------------------------------------------
# adjoint (a12)
# --------------------------------------------------------------------------

ADJ_CASES = [
    ((17,), (3,), (301,), 2),
    ((12, 10), (3, 4), (26, 73), 3),
    ((40, 37), (2, 3), (257, 191), 1),
    ((8, 7), (0, 1), (65, 70), 2),
    ((13, 11), (5, 5), (129, 64), 4),
    ((11, 9), (7, 2), (80, 60), 1),
    ((10, 9, 8), (2, 3, 2), (70, 33, 21), 3),
    ((14, 6, 9), (3, 3, 4), (130, 150, 7), 5),
    ((6, 5, 5, 4), (2, 1, 2, 1), (9, 8, 7, 6), 2),
]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("case", ADJ_CASES, ids=lambda c: f"n{c[0]}p{c[1]}")
def test_adjoint_vs_oracle(gpu, case, dtype):
    sg = gpu
    n, p, m, nout = case
    D = [min(1, pi) for pi in p]
    dims = make_dims(sg, n, p, m, D=D, dtype=dtype)
    grid = sg.SplineGrid(tuple(dims), nout)
    rng = np.random.default_rng(sum(m))
    E = rng.standard_normal(tuple(m) + (nout,)).astype(dtype)
    ebar = sg.from_numpy_colmajor(E, grid.dtype, "cuda")
    Pbar = sg.colmajor_zeros(tuple(n) + (nout,), grid.dtype, "cuda")
    od = oracle_dims(dims)
    for deriv in [tuple(0 for _ in n), tuple(D)]:
        sg.evaluate_adjoint_(grid, derivative_order=deriv, eval=ebar, control_points=Pbar)
        want = O.adjoint_direct(od, E.astype(np.float64), deriv=deriv)
        err = rel_err(eval_host(sg, Pbar), want)
        assert err < TOL[dtype], (deriv, err)


def test_adjoint_dot_product_identity(gpu):
    sg = gpu
    n, p, m = (15, 12, 9), (3, 2, 3), (61, 47, 33)
    dims = make_dims(sg, n, p, m, D=(1, 1, 1), dtype=np.float64)
    grid = sg.SplineGrid(tuple(dims), 2)
    rng = np.random.default_rng(9)
    x = sg.from_numpy_colmajor(rng.standard_normal(n + (2,)), torch.float64, "cuda")
    y = sg.from_numpy_colmajor(rng.standard_normal(m + (2,)), torch.float64, "cuda")
    for deriv in [(0, 0, 0), (1, 0, 1)]:
        Mx = sg.colmajor_zeros(m + (2,), torch.float64, "cuda")
        MTy = sg.colmajor_zeros(n + (2,), torch.float64, "cuda")
        sg.evaluate_(grid, derivative_order=deriv, control_points=x, eval=Mx)
        sg.evaluate_adjoint_(grid, derivative_order=deriv, eval=y, control_points=MTy)
        lhs, rhs = float((Mx * y).sum()), float((x * MTy).sum())
        assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)


def test_adjoint_accumulate(gpu):
    sg = gpu
    dims = make_dims(sg, (10, 9), (3, 2), (80, 70), dtype=np.float64)
    grid = sg.SplineGrid(tuple(dims), 2)
    rng = np.random.default_rng(2)
    ebar = sg.from_numpy_colmajor(rng.standard_normal((80, 70, 2)), torch.float64, "cuda")
    base = sg.from_numpy_colmajor(rng.standard_normal((10, 9, 2)), torch.float64, "cuda")
    acc = base.clone(memory_format=torch.preserve_format)
    acc = sg.colmajor(acc)
    out = sg.colmajor_zeros((10, 9, 2), torch.float64, "cuda")
    sg.evaluate_adjoint_(grid, eval=ebar, control_points=out)
    sg.evaluate_adjoint_(grid, eval=ebar, control_points=acc, accumulate=True)
    torch.testing.assert_close(acc, base + out, rtol=1e-14, atol=1e-14)


def test_g2_error_informed_index_set(gpu, goldens):
    # test/test_local_refinement.jl:48-67: the adjoint maps the error field
    # onto the refined 10x10 control grid; entries above the mean are the
    # 18 refinement indices.
    sg = gpu
    g = goldens["G2"]
    kv = sg.KnotVector(np.arange(9, dtype=np.float32) / 8,
                       np.array([3, 1, 1, 1, 1, 1, 1, 1, 3], dtype=np.int32))
    dims = [sg.SplineDimension(10, 2, 50, knot_vector=kv, device="cuda") for _ in range(2)]
    grid = sg.SplineGrid(tuple(dims), 3)
    err = torch.zeros_like(grid.eval)
    (a0, a1), (b0, b1) = g["error_ranges_1based"]
    err[a0 - 1:a1, b0 - 1:b1, g["error_channel_1based"] - 1] = 1
    cp_err = sg.colmajor_zeros((10, 10, 3), torch.float32, "cuda")
    sg.evaluate_adjoint_(grid, eval=err, control_points=cp_err)
    e = cp_err.sum(dim=2).cpu().numpy()
    thr = e.sum() / e.size
    ij = (np.argwhere(e.T > thr)[:, ::-1] + 1).tolist()
    assert ij == g["refinement_indices_1based"]


def test_determinism(gpu):
    sg = gpu
    dims = make_dims(sg, (30, 25, 12), (3, 3, 2), (200, 150, 40), dtype=np.float32)
    grid = sg.SplineGrid(tuple(dims), 3)
    rng = np.random.default_rng(4)
    grid.control_points.copy_(random_cp(rng, (30, 25, 12), 3, np.float32))
    sg.evaluate_(grid)
    first = grid.eval.clone()
    sg.evaluate_(grid)
    assert torch.equal(first, grid.eval)
    a = sg.colmajor_zeros((30, 25, 12, 3), torch.float32, "cuda")
    b = sg.colmajor_zeros((30, 25, 12, 3), torch.float32, "cuda")
    sg.evaluate_adjoint_(grid, control_points=a)
    sg.evaluate_adjoint_(grid, control_points=b)
    assert torch.equal(a, b)


# --------------------------------------------------------------------------
# validation and error behaviour (a9)
# --------------------------------------------------------------------------

def test_derivative_order_validation(gpu, caplog):
    # test/test_spline_grid.jl:55-83
    sg = gpu
    dims = make_dims(sg, (5, 6), (3, 2), (7, 9), D=(1, 1))
    grid = sg.SplineGrid(tuple(dims), 2)
    with caplog.at_level(logging.ERROR, logger="splinegrids_amd"):
        with pytest.raises(ValueError, match=r"Invalid derivative order\(s\) supplied\. If you "
                           r"want to evaluate \(higher order\) derivatives, specify this at "
                           r"construction as SplineDimension\(\.\.\.; max_derivative_order\)\."):
            sg.evaluate_(grid, derivative_order=(2, 1))
    assert [r.getMessage() for r in caplog.records] == [
        "The maximum derivative order available for spline dimension 1 is 1, got 2."]


def test_nurbs_validation(gpu):
    # test/test_nurbs_grid.jl:11-23 and src/adjoint.jl:57
    sg = gpu
    s = sg.SplineDimension(4, 2, 100, max_derivative_order=1, device="cuda")
    grid = sg.NURBSGrid(s, 1)
    with pytest.raises(ValueError, match="Computing derivatives of NURBS is currently not supported."):
        sg.evaluate_(grid, derivative_order=(1,))
    with pytest.raises(AssertionError, match="Adjoint evaluation not supported for NURBS."):
        sg.evaluate_adjoint_(grid)


def test_c_abi_rejects_bad_arguments(gpu):
    sg = gpu
    from splinegrids_amd import _lib

    L = _lib.lib()
    dims = make_dims(sg, (5, 6), (3, 2), (7, 9))
    grid = sg.SplineGrid(tuple(dims), 2)
    plan = grid.plan()
    rc = L.sg_evaluate(plan.handle, _lib.i32_array([1, 0]), grid.eval.data_ptr(), None,
                       grid.eval.data_ptr(), 0, 0, None)
    assert rc == _lib.SG_ERR_INVALID
    assert "spline dimension 1 is 0, got 1" in L.sg_last_error().decode()
    rc = L.sg_evaluate(plan.handle, _lib.i32_array([0, 0]), None, None,
                       grid.eval.data_ptr(), 0, 0, None)
    assert rc == _lib.SG_ERR_INVALID
    rc = L.sg_evaluate(None, _lib.i32_array([0, 0]), None, None, None, 0, 0, None)
    assert rc == _lib.SG_ERR_INVALID
