"""Refinement matrices, knot insertion and THB control points on the GPU.

The reference's own tests, run through the HIP library:
test/test_refinement_matrix.jl (identity, diagonal and banded products, the
invalid-row messages), test/test_refinement.jl (geometry kept by knot
insertion and refinement), test/test_local_refinement.jl (36 -> 63 control
points with the geometry kept, and the error-informed index set G2) and
test/test_LinearMapsExt.jl:34-62 (lsqr through a THB grid).  Beside them,
parity with the oracle (oracle/refinement.py): FP64 refinement matrices,
mode products, transposed products and THB chains bit-exact (the kernels sum
in the order a sequential run of the reference's loops adds), the Flag
scatter and every index set exact; FP32 results within 1e-6 of FP64.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle as O
import refinement as R
from helpers import make_dims

pytestmark = pytest.mark.gpu

TORCH = {np.float32: torch.float32, np.float64: torch.float64}


def _to_oracle(A) -> R.RefMat:
    return R.RefMat(A.m, A.n, A.row_pointer.cpu().numpy().astype(np.int32),
                    A.column_start.cpu().numpy().astype(np.int32),
                    A.nzval.cpu().numpy().astype(np.float64))


def _to_gpu(sg, M: R.RefMat, dtype=np.float64):
    return sg.RefinementMatrix(M.m, M.n, M.row_pointer, M.column_start,
                               M.nzval.astype(dtype), device="cuda")


def _same(A, M: R.RefMat) -> None:
    """A (device) equals the oracle's M: structure and FP64 values exact."""
    assert (A.m, A.n) == (M.m, M.n)
    assert np.array_equal(A.row_pointer.cpu().numpy(), M.row_pointer)
    assert np.array_equal(A.column_start.cpu().numpy(), M.column_start)
    assert np.array_equal(A.nzval.cpu().numpy(), M.nzval)


def _banded(rng, m, n, lo, hi):
    A = np.zeros((m, n))
    for i in range(m):
        for j in range(max(0, i - lo), min(n, i + hi + 1)):
            A[i, j] = rng.uniform(0.1, 1.0)
    return A


def _colmajor(sg, a: np.ndarray, dtype=torch.float64):
    return sg.from_numpy_colmajor(a, dtype, "cuda")


def _host(sg, t: torch.Tensor) -> np.ndarray:
    return sg.to_numpy_colmajor(t)


# ---- test/test_refinement_matrix.jl ----

def test_unit_matrix(gpu):
    sg = gpu
    N = 100
    M = sg.rmeye(N)
    assert M[25, 25] == 1
    assert M[25, 26] == 0
    with pytest.raises(IndexError,
                       match=r"Index \(101, 1\) out of bounds for refinement matrix of size "
                             r"\(100, 100\)\."):
        M[N + 1, 1]
    M2 = M @ M
    assert isinstance(M2, sg.RefinementMatrix)
    assert (M2.m, M2.n) == (M.m, M.n)
    assert torch.equal(M2.row_pointer, M.row_pointer)
    assert torch.equal(M2.column_start, M.column_start)
    assert torch.equal(M2.nzval, M.nzval)
    assert M2 == M
    assert M2.size() == (N, N) and len(M2) == N * N


def test_diagonal_product(gpu):
    sg = gpu
    rng = np.random.default_rng(1)
    a, b = rng.uniform(size=100), rng.uniform(size=100)
    A, B = sg.RefinementMatrix(np.diag(a)), sg.RefinementMatrix(np.diag(b))
    C = A @ B
    assert np.array_equal(C.nzval.cpu().numpy(), a * b)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_banded_product(gpu, dtype):
    sg = gpu
    rng = np.random.default_rng(1)
    Ad, Bd = _banded(rng, 100, 100, 4, 3), _banded(rng, 100, 200, 3, 5)
    A = sg.RefinementMatrix(Ad.astype(dtype))
    B = sg.RefinementMatrix(Bd.astype(dtype))
    C = A @ B
    dense = Ad.astype(dtype).astype(np.float64) @ Bd.astype(dtype).astype(np.float64)
    if dtype == np.float64:
        _same(C, R.product(_to_oracle(A), _to_oracle(B)))
        np.testing.assert_allclose(C.collect(), dense, rtol=1e-14, atol=1e-14)
    else:
        np.testing.assert_allclose(C.collect(), dense, rtol=1e-6, atol=1e-6)


def test_invalid_rows(gpu):
    sg = gpu
    M = np.zeros((2, 50))
    M[0, :20] = 1
    M[1, 38:50] = 1
    with pytest.raises(ValueError, match=r"^Invalid rows: \[2\]\.$"):
        sg.RefinementMatrix(M)
    M = np.zeros((3, 50))
    M[0, :25] = 1
    M[2, 25:50] = 1
    with pytest.raises(ValueError, match=r"^Invalid rows: \[2, 3\]\.$"):
        sg.RefinementMatrix(M)


# ---- test/test_refinement.jl ----

def _refinement_grid(sg, dtype):
    n, p, m = (5, 8, 6), (4, 2, 3), (15, 20, 25)
    rng = np.random.default_rng(1)
    dims = [sg.SplineDimension(n[d], p[d], m[d], distribution="random", rng=rng,
                               float_type=dtype, device="cuda") for d in range(3)]
    grid = sg.SplineGrid(tuple(dims), 2)
    grid.control_points.copy_(rng.random(n + (2,)).astype(dtype))
    sg.evaluate_(grid)
    return grid


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 1e-5)])
def test_knot_insertion_keeps_geometry(gpu, dtype, tol):
    sg = gpu
    grid = _refinement_grid(sg, dtype)
    ref = grid.eval.clone()
    for d in (1, 2, 3):
        s = grid.spline_dimensions[d - 1]
        knots_before = s.host_knots()
        grid, A = sg.insert_knot(grid, d, 0.25)
        sg.evaluate_(grid)
        torch.testing.assert_close(grid.eval, ref, rtol=tol, atol=tol)
        size = grid.control_points.shape
        assert A.size() == (size[d - 1], size[d - 1] - 1)
        if dtype == np.float64:
            # the oracle's insertion matrix, bit for bit
            v, u = s.knot_vector.knot_values, s.knot_vector.multiplicities
            v2, u2, span = R.insert_knot_values(v, u, 0.25)
            _same(A, R.insert_knot_matrix(knots_before, s.degree, span, 0.25))
            assert np.array_equal(grid.spline_dimensions[d - 1].host_knots(),
                                  O.expand_knots(v2, u2))


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 1e-5)])
def test_refinement_keeps_geometry(gpu, dtype, tol):
    sg = gpu
    grid = _refinement_grid(sg, dtype)
    ref = grid.eval.clone()
    size_old = grid.control_points.shape
    for d in (1, 2, 3):
        s = grid.spline_dimensions[d - 1]
        grid, A = sg.refine(grid, d)
        sg.evaluate_(grid)
        torch.testing.assert_close(grid.eval, ref, rtol=tol, atol=tol)
        size = grid.control_points.shape
        assert A.size() == (size[d - 1], size_old[d - 1])
        if dtype == np.float64:
            v, u = s.knot_vector.knot_values, s.knot_vector.multiplicities
            v2, u2, M = R.refine_dimension(v, u, s.degree)
            _same(A, M)
            assert np.array_equal(grid.spline_dimensions[d - 1].knot_vector.knot_values, v2)


def test_knot_insertion_rejects_nurbs_and_existing_knots(gpu):
    sg = gpu
    dims = make_dims(sg, (5, 6), (2, 2), (9, 9), dtype=np.float64)
    grid = sg.NURBSGrid(tuple(dims), 2)
    with pytest.raises(AssertionError, match="Knot insertion not supported for NURBS"):
        sg.insert_knot(grid, 1, 0.3)
    with pytest.raises(AssertionError):
        sg.insert_knot(dims[0], float(dims[0].knot_vector.knot_values[1]))


# ---- mult! and mult_adjoint! ----

def _refmats(rng, shape_b, dims):
    mats = []
    for d in dims:
        nb = shape_b[d - 1]
        A = _banded(rng, nb + 3, nb, 2, 1)
        A[-3:, :] = 0
        A[-3:, -2:] = rng.uniform(0.1, 1, (3, 2))
        mats.append(R.RefMat.from_dense(A))
    return mats


@pytest.mark.parametrize("dims", [[1], [2, 3], [1, 2, 3], [3, 1]])
def test_mult_and_adjoint_match_oracle(gpu, dims):
    sg = gpu
    rng = np.random.default_rng(3)
    B = np.asfortranarray(rng.standard_normal((7, 6, 5, 2)))
    mats = _refmats(rng, B.shape, dims)
    Y_or = R.mult(mats, dims, B)
    Z = np.asfortranarray(rng.standard_normal(Y_or.shape))
    Bt_or = R.mult_adjoint(mats, dims, Z)
    for dtype in (np.float64, np.float32):
        tt = TORCH[dtype]
        gm = [_to_gpu(sg, M, dtype) for M in mats]
        Y = sg.colmajor_empty(Y_or.shape, tt, "cuda")
        sg.mult_(Y, gm, _colmajor(sg, B, tt), dims)
        Bt = sg.colmajor_empty(B.shape, tt, "cuda")
        Bt.fill_(7)  # overwritten
        sg.mult_adjoint_(Bt, gm, _colmajor(sg, Z, tt), dims)
        if dtype == np.float64:
            assert np.array_equal(_host(sg, Y), Y_or)
            assert np.array_equal(_host(sg, Bt), Bt_or)
            # accumulate adds the transposed product to what the array holds
            acc = _colmajor(sg, B, tt)
            sg.mult_adjoint_(acc, gm, _colmajor(sg, Z, tt), dims, accumulate=True)
            np.testing.assert_allclose(_host(sg, acc), B + Bt_or, rtol=1e-14, atol=1e-14)
        else:
            scale = np.abs(Y_or).max()
            assert np.abs(_host(sg, Y) - Y_or).max() <= 1e-6 * scale
            scale = np.abs(Bt_or).max()
            assert np.abs(_host(sg, Bt) - Bt_or).max() <= 1e-6 * scale


def test_flag_adjoint_matches_oracle(gpu):
    sg = gpu
    rng = np.random.default_rng(4)
    shape_b = (6, 5, 1)
    dims = [1, 2]
    mats = _refmats(rng, shape_b, dims)
    mats[0].nzval[0] = 0.0  # a stored zero still counts (structure, not value)
    Y = np.asfortranarray((rng.uniform(size=(mats[0].m, mats[1].m, 1)) < 0.3).astype(np.uint8))
    want = R.mult_adjoint_flags(mats, dims, Y)
    gm = [_to_gpu(sg, M) for M in mats]
    out = sg.colmajor_zeros(shape_b, torch.uint8, "cuda")
    sg.mult_adjoint_(out, gm, sg.from_numpy_colmajor(Y, torch.uint8, "cuda"), dims)
    assert np.array_equal(_host(sg, out), want)


# ---- test/test_local_refinement.jl ----

def _ranges_oracle(*rs):
    import itertools
    return np.array([list(t)[::-1] for t in itertools.product(*[list(r) for r in rs[::-1]])],
                    np.int32)


def _thb_oracle(grid, sg) -> R.Thb:
    """The oracle's THB state equal to a THB grid's (FP64)."""
    cp = grid.control_points
    arrays = [_host(sg, a).astype(np.float64) for a in cp.control_points_refined]
    levels = []
    for lr in cp.local_refinements:
        levels.append(R.Level(list(lr.dims_refinement),
                              [_to_oracle(A) for A in lr.refinement_matrices],
                              _host(sg, lr.refinement_indices).astype(np.int32).reshape(-1, grid.Nin),
                              _host(sg, lr.refinement_values).astype(np.float64).reshape(-1, grid.Nout)))
    return R.Thb(arrays, levels)


def test_local_refinement_counts_and_geometry(gpu):
    sg = gpu
    dims = make_dims(sg, (6, 6), (2, 2), (500, 500))
    grid = sg.SplineGrid(tuple(dims), 3)
    assert sg.get_n_control_points(grid) == 36
    assert sg.evaluate_(grid.control_points) is None
    grid = sg.add_default_local_refinement(grid)
    cp = grid.control_points
    assert isinstance(cp, sg.LocallyRefinedControlPoints)
    assert ("LocallyRefinedControlPoints for final grid of size (10, 10) in ℝ³ "
            "(Float32). Local refinements:") in repr(cp)
    assert sg.get_n_control_points(grid) == 36
    before = sg.obtain(cp).clone()
    sg.activate_local_control_point_range_(grid, range(1, 5), range(1, 7))
    sg.activate_local_control_point_range_(grid, (1, 6), (1, 2))
    sg.activate_local_control_point_range_(grid, range(9, 11), range(7, 11))
    sg.deactivate_overwritten_control_points_(grid.control_points)
    sg.evaluate_(grid.control_points)
    assert sg.get_n_control_points(grid) == 63
    torch.testing.assert_close(sg.obtain(cp), before, rtol=1e-6, atol=1e-6)
    grid = sg.add_default_local_refinement(grid)
    assert len(grid.control_points.local_refinements) == 3
    assert tuple(sg.obtain(grid.control_points).shape) == (18, 18, 3)
    sg.evaluate_(grid.control_points)
    sg.evaluate_(grid)


def test_thb_chain_matches_oracle(gpu):
    """The same activations on the GPU and in the oracle (FP64): index sets
    equal, the evaluated chain and its adjoint bit for bit."""
    sg = gpu
    rng = np.random.default_rng(5)
    dims = make_dims(sg, (6, 7), (2, 3), (40, 30), dtype=np.float64)
    grid = sg.SplineGrid(tuple(dims), 2)
    grid.control_points.copy_(rng.uniform(-1, 1, (6, 7, 2)))
    grid = sg.add_default_local_refinement(grid)
    thb = _thb_oracle(grid, sg)
    for rs in [(range(1, 5), range(1, 7)), (range(1, 7), range(1, 3)), (range(8, 11), range(7, 12))]:
        sg.activate_local_control_point_range_(grid, *rs)
        thb.activate(_ranges_oracle(*rs))
    sg.deactivate_overwritten_control_points_(grid.control_points)
    thb.deactivate_overwritten()
    cp = grid.control_points
    for lr, lv in zip(cp.local_refinements, thb.levels):
        assert np.array_equal(_host(sg, lr.refinement_indices).reshape(-1, 2), lv.indices)
        assert np.array_equal(_host(sg, lr.refinement_values).reshape(-1, 2), lv.values)
    # new values for every active control point, then the chain
    x = rng.standard_normal((sg.get_n_control_points(grid), 2))
    cp.copy_(torch.from_numpy(x))
    thb.set_flat_values(x)
    sg.evaluate_(cp)
    want = thb.evaluate()
    assert np.array_equal(_host(sg, sg.obtain(cp)), want)
    # the adjoint from a cotangent on the last level
    y = np.asfortranarray(rng.standard_normal(want.shape))
    sg.obtain(cp).copy_(_colmajor(sg, y))
    sg.evaluate_adjoint_(cp)
    got = cp.get_control_point_view().cpu().numpy()
    assert np.array_equal(got, thb.evaluate_adjoint(y))
    # and it is the transpose: <M x, y> = <x, M^T y>
    lhs, rhs = float((want * y).sum()), float((x * got).sum())
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_error_informed_index_set(gpu, goldens, dtype):
    # test/test_local_refinement.jl:43-67 (G2)
    sg = gpu
    g = goldens["G2"]
    dims = make_dims(sg, (6, 6), (2, 2), (50, 50), dtype=dtype)
    grid = sg.SplineGrid(tuple(dims), 3)
    grid = sg.add_default_local_refinement(grid)
    error = torch.zeros_like(grid.eval)
    (a0, a1), (b0, b1) = g["error_ranges_1based"]
    error[a0 - 1:a1, b0 - 1:b1, g["error_channel_1based"] - 1] = 1
    sg.error_informed_local_refinement_(grid, sg.colmajor(error))
    got = _host(sg, grid.control_points.local_refinements[-1].refinement_indices)
    assert got.reshape(-1, 2).tolist() == g["refinement_indices_1based"]


def test_thb_linear_map_lsqr(gpu):
    # test/test_LinearMapsExt.jl:34-62
    from scipy.sparse.linalg import lsqr

    sg = gpu
    rng = np.random.default_rng(1)
    n, p, m, nout = (12, 10), (3, 4), (26, 73), 3
    dims = make_dims(sg, n, p, m, dtype=np.float64)
    grid = sg.SplineGrid(tuple(dims), nout)
    grid = sg.add_default_local_refinement(grid)
    sg.activate_local_control_point_range_(grid, range(1, 5), range(1, 7))
    sg.activate_local_control_point_range_(grid, range(1, 7), range(1, 3))
    sg.activate_local_control_point_range_(grid, range(9, 11), range(7, 11))
    sg.deactivate_overwritten_control_points_(grid.control_points)
    n_cp = sg.get_n_control_points(grid)
    P = rng.random((n_cp, nout))
    grid.control_points.copy_(torch.from_numpy(P))
    sg.evaluate_(grid.control_points)
    sg.evaluate_(grid)
    target = sg.vec_colmajor(grid.eval).cpu().numpy().copy()
    A = sg.LinearMap(grid)
    assert A.shape == (grid.eval.numel(), n_cp * nout)
    x = rng.standard_normal(A.shape[1])
    y = rng.standard_normal(A.shape[0])
    lhs, rhs = float(y @ A.matvec(x)), float(x @ A.rmatvec(y))
    assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(lhs))
    fit = lsqr(A, target, atol=1e-14, btol=1e-14, iter_lim=5000)[0]
    np.testing.assert_allclose(fit, P.reshape(-1, order="F"), rtol=1e-5)
