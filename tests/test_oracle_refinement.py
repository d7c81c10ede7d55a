"""The refinement oracle (oracle/refinement.py) pinned on the CPU.

Pins: the reference's own refinement-matrix tests (test/test_refinement_
matrix.jl: identity, diagonal and banded products, the invalid-row
messages), geometry invariance under knot insertion and refinement
(test/test_refinement.jl:29-58), dense mode products and the transpose
identity for mult! / mult_adjoint!, brute force for the Flag scatter, and
the reference's local-refinement counts and index set
(test/test_local_refinement.jl:11-67).  Planted mistakes (a dropped alpha,
a wrong span index) must fail the geometry check.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
import refinement as R


def _banded(rng, m, n, lo, hi):
    """A random banded matrix with lower bandwidth lo and upper bandwidth hi
    (every entry of the band non-zero), as BandedMatrices.brand does."""
    A = np.zeros((m, n))
    for i in range(m):
        for j in range(max(0, i - lo), min(n, i + hi + 1)):
            A[i, j] = rng.uniform(0.1, 1.0)
    return A


def test_eye_and_diagonal():
    # test/test_refinement_matrix.jl:17-44
    M = R.RefMat.eye(100)
    assert M.invalid_rows() == []
    M2 = R.product(M, M)
    assert (M2.m, M2.n) == (100, 100)
    assert np.array_equal(M2.row_pointer, M.row_pointer)
    assert np.array_equal(M2.column_start, M.column_start)
    assert np.array_equal(M2.nzval, M.nzval)
    rng = np.random.default_rng(1)
    a, b = rng.uniform(size=100), rng.uniform(size=100)
    C = R.product(R.RefMat.from_dense(np.diag(a)), R.RefMat.from_dense(np.diag(b)))
    assert np.array_equal(C.nzval, a * b)


def test_banded_products_equal_dense():
    # test/test_refinement_matrix.jl:46-59
    rng = np.random.default_rng(1)
    A = _banded(rng, 100, 100, 4, 3)
    B = _banded(rng, 100, 200, 3, 5)
    RA, RB = R.RefMat.from_dense(A), R.RefMat.from_dense(B)
    assert RA.invalid_rows() == [] and RB.invalid_rows() == []
    assert np.array_equal(RA.dense(), A)
    C = R.product(RA, RB)
    np.testing.assert_allclose(C.dense(), A @ B, rtol=1e-14, atol=1e-14)


def test_invalid_rows_messages():
    # test/test_refinement_matrix.jl:61-75
    M = np.zeros((2, 50))
    M[0, :20] = 1
    M[1, 38:50] = 1
    assert R.RefMat.from_dense(M).invalid_rows() == [2]
    M = np.zeros((3, 50))
    M[0, :25] = 1
    M[2, 25:50] = 1
    assert R.RefMat.from_dense(M).invalid_rows() == [2, 3]


def _grid_eval(values_list, mults_list, degrees, P, m):
    dims = []
    for vals, mults, p, mm in zip(values_list, mults_list, degrees, m):
        knots = O.expand_knots(np.asarray(vals, np.float64), mults)
        t = O.sample_points((knots[0], knots[-1]), mm)
        dims.append(O.Dim(knots, t, p))
    return O.eval_direct(dims, P)


def _random_knot_vector(rng, n, p):
    """Clamped knots with random interior values (distribution = :random)."""
    k = n - p + 1
    inner = np.sort(rng.uniform(0, 1, k - 2))
    vals = np.concatenate([[0.0], inner, [1.0]])
    mults = np.ones(k, np.int32)
    mults[0] = mults[-1] = p + 1
    return vals, mults


@pytest.mark.parametrize("plant", [None, "alpha", "span"])
def test_knot_insertion_keeps_geometry(plant):
    # test/test_refinement.jl:29-40
    rng = np.random.default_rng(1)
    n, p, m = (5, 8, 6), (4, 2, 3), (15, 20, 25)
    kv = [_random_knot_vector(rng, n[d], p[d]) for d in range(3)]
    P = rng.uniform(size=n + (2,))
    ref = _grid_eval([v for v, _ in kv], [u for _, u in kv], p, P, m)
    vals, mults = [list(x) for x in zip(*kv)]
    valid = True
    for d in range(3):
        knots_all = O.expand_knots(vals[d], mults[d])
        new_v, new_u, span = R.insert_knot_values(vals[d], mults[d], 0.25)
        if plant == "span":
            span += 1
        with np.errstate(divide="ignore", invalid="ignore"):  # a wrong span may divide by zero
            A = R.insert_knot_matrix(knots_all, p[d], span, 0.25)
        if plant == "alpha":
            A.nzval[A.nzval < 1] = 0.5
        assert (A.m, A.n) == (n[d] + 1, n[d])
        valid = valid and A.invalid_rows() == []
        if plant is None:
            assert valid
            np.testing.assert_allclose(A.dense().sum(axis=1), 1.0, rtol=1e-14)
        P = R.mult([A], [d + 1], P)
        vals[d], mults[d] = new_v, new_u
        n = tuple(P.shape[:3])
    got = _grid_eval(vals, mults, p, P, m)
    err = np.abs(got - ref).max()
    if plant is None:
        assert err < 1e-13
    else:
        assert not valid or err > 1e-6  # the planted mistake is caught


def test_refinement_keeps_geometry_and_composes():
    # test/test_refinement.jl:42-58
    rng = np.random.default_rng(2)
    n, p, m = (5, 8, 6), (4, 2, 3), (15, 20, 25)
    kv = [_random_knot_vector(rng, n[d], p[d]) for d in range(3)]
    P = rng.uniform(size=n + (2,))
    ref = _grid_eval([v for v, _ in kv], [u for _, u in kv], p, P, m)
    vals, mults = [list(x) for x in zip(*kv)]
    for d in range(3):
        new_v, new_u, A = R.refine_dimension(vals[d], mults[d], p[d])
        assert (A.m, A.n) == (n[d] + len(vals[d]) - 1, n[d])
        # the product of single insertions, done densely
        v, u = np.asarray(vals[d], np.float64), np.asarray(mults[d], np.int32)
        D = np.eye(n[d])
        for k in v[:-1] + np.diff(v) / 2:
            ka = O.expand_knots(v, u)
            v, u, span = R.insert_knot_values(v, u, float(k))
            D = R.insert_knot_matrix(ka, p[d], span, float(k)).dense() @ D
        np.testing.assert_allclose(A.dense(), D, rtol=1e-13, atol=1e-15)
        P = R.mult([A], [d + 1], P)
        vals[d], mults[d] = new_v, new_u
        n = tuple(P.shape[:3])
    got = _grid_eval(vals, mults, p, P, m)
    assert np.abs(got - ref).max() < 1e-13


def _refmats(rng, shape_b, dims):
    mats = []
    for d in dims:
        nb = shape_b[d - 1]
        A = _banded(rng, nb + 3, nb, 2, 1)
        A[-3:, :] = 0
        A[-3:, -2:] = rng.uniform(0.1, 1, (3, 2))
        mats.append(R.RefMat.from_dense(A))
    return mats


def test_mult_is_the_mode_product_and_adjoint_its_transpose():
    rng = np.random.default_rng(3)
    B = rng.standard_normal((7, 6, 5, 2))
    for dims in ([1], [2, 3], [1, 2, 3], [3, 1]):
        mats = _refmats(rng, B.shape, dims)
        for A in mats:
            assert A.invalid_rows() == []
        Y = R.mult(mats, dims, B)
        want = B.copy()
        for A, d in zip(mats, dims):
            want = np.moveaxis(np.tensordot(A.dense(), want, axes=(1, d - 1)), 0, d - 1)
        np.testing.assert_allclose(Y, want, rtol=1e-13, atol=1e-13)
        Z = rng.standard_normal(Y.shape)
        Bt = R.mult_adjoint(mats, dims, Z)
        assert Bt.shape == B.shape
        lhs, rhs = float((Y * Z).sum()), float((B * Bt).sum())
        assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_flag_scatter_brute_force():
    rng = np.random.default_rng(4)
    shape_b = (6, 5, 1)
    dims = [1, 2]
    mats = _refmats(rng, shape_b, dims)
    mats[0].nzval[0] = 0.0  # a stored zero still counts (structure, not value)
    Y = (rng.uniform(size=(mats[0].m, mats[1].m, 1)) < 0.3).astype(np.uint8)
    got = R.mult_adjoint_flags(mats, dims, Y)
    S = [np.zeros((A.m, A.n), bool) for A in mats]
    for A, s in zip(mats, S):
        for i in range(1, A.m + 1):
            cs, ce = A.row_range(i)
            s[i - 1, cs - 1:ce] = True
    want = np.zeros(shape_b, np.uint8)
    for i1, i2 in zip(*np.nonzero(Y[..., 0])):
        want[np.ix_(S[0][i1], S[1][i2], [0])] = 1
    assert np.array_equal(got, want)


def _default_local_refinement(rng, n, p, nout):
    """add_default_local_refinement (src/control_points.jl:364-425) on a grid
    with default knots: every dimension refined at its span midpoints."""
    vals, mults, mats = [], [], []
    for nd, pd in zip(n, p):
        v, u = O.default_knot_values(nd, pd)
        v2, u2, A = R.refine_dimension(v, u, pd)
        vals.append(v2)
        mults.append(u2)
        mats.append(A)
    base = np.asfortranarray(rng.uniform(size=tuple(n) + (nout,)))
    all_idx = np.array([[i + 1, j + 1] for j in range(n[1]) for i in range(n[0])], np.int32)
    lv0 = R.Level([], [], all_idx, base.reshape(-1, nout, order="F").copy())
    refined = R.mult(mats, [1, 2], base)
    lv1 = R.Level([1, 2], mats, np.zeros((0, 2), np.int32), np.zeros((0, nout)))
    return R.Thb([base, refined], [lv0, lv1]), vals, mults


def _ranges(*rs):
    """activate_local_control_point_range! (src/control_points.jl:531-543):
    the product of the ranges, first range fastest."""
    import itertools
    return np.array([list(t)[::-1] for t in itertools.product(*[list(r) for r in rs[::-1]])],
                    np.int32)


def test_local_refinement_counts_and_geometry():
    # test/test_local_refinement.jl:11-41
    rng = np.random.default_rng(5)
    thb, _, _ = _default_local_refinement(rng, (6, 6), (2, 2), 3)
    assert thb.arrays[1].shape[:2] == (10, 10)
    assert thb.n_control_points() == 36
    before = thb.evaluate().copy()
    thb.activate(_ranges(range(1, 5), range(1, 7)))
    thb.activate(_ranges(range(1, 7), range(1, 3)))
    thb.activate(_ranges(range(9, 11), range(7, 11)))
    thb.deactivate_overwritten()
    after = thb.evaluate()
    assert thb.n_control_points() == 63
    np.testing.assert_allclose(after, before, rtol=1e-14, atol=1e-14)


def test_local_refinement_transpose_identity():
    rng = np.random.default_rng(6)
    thb, _, _ = _default_local_refinement(rng, (6, 7), (2, 3), 2)
    thb.activate(_ranges(range(2, 6), range(3, 7)))
    thb.deactivate_overwritten()
    x = rng.standard_normal((thb.n_control_points(), 2))
    thb.set_flat_values(x)
    y_shape = thb.arrays[-1].shape
    Mx = thb.evaluate().copy()
    y = rng.standard_normal(y_shape)
    MTy = thb.evaluate_adjoint(y)
    lhs, rhs = float((Mx * y).sum()), float((x * MTy).sum())
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_error_informed_index_set(goldens):
    # test/test_local_refinement.jl:43-67: the grid adjoint of the error onto the
    # default-refined 10 x 10 control grid, summed over channels, above its mean
    g = goldens["G2"]
    rng = np.random.default_rng(7)
    thb, vals, mults = _default_local_refinement(rng, (6, 6), (2, 2), 3)
    dims = []
    for v, u in zip(vals, mults):
        knots = O.expand_knots(np.asarray(v, np.float64), u)
        dims.append(O.Dim(knots, O.sample_points((0.0, 1.0), 50), 2))
    err = np.zeros((50, 50, 3))
    (a0, a1), (b0, b1) = g["error_ranges_1based"]
    err[a0 - 1:a1, b0 - 1:b1, g["error_channel_1based"] - 1] = 1
    e = O.adjoint_direct(dims, err).sum(axis=2)
    thr = e.sum() / e.size
    idx = np.argwhere(e.T > thr)[:, ::-1] + 1
    thb.activate(idx.astype(np.int32))
    assert thb.levels[-1].indices.tolist() == g["refinement_indices_1based"]
