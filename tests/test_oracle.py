"""Pins the CPU oracle before anything is checked against it.

Each oracle function is checked against the reference's own golden vectors
(G1-G4), against scipy.interpolate (an independent de Boor implementation),
and against what the mathematics fixes (partition of unity, Kronecker
products, the transpose identity).  The last tests plant mistakes in a copy of
the oracle's output and show that the same checks catch them.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


# --------------------------------------------------------------------------
# a1-a3: knot vectors and sample points
# --------------------------------------------------------------------------


def test_g3_knot_vector(goldens):
    g = goldens["G3"]
    vals, mults = O.default_knot_values(g["n"], g["p"])
    np.testing.assert_allclose(vals, g["knot_values"], rtol=0, atol=1e-16)
    assert mults.tolist() == g["multiplicities"]
    knots = O.expand_knots(vals, mults)
    np.testing.assert_allclose(knots, g["knots_all"], rtol=0, atol=1e-16)
    vals2, mults2 = O.default_knot_values(g["n"], g["p"], extent=g["extent2"])
    np.testing.assert_allclose(vals2, 2 * vals + 5.0, rtol=1e-15)
    assert (mults2 == mults).all()


def test_sample_points_closed_form():
    t = O.sample_points((-2.0, 3.0), 11)
    assert t[0] == -2.0 and t[-1] == 3.0
    np.testing.assert_allclose(np.diff(t), 0.5, rtol=1e-14)


def test_sample_indices_brute_force():
    rng = np.random.default_rng(1)
    for p in range(0, 5):
        inner = np.sort(rng.uniform(0, 1, 9))
        vals = np.concatenate([[0.0], inner[:4], [inner[4]], inner[5:], [1.0]])
        mults = np.ones(len(vals), dtype=np.int32)
        mults[0] = mults[-1] = p + 1
        mults[5] = min(p, 2) if p > 0 else 1  # an interior repeated knot
        knots = O.expand_knots(vals, mults)
        t = np.concatenate([rng.uniform(-0.1, 1.1, 200), knots, [0.0, 1.0]])
        got = O.sample_indices(t, knots, p)
        want = np.clip(np.searchsorted(knots, t, side="right"), p + 1,
                       len(knots) - p - 1)
        assert (got == want).all()
        inside = (t >= 0) & (t < 1)
        i0 = got[inside] - 1
        assert (knots[i0] <= t[inside]).all() and (t[inside] < knots[i0 + 1]).all()


# --------------------------------------------------------------------------
# a5: basis tables
# --------------------------------------------------------------------------


CASES = ["uniform_p3", "uniform_p0", "uniform_p1", "extent_p2", "doubled_p3",
         "random_p5"]


@pytest.mark.parametrize("name", CASES)
def test_basis_vs_scipy(scipy_vectors, name):
    v = scipy_vectors
    p, D = (int(x) for x in v[f"basis_{name}_meta"])
    d = O.Dim(v[f"basis_{name}_knots"], v[f"basis_{name}_x"], p, D)
    for nu in range(D + 1):
        want = v[f"basis_{name}_nu{nu}"]
        assert _rel(d.dense(nu), want) < 1e-12, (name, nu)


@pytest.mark.parametrize("degree", range(6))
@pytest.mark.parametrize("random", [False, True])
def test_partition_of_unity(degree, random):
    # test/test_spline_dimension.jl:10-29
    n, m = 25, 500
    if random:
        rng = np.random.default_rng(degree)
        v = np.cumsum(rng.uniform(size=n - degree + 1))
        v = (v - v[0]) / (v[-1] - v[0])
        mults = np.ones(len(v), dtype=np.int32)
        mults[0] = mults[-1] = degree + 1
        knots = O.expand_knots(v, mults)
        d = O.Dim(knots, O.sample_points((0.0, 1.0), m), degree)
    else:
        d = O.Dim.default(n, degree, m)
    B = d.table[:, :, 0]
    assert B.shape == (m, degree + 1)
    assert (B >= 0).all()
    np.testing.assert_allclose(B.sum(axis=1), 1.0, rtol=1e-14)


def test_fd_derivatives():
    # test/test_spline_dimension.jl:31-66
    d = O.Dim.default(10, 3, 5000, max_deriv=2)
    t = d.t
    dt = np.diff(t)
    B0, B1, B2 = d.dense(0), d.dense(1), d.dense(2)
    fd1 = np.diff(B0, axis=0) / dt[:, None]
    fd2 = np.diff(fd1, axis=0) / dt[1:, None]
    np.testing.assert_allclose(B1[:-1], fd1, rtol=1e-2, atol=1e-2 * np.abs(fd1).max())
    np.testing.assert_allclose(B2[2:], fd2, rtol=1e-2, atol=1e-2 * np.abs(fd2).max())


# --------------------------------------------------------------------------
# a10/a11: forward
# --------------------------------------------------------------------------


def _g1(goldens):
    g = goldens["G1"]
    n, p, m = g["n_control_points"], g["degree"], g["n_sample_points"]
    P = np.array(g["control_points_colmajor"]).reshape(n + [g["Nout"]], order="F")
    E = np.array(g["eval_colmajor"]).reshape(m + [g["Nout"]], order="F")
    dims = [O.Dim.default(n[i], p[i], m[i]) for i in range(2)]
    return dims, P, E


def test_g1_golden(goldens):
    dims, P, E = _g1(goldens)
    assert _rel(O.eval_direct(dims, P), E) < 1e-14


def test_constant_control_points():
    # test/test_spline_grid.jl:11-27
    for n in range(2, 11):
        for p in range(1, n):
            d = O.Dim.default(n, p, 100)
            out = O.eval_direct([d], np.ones((n, 1)), dtype=np.float32)
            np.testing.assert_allclose(out, 1.0, rtol=1e-6)


def test_grid2_mixed_partials_vs_scipy(scipy_vectors):
    v = scipy_vectors
    dims = [O.Dim(v["grid2_k1"], v["grid2_x1"], 3, 1),
            O.Dim(v["grid2_k2"], v["grid2_x2"], 2, 1)]
    for o1 in range(2):
        for o2 in range(2):
            got = O.eval_direct(dims, v["grid2_P"], deriv=(o1, o2))
            assert _rel(got, v[f"grid2_eval_{o1}{o2}"]) < 1e-12, (o1, o2)


def test_grid3_vs_scipy(scipy_vectors):
    v = scipy_vectors
    dims = [O.Dim(v["grid3_k1"], v["grid3_x1"], 2, 1),
            O.Dim(v["grid3_k2"], v["grid3_x2"], 3, 1),
            O.Dim(v["grid3_k3"], v["grid3_x3"], 1, 1)]
    assert _rel(O.eval_direct(dims, v["grid3_P"]), v["grid3_eval_000"]) < 1e-12
    got = O.eval_direct(dims, v["grid3_P"], deriv=(1, 0, 1))
    assert _rel(got, v["grid3_eval_101"]) < 1e-12


def test_nurbs_vs_scipy(scipy_vectors):
    v = scipy_vectors
    dims = [O.Dim(v["nurbs2_k1"], v["nurbs2_x1"], 2), O.Dim(v["nurbs2_k2"], v["nurbs2_x2"], 3)]
    got = O.eval_direct(dims, v["nurbs2_P"], W=v["nurbs2_W"])
    assert _rel(got, v["nurbs2_eval"]) < 1e-12


def test_g4_nurbs_circle(goldens):
    g = goldens["G4"]
    knots = O.expand_knots(np.array(g["knot_values"]), np.array(g["multiplicities"]))
    t = O.sample_points((knots[0], knots[-1]), g["n_sample_points"], np.float32)
    d = O.Dim(knots, t.astype(np.float64), g["degree"])
    P = np.array(g["control_points"])
    W = np.ones(g["n_control_points"])
    W[1::2] = g["weights_even_1based"]
    out = O.eval_direct([d], P, W=W)
    r2 = out[:, 0] ** 2 + out[:, 1] ** 2
    np.testing.assert_allclose(r2, 1.0, rtol=1e-12)
    assert len({tuple(np.round(x, 12)) for x in out[1:]}) == len(out) - 1


def test_dense_kronecker():
    rng = np.random.default_rng(3)
    dims = [O.Dim.default(6, 2, 9, 1), O.Dim.default(5, 3, 8, 1), O.Dim.default(4, 1, 7, 1)]
    P = rng.standard_normal((6, 5, 4, 2))
    for deriv in [(0, 0, 0), (1, 0, 1), (0, 1, 0)]:
        got = O.eval_direct(dims, P, deriv=deriv)
        want = np.einsum("ia,jb,kc,abco->ijko", dims[0].dense(deriv[0]),
                         dims[1].dense(deriv[1]), dims[2].dense(deriv[2]), P)
        assert _rel(got, want) < 1e-13


def test_eval_points_and_slab():
    rng = np.random.default_rng(4)
    dims = [O.Dim.default(7, 3, 11), O.Dim.default(6, 2, 10), O.Dim.default(5, 2, 9)]
    P = rng.standard_normal((7, 6, 5, 3))
    full = O.eval_direct(dims, P)
    lin = rng.integers(0, 11 * 10 * 9, 50)
    pts = O.eval_points(dims, P, lin, nthreads=4)
    assert np.array_equal(pts, full.reshape(-1, 3, order="F")[lin])
    slab = O.eval_direct(dims, P, last_range=(3, 7), nthreads=4)
    assert np.array_equal(slab, full[:, :, 3:7])


# --------------------------------------------------------------------------
# a12: adjoint
# --------------------------------------------------------------------------


def test_adjoint_is_transpose():
    rng = np.random.default_rng(5)
    dims = [O.Dim.default(6, 2, 13, 1), O.Dim.default(7, 3, 9, 1)]
    x = rng.standard_normal((6, 7, 2))
    y = rng.standard_normal((13, 9, 2))
    for deriv in [(0, 0), (1, 0), (1, 1)]:
        Mx = O.eval_direct(dims, x, deriv=deriv)
        MTy = O.adjoint_direct(dims, y, deriv=deriv)
        lhs, rhs = np.sum(Mx * y), np.sum(x * MTy)
        assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1)
        M = np.kron(dims[1].dense(deriv[1]), dims[0].dense(deriv[0]))  # col-major
        want = (M.T @ y.reshape(-1, 2, order="F")).reshape(6, 7, 2, order="F")
        assert _rel(MTy, want) < 1e-13


def test_adjoint_f32_matches_f64():
    rng = np.random.default_rng(6)
    dims = [O.Dim.default(8, 3, 40), O.Dim.default(6, 2, 30)]
    y = rng.standard_normal((40, 30, 3))
    a64 = O.adjoint_direct(dims, y)
    a32 = O.adjoint_direct(dims, y, dtype=np.float32, nthreads=4)
    assert _rel(a32, a64) < 1e-5


def g2_refined_dims():
    """The 10x10 grid that add_default_local_refinement produces from the 6x6,
    p=2 grid of test/test_local_refinement.jl:48-67: span midpoints inserted
    (src/refinement.jl:256-280), 50 uniform samples, Float32 values."""
    vals = np.arange(9, dtype=np.float64) / 8
    mults = np.array([3, 1, 1, 1, 1, 1, 1, 1, 3], dtype=np.int32)
    knots = O.expand_knots(vals, mults)
    t = O.sample_points((0.0, 1.0), 50, np.float32).astype(np.float64)
    return [O.Dim(knots, t, 2), O.Dim(knots, t, 2)]


def g2_indices(Pbar):
    """src/control_points.jl:560-575: channel sum, > mean, findall order."""
    e = Pbar.sum(axis=-1)
    thr = e.sum() / e.size
    ij = np.argwhere(e.T > thr)[:, ::-1] + 1  # column-major order, 1-based
    return ij.tolist()


def g2_error(goldens):
    g = goldens["G2"]
    err = np.zeros((50, 50, 3))
    (a0, a1), (b0, b1) = g["error_ranges_1based"]
    err[a0 - 1:a1, b0 - 1:b1, g["error_channel_1based"] - 1] = 1.0
    return err


def test_g2_adjoint_index_set(goldens):
    dims = g2_refined_dims()
    Pbar = O.adjoint_direct(dims, g2_error(goldens))
    assert g2_indices(Pbar) == goldens["G2"]["refinement_indices_1based"]


# --------------------------------------------------------------------------
# planted mistakes: the checks above catch plausible oracle bugs
# --------------------------------------------------------------------------


def test_planted_mistakes_are_caught(scipy_vectors, goldens):
    v = scipy_vectors
    p, D = (int(x) for x in v["basis_doubled_p3_meta"])
    d = O.Dim(v["basis_doubled_p3_knots"], v["basis_doubled_p3_x"], p, D)
    good = d.dense(1)
    want = v["basis_doubled_p3_nu1"]
    assert _rel(good, want) < 1e-12
    d.table[:, :, 1] *= -1                       # wrong derivative sign
    assert _rel(d.dense(1), want) > 1e-3
    d.table[:, :, 1] *= -1
    d.idx1 = d.idx1 + 1                          # off-by-one span
    d.idx1[d.idx1 > d.n] = d.n
    assert _rel(d.dense(0), v["basis_doubled_p3_nu0"]) > 1e-3

    dims, P, E = _g1(goldens)
    assert _rel(O.eval_direct(dims, P), E) < 1e-14
    assert _rel(O.eval_direct(dims[::-1], P.transpose(1, 0, 2)).transpose(1, 0, 2), E) < 1e-14
    Pt = P.copy()
    Pt[:, :, [0, 1]] = Pt[:, :, [1, 0]]          # swapped channels
    assert _rel(O.eval_direct(dims, Pt), E) > 1e-3

    dims2 = g2_refined_dims()
    err = g2_error(goldens)
    Pbar = O.adjoint_direct(dims2, err)
    Pbar_bad = np.roll(Pbar, 1, axis=0)          # shifted band ownership
    assert g2_indices(Pbar_bad) != goldens["G2"]["refinement_indices_1based"]


# ---- rational derivatives (NURBS): pinned by the mathematics ----

def test_rational_derivs_unit_weights_are_bspline_derivs():
    rng = np.random.default_rng(31)
    dims = [O.Dim.default(7, 3, 13, 2), O.Dim.default(6, 2, 11, 2)]
    P = rng.standard_normal((7, 6, 2))
    W = np.ones((7, 6))
    for deriv in [(1, 0), (0, 1), (1, 1), (2, 1)]:
        got = O.rational_derivs(dims, P, W, deriv)
        want = O.eval_direct(dims, P, deriv=deriv)
        assert _rel(got, want) < 1e-12, deriv


def test_rational_derivs_circle_tangent(goldens):
    """|R| = 1 on the NURBS circle, so R . dR/dt = 0 and dR/dt != 0."""
    g = goldens["G4"]
    knots = O.expand_knots(np.array(g["knot_values"]), np.array(g["multiplicities"]))
    t = O.sample_points((knots[0], knots[-1]), g["n_sample_points"], np.float64)
    d = O.Dim(knots, t, g["degree"], 1)
    P = np.array(g["control_points"])
    W = np.ones(g["n_control_points"])
    W[1::2] = g["weights_even_1based"]
    R = O.eval_direct([d], P, W=W)
    dR = O.rational_derivs([d], P, W, (1,))
    speed = np.linalg.norm(dR, axis=1)
    assert speed.min() > 0.5
    np.testing.assert_allclose(np.sum(R * dR, axis=1) / speed, 0.0, atol=1e-12)


def test_rational_derivs_finite_differences():
    """Central differences of the NURBS map in each sample coordinate (no
    sample on an interior knot, where a higher derivative jumps)."""
    rng = np.random.default_rng(32)
    n, p, m = (6, 5, 4), (3, 2, 2), (9, 8, 8)
    dims = [O.Dim.default(n[i], p[i], m[i], 2) for i in range(3)]
    P = rng.standard_normal(n + (2,))
    W = rng.uniform(0.5, 2.0, n)
    h = 1e-5

    def shifted(axis, dt):
        out = []
        for i, d in enumerate(dims):
            t = d.t + (dt if i == axis else 0.0)
            t = np.clip(t, d.knots_all[0], d.knots_all[-1])
            out.append(O.Dim(d.knots_all, t, d.degree, d.max_deriv))
        return out

    def interior(axis):
        sl = [slice(None)] * 3
        sl[axis] = slice(1, -1)  # the shift clips at the ends of the domain
        return tuple(sl)

    for axis in range(3):
        e = [0, 0, 0]
        e[axis] = 1
        got = O.rational_derivs(dims, P, W, tuple(e))
        fd = (O.eval_direct(shifted(axis, h), P, W=W)
              - O.eval_direct(shifted(axis, -h), P, W=W)) / (2 * h)
        assert _rel(got[interior(axis)], fd[interior(axis)]) < 1e-6, axis
    # a mixed partial: differences of the first derivative along dimension 2
    got = O.rational_derivs(dims, P, W, (1, 1, 0))
    fd = (O.rational_derivs(shifted(1, h), P, W, (1, 0, 0))
          - O.rational_derivs(shifted(1, -h), P, W, (1, 0, 0))) / (2 * h)
    assert _rel(got[interior(1)], fd[interior(1)]) < 1e-6


def test_adjoint_rational_transpose_identity():
    """<R(P), e> = <P, R^T e> for the NURBS map (linear in P, weights fixed)."""
    rng = np.random.default_rng(33)
    dims = [O.Dim.default(7, 3, 19), O.Dim.default(6, 2, 15), O.Dim.default(5, 2, 11)]
    P = rng.standard_normal((7, 6, 5, 2))
    W = rng.uniform(0.5, 2.0, (7, 6, 5))
    E = rng.standard_normal((19, 15, 11, 2))
    lhs = np.sum(O.eval_direct(dims, P, W=W) * E)
    rhs = np.sum(P * O.adjoint_rational(dims, W, E))
    assert abs(lhs - rhs) < 1e-12 * max(abs(lhs), 1.0)
    # unit weights: the B-spline adjoint
    np.testing.assert_allclose(O.adjoint_rational(dims, np.ones_like(W), E),
                               O.adjoint_direct(dims, E), rtol=1e-13, atol=1e-13)
