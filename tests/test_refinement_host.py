"""Host-side logic of the refinement API, checked against the oracle on the
CPU: knot-vector insertion (src/refinement.jl:107-133), the stored form of a
dense matrix (src/refinement_matrix.jl:476-504), the first-occurrence
``unique`` of activate_local_refinement! (src/control_points.jl:481-493)
and the index order of activate_local_control_point_range!
(src/control_points.jl:531-539).  No GPU is used."""
from __future__ import annotations

import itertools

import numpy as np
import pytest

import refinement as R


@pytest.fixture(scope="module")
def mods(sg):
    import importlib

    return (importlib.import_module(sg.__name__ + ".refinement"),
            importlib.import_module(sg.__name__ + ".refinement_matrix"))


@pytest.mark.parametrize("knot_new", [0.05, 0.25, 0.6, 0.95])
def test_insert_knot_vector_matches_oracle(sg, mods, knot_new):
    ref, _ = mods
    kv = sg.KnotVector(np.array([0.0, 0.2, 0.5, 0.7, 1.0]), np.array([3, 1, 2, 1, 3], np.int32))
    new, span = ref.insert_knot_vector(kv, knot_new)
    v, u, span_o = R.insert_knot_values(kv.knot_values, kv.multiplicities, knot_new)
    assert span == span_o
    assert np.array_equal(new.knot_values, v)
    assert np.array_equal(new.multiplicities, u)
    # knots_all[span] <= knot_new < knots_all[span + 1] (1-based)
    assert kv.knots_all[span - 1] <= knot_new < kv.knots_all[span]


def test_insert_existing_knot_is_refused(sg, mods):
    ref, _ = mods
    kv = sg.KnotVector(np.array([0.0, 0.5, 1.0]), np.array([2, 1, 2], np.int32))
    with pytest.raises(AssertionError):
        ref.insert_knot_vector(kv, 0.5)


def test_dense_to_stored_form_matches_oracle(mods):
    _, rm = mods
    rng = np.random.default_rng(1)
    A = np.zeros((9, 7))
    for i in range(9):
        a = min(6, i * 7 // 9)
        A[i, a:min(7, a + 3)] = rng.uniform(0.1, 1, min(7, a + 3) - a)
    A[4, :] = 0  # an empty row: column_start 0, as the reference stores it
    m, n, rp, cs, nz = rm._from_dense(A)
    o = R.RefMat.from_dense(A)
    assert (m, n) == (o.m, o.n)
    assert np.array_equal(np.asarray(rp), o.row_pointer)
    assert np.array_equal(np.asarray(cs), o.column_start)
    assert np.array_equal(np.asarray(nz), o.nzval)


def test_unique_rows_keep_first_occurrence(mods):
    ref, _ = mods
    rows = np.array([[3, 1], [1, 2], [3, 1], [2, 2], [1, 2], [0, 5]])
    assert ref._unique_rows(rows).tolist() == [[3, 1], [1, 2], [2, 2], [0, 5]]
    assert ref._unique_rows(np.zeros((0, 2), np.int64)).shape == (0, 2)


def test_range_product_order(mods):
    ref, _ = mods
    # Iterators.product(1:4, 1:6): the first range fastest
    want = [[i, j] for j in range(1, 7) for i in range(1, 5)]
    assert ref._range_product([range(1, 5), (1, 6)]).tolist() == want
    want3 = [[i, j, k] for k in (5, 6) for j in range(2, 4) for i in (1,)]
    assert ref._range_product([(1, 1), range(2, 4), (5, 6)]).tolist() == want3
    # the oracle's helper forms the same order
    orc = np.array([list(t)[::-1] for t in itertools.product(range(1, 7), range(1, 5))])
    assert orc.tolist() == want
