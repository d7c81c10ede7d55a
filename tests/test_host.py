"""Host logic and the C-ABI library, without a GPU.

Knot vectors (test/test_knot_vector.jl), the reference's validation
messages (src/validation.jl), the column-major helpers, and that
libsplinegrids_hip.so loads and exports every symbol include/*.h declares.
"""
from __future__ import annotations

import logging
import os
import re
import types

import numpy as np
import pytest
import torch

import oracle as O
from conftest import ROOT


def test_knot_vector_validation(sg):
    # test/test_knot_vector.jl:11-24
    with pytest.raises(AssertionError, match="knot_values and multiplicities must be of the same length."):
        sg.KnotVector(np.arange(1, 6), np.arange(1, 4))
    with pytest.raises(AssertionError, match="knot_values must be sorted."):
        sg.KnotVector(np.array([3, 2, 1]), np.arange(1, 4))
    with pytest.raises(AssertionError, match="knot_values must be unique."):
        sg.KnotVector(np.array([1, 1, 2]), np.arange(1, 4))


def test_knot_vector_construction(sg, goldens):
    # test/test_knot_vector.jl:26-42 (G3)
    g = goldens["G3"]
    kv1 = sg.KnotVector(5, 2)
    np.testing.assert_allclose(kv1.knot_values, np.arange(4) / 3, rtol=1e-6)
    assert kv1.multiplicities.tolist() == [3, 1, 1, 3]
    np.testing.assert_allclose(kv1.knots_all, g["knots_all"], rtol=1e-6)
    kv2 = sg.KnotVector(5, 2, extent=(5.0, 7.0))
    np.testing.assert_allclose(kv2.knot_values, 2 * kv1.knot_values + 5.0, rtol=1e-6)
    np.testing.assert_allclose(kv2.knots_all, 2 * kv1.knots_all + 5.0, rtol=1e-6)
    kv3 = sg.KnotVector(10, 3, extent=(4, 8), distribution="random",
                        rng=np.random.default_rng(0))
    assert kv3.knot_values[0] == 4 and kv3.knot_values[-1] == 8
    with pytest.raises(AssertionError):
        sg.KnotVector(5, 10)
    with pytest.raises(ValueError, match="Unsupported knot distribution type"):
        sg.KnotVector(5, 2, distribution="chebyshev")


@pytest.mark.parametrize("n,p,dtype", [(5, 2, np.float32), (25, 3, np.float64), (128, 3, np.float32)])
def test_host_setup_matches_oracle(sg, n, p, dtype):
    """The package's default knots and sample points equal the oracle's own
    restatement of src/knot_vector.jl:75-113 and src/spline_dimension.jl:121-128."""
    from splinegrids_amd.knot_vector import uniform_range

    kv = sg.KnotVector(n, p, float_type=dtype)
    vals, mults = O.default_knot_values(n, p, dtype=dtype)
    assert np.array_equal(kv.knot_values, vals) and np.array_equal(kv.multiplicities, mults)
    assert np.array_equal(kv.knots_all.astype(np.float64), O.expand_knots(vals, mults))
    t = uniform_range(float(kv.extent[0]), float(kv.extent[1]), 77).astype(dtype)
    assert np.array_equal(t, O.sample_points(kv.extent, 77, dtype))


def _dims(*maxd):
    return tuple(types.SimpleNamespace(max_derivative_order=d) for d in maxd)


def test_validate_partial_derivatives_messages(sg, caplog):
    # src/validation.jl:52-73; test/test_spline_grid.jl:73-82
    with caplog.at_level(logging.ERROR, logger="splinegrids_amd"):
        with pytest.raises(ValueError, match=r"^Invalid derivative order\(s\) supplied\."):
            sg.validate_partial_derivatives(_dims(1, 1), (2, 1))
    assert [r.getMessage() for r in caplog.records] == [
        "The maximum derivative order available for spline dimension 1 is 1, got 2."]
    caplog.clear()
    with caplog.at_level(logging.ERROR, logger="splinegrids_amd"):
        with pytest.raises(ValueError):
            sg.validate_partial_derivatives(_dims(0, 2, 1), (1, 3, -1))
    assert len(caplog.records) == 3
    sg.validate_partial_derivatives(_dims(1, 2), (1, 2))


def test_validate_spline_grid_messages(sg, caplog):
    # src/validation.jl:1-50
    dims = tuple(types.SimpleNamespace(n_basis_functions=n) for n in (5, 6))
    cp = torch.zeros(5, 7, 2)
    ev = torch.zeros(9, 9, 3)
    with caplog.at_level(logging.ERROR, logger="splinegrids_amd"):
        with pytest.raises(ValueError, match="Errors encountered when validating spline grid "
                           "constructor inputs."):
            sg.validate_spline_grid(dims, cp, None, ev)
    assert [r.getMessage() for r in caplog.records] == [
        "The number of output dimensions from the control points 2 and eval 3 must agree.",
        "The control point grid sizes from the spline dimensions (5, 6) and the control "
        "points (5, 7) must agree."]
    caplog.clear()
    with caplog.at_level(logging.ERROR, logger="splinegrids_amd"):
        with pytest.raises(ValueError):
            sg.validate_spline_grid(dims, torch.zeros(5, 6, 2), torch.zeros(5, 6, 1), torch.zeros(9, 9, 2))
    assert [r.getMessage() for r in caplog.records] == [
        "The number of input dimensions from the spline dimensions, control_points, eval and "
        "weights must agree, got (2, 2, 2, 3) respectively."]


def test_colmajor_helpers(sg):
    a = np.arange(24.0).reshape(2, 3, 4, order="F")
    t = sg.from_numpy_colmajor(a, torch.float64, "cpu")
    assert sg.utils.is_colmajor(t) if hasattr(sg, "utils") else True
    assert t[1, 2, 3].item() == a[1, 2, 3]
    flat = torch.arange(24.0)
    r = sg.reshape_colmajor(flat, (2, 3, 4))
    assert np.array_equal(r.numpy(), np.arange(24.0).reshape(2, 3, 4, order="F"))
    assert torch.equal(sg.vec_colmajor(r), flat)
    z = sg.colmajor_zeros((3, 5), torch.float32, "cpu")
    assert z.stride() == (1, 3)


def _header_functions():
    with open(os.path.join(ROOT, "include", "splinegrids_hip.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(sg_\w+)\s*\(", text, re.M)))


def test_library_exports_header_symbols(sg):
    from splinegrids_amd import _lib

    names = _header_functions()
    assert len(names) >= 10
    assert sorted(_lib.EXPORTS) == names
    L = _lib.load()  # opening the library needs no GPU
    for name in names:
        assert hasattr(L, name), name
    major, minor, patch = (__import__("ctypes").c_int32() for _ in range(3))
    import ctypes

    assert L.sg_version(ctypes.byref(major), ctypes.byref(minor), ctypes.byref(patch)) == 0
    assert (major.value, minor.value) == (0, 1)


def test_product_fails_loudly_without_gpu(sg):
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(sg.SplineGridsError, match="no HIP device"):
        sg.SplineDimension(5, 2, 10, device="cuda")
