"""Shared test set-up.

Two backends run one test corpus, as in the reference (test files take
``--gpu_backend``, e.g. test/test_spline_grid.jl:5-9): the CPU oracle under
``oracle/`` is the checker, and the HIP library is what is checked.  Tests
that need the GPU are marked ``gpu``.
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def goldens():
    with open(os.path.join(GOLDEN, "reference_goldens.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def scipy_vectors():
    with np.load(os.path.join(GOLDEN, "scipy_vectors.npz")) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def sg():
    """The product package (splinegrids.jl_amd/), imported as splinegrids_amd."""
    import __graft_entry__

    return __graft_entry__.package()


@pytest.fixture(scope="session")
def gpu(sg):
    """The product package with its HIP library loaded on a visible GPU."""
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    sg.lib()  # raises if the HIP library is missing
    return sg
