"""sg_dimension_evaluate_noalloc against sg_dimension_evaluate.

The no-allocation variant runs the same kernels with its scratch in a fixed
device array instead of per-call hipMalloc; both are called through the
C-ABI on the same inputs and must agree bit for bit, for two calls in a row
(the second with a smaller input, so that the shared scratch holds the first
call's contents), in FP64 and FP32, with derivative orders.  Sizes beyond
the fixed scratch are refused with a message.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _inputs(rng, n, p, m, dtype):
    inner = np.sort(rng.uniform(0, 1, n - p - 1))
    knots = np.concatenate([np.zeros(p + 1), inner, np.ones(p + 1)]).astype(dtype)
    t = np.concatenate([rng.uniform(-0.1, 1.1, m - 3), [0.0, inner[len(inner) // 2], 1.0]])
    return (torch.from_numpy(knots).cuda(), torch.from_numpy(t.astype(dtype)).cuda())


def _run(fn, knots, t, p, D, dtype_code, tdtype):
    m = t.numel()
    idx = torch.full((m,), -7, dtype=torch.int32, device="cuda")
    ev = torch.full((m * (p + 1) * (D + 1),), float("nan"), dtype=tdtype, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    rc = fn(knots.data_ptr(), knots.numel(), t.data_ptr(), m, p, D, idx.data_ptr(),
            ev.data_ptr(), dtype_code, stream)
    torch.cuda.synchronize()
    return rc, idx, ev


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_noalloc_matches_allocating_entry_point(gpu, dtype):
    from splinegrids_amd import _lib

    L = _lib.lib()
    fn = L.sg_dimension_evaluate_noalloc
    code, tdtype = (_lib.SG_F64, torch.float64) if dtype == np.float64 else \
        (_lib.SG_F32, torch.float32)
    rng = np.random.default_rng(3)
    for n, p, D, m in [(40, 3, 2, 5001), (9, 2, 1, 37)]:  # second call smaller
        knots, t = _inputs(rng, n, p, m, dtype)
        rc_ref, idx_ref, ev_ref = _run(L.sg_dimension_evaluate, knots, t, p, D, code, tdtype)
        assert rc_ref == _lib.SG_OK, L.sg_last_error()
        rc, idx, ev = _run(fn, knots, t, p, D, code, tdtype)
        assert rc == _lib.SG_OK, L.sg_last_error()
        assert torch.equal(idx, idx_ref), (n, p, m)
        assert torch.equal(ev, ev_ref), (n, p, m)  # bit for bit, no NaN left


def test_noalloc_refuses_sizes_beyond_its_scratch(gpu):
    from splinegrids_amd import _lib

    L = _lib.lib()
    fn = L.sg_dimension_evaluate_noalloc
    rng = np.random.default_rng(4)
    knots, t = _inputs(rng, 12, 3, 70000, np.float32)  # 70000 * 4 * 4 * 2 > 2^20
    rc, idx, _ = _run(fn, knots, t, 3, 3, _lib.SG_F32, torch.float32)
    assert rc == _lib.SG_ERR_INVALID
    assert "exceeds the fixed scratch" in L.sg_last_error().decode()
    assert bool((idx == -7).all())  # nothing was launched
