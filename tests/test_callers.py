"""Callers of the hot path: LinearMap + lsqr and differentiation.

test/test_LinearMapsExt.jl:14-32 (lsqr recovers random control points at
rtol 1e-5, so the adjoint is the exact transpose of the forward map) and
test/test_EnzymeExt.jl:13-68 (reverse mode gives a non-zero gradient),
plus central finite differences of a least-squares loss.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from helpers import make_dims

pytestmark = pytest.mark.gpu


def test_linear_map_lsqr_recovers_control_points(gpu):
    from scipy.sparse.linalg import lsqr

    sg = gpu
    rng = np.random.default_rng(1)
    n, p, m, nout = (12, 10), (3, 4), (26, 73), 3
    dims = make_dims(sg, n, p, m, dtype=np.float64)
    grid = sg.SplineGrid(tuple(dims), nout)
    P = rng.random(n + (nout,))
    grid.control_points.copy_(P)
    sg.evaluate_(grid)
    target = sg.vec_colmajor(grid.eval).cpu().numpy().copy()
    A = sg.LinearMap(grid)
    fit = lsqr(A, target, atol=1e-14, btol=1e-14, iter_lim=2000)[0]
    np.testing.assert_allclose(fit, P.reshape(-1, order="F"), rtol=1e-5)


def test_linear_map_is_transpose(gpu):
    sg = gpu
    rng = np.random.default_rng(2)
    dims = make_dims(sg, (9, 8, 7), (2, 3, 2), (31, 29, 17), D=(1, 1, 1), dtype=np.float64)
    grid = sg.SplineGrid(tuple(dims), 2)
    A = sg.LinearMap(grid, derivative_order=(1, 0, 1))
    x = rng.standard_normal(A.shape[1])
    y = rng.standard_normal(A.shape[0])
    lhs, rhs = float(y @ A.matvec(x)), float(x @ A.rmatvec(y))
    assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(lhs))


def test_reverse_mode_gradient_nonzero(gpu):
    # test/test_EnzymeExt.jl:13-68: loss = sum(eval), gradient w.r.t. the control points
    sg = gpu
    dims = make_dims(sg, (10, 10), (2, 2), (50, 50))
    grid = sg.SplineGrid(tuple(dims), 2)
    P = torch.rand(grid.control_points.shape, device="cuda", requires_grad=True)
    loss = sg.spline_eval(grid, P).sum()
    loss.backward()
    assert P.grad is not None and bool((P.grad != 0).any())
    # d sum(M P) / dP = M^T 1 = column sums of M, which are positive
    assert bool((P.grad > 0).all())


def test_gradient_vs_central_differences(gpu):
    sg = gpu
    rng = np.random.default_rng(4)
    n, p, m, nout = (11, 9, 7), (3, 3, 3), (40, 36, 20), 3
    dims = make_dims(sg, n, p, m, dtype=np.float64)
    grid = sg.SplineGrid(tuple(dims), nout)
    P0 = sg.from_numpy_colmajor(rng.uniform(-1, 1, n + (nout,)), torch.float64, "cuda")
    target = sg.from_numpy_colmajor(rng.standard_normal(m + (nout,)), torch.float64, "cuda")

    def loss(P):
        return 0.5 * ((sg.spline_eval(grid, P) - target) ** 2).sum()

    P = P0.clone().requires_grad_(True)
    loss(P).backward()
    g = P.grad.detach()
    h = 1e-3
    flat = sg.vec_colmajor(P0)
    for q in rng.integers(0, flat.numel(), 24):
        e = torch.zeros_like(flat)
        e[q] = h
        Pp = sg.reshape_colmajor(flat + e, P0.shape)
        Pm = sg.reshape_colmajor(flat - e, P0.shape)
        fd = (loss(Pp) - loss(Pm)).item() / (2 * h)
        assert abs(fd - sg.vec_colmajor(g)[q].item()) <= 1e-7 * max(1.0, abs(fd))


def test_torch_gradcheck_and_forward_mode(gpu):
    sg = gpu
    dims = make_dims(sg, (5, 4), (2, 1), (9, 7), D=(1, 1), dtype=np.float64)
    grid = sg.SplineGrid(tuple(dims), 2)
    P = sg.colmajor(torch.randn(5, 4, 2, dtype=torch.float64, device="cuda")).requires_grad_(True)
    f = lambda x: sg.spline_eval(grid, x, derivative_order=(1, 1))
    assert torch.autograd.gradcheck(f, (P,), eps=1e-6, atol=1e-8)
    import torch.autograd.forward_ad as fwAD

    t = sg.colmajor(torch.randn(5, 4, 2, dtype=torch.float64, device="cuda"))
    with fwAD.dual_level():
        out = f(fwAD.make_dual(P.detach(), t))
        tangent = fwAD.unpack_dual(out).tangent
    torch.testing.assert_close(tangent, f(t), rtol=1e-13, atol=1e-13)
