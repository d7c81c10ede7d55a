"""Write tests/golden/scipy_vectors.npz: FP64 B-spline vectors from scipy.

scipy.interpolate is an independent implementation of the same mathematics
(de Boor's algorithm), so these vectors pin the oracle's basis tables,
derivative orders, tensor-product contraction and NURBS quotient without
using the reference.  Run:  python tests/golden/make_scipy_vectors.py

Cases (inputs are stored alongside the outputs):
  basis_<name>   dense (m, n) basis matrices for derivative orders 0..D
  grid2_*        2-D -> 2 evaluation with all mixed partials up to (1, 1)
  grid3_*        3-D -> 3 evaluation, orders (0,0,0) and (1,0,1)
  nurbs2_*       2-D -> 3 NURBS value: NdBSpline(w*P) / NdBSpline(w)
"""
from __future__ import annotations

import os

import numpy as np
from scipy.interpolate import BSpline, NdBSpline

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "scipy_vectors.npz")


def clamped(values, p):
    values = np.asarray(values, dtype=np.float64)
    return np.concatenate([[values[0]] * p, values, [values[-1]] * p])


def uniform_samples(lo, hi, m):
    t = np.array([lo + (hi - lo) * (i / (m - 1)) for i in range(m)])
    t[0], t[-1] = lo, hi
    return t


def dense_basis(knots, p, x, nu):
    n = len(knots) - p - 1
    return BSpline(knots, np.eye(n), p)(x, nu=nu)


def main() -> None:
    rng = np.random.default_rng(20240607)
    out = {}

    # --- per-dimension basis tables -------------------------------------
    cases = {}
    cases["uniform_p3"] = (clamped(np.linspace(0, 1, 5), 3), 3, 23, 2)
    cases["uniform_p0"] = (clamped(np.linspace(0, 1, 6), 0), 0, 17, 0)
    cases["uniform_p1"] = (clamped(np.linspace(0, 1, 7), 1), 1, 19, 1)
    cases["extent_p2"] = (clamped(np.linspace(-2, 3, 6), 2), 2, 29, 2)
    inner = np.sort(rng.uniform(0, 1, 7))
    vals = np.concatenate([[0.0], inner[:3], [inner[3], inner[3]], inner[4:], [1.0]])
    cases["doubled_p3"] = (clamped(vals, 3), 3, 41, 3)
    inner = np.sort(rng.uniform(0, 1, 9))
    cases["random_p5"] = (clamped(np.concatenate([[0.0], inner, [1.0]]), 5), 5, 37, 3)
    for name, (knots, p, m, D) in cases.items():
        x = uniform_samples(knots[0], knots[-1], m)
        out[f"basis_{name}_knots"] = knots
        out[f"basis_{name}_x"] = x
        out[f"basis_{name}_meta"] = np.array([p, D], dtype=np.int64)
        for nu in range(D + 1):
            out[f"basis_{name}_nu{nu}"] = dense_basis(knots, p, x, nu)

    # --- 2-D grid, all mixed partials up to (1, 1) ------------------------
    k1 = clamped(np.linspace(0, 1, 6), 3)                 # n1 = 8
    inner = np.sort(rng.uniform(0, 1, 4))
    k2 = clamped(np.concatenate([[0.0], inner[:2], [inner[2]] * 2, [1.0]]), 2)  # n2 = 7
    n1, n2 = len(k1) - 4, len(k2) - 3
    P = rng.standard_normal((n1, n2, 2))
    x1, x2 = uniform_samples(0, 1, 13), uniform_samples(0, 1, 11)
    X = np.stack(np.meshgrid(x1, x2, indexing="ij"), axis=-1)
    spl = NdBSpline((k1, k2), P, (3, 2))
    out["grid2_k1"], out["grid2_k2"], out["grid2_P"] = k1, k2, P
    out["grid2_x1"], out["grid2_x2"] = x1, x2
    for o1 in range(2):
        for o2 in range(2):
            out[f"grid2_eval_{o1}{o2}"] = spl(X, nu=(o1, o2))

    # --- 3-D grid --------------------------------------------------------
    k1 = clamped(np.linspace(0, 1, 4), 2)                 # n1 = 5
    k2 = clamped(np.linspace(-1, 1, 3), 3)                # n2 = 5
    k3 = clamped(np.linspace(0, 2, 5), 1)                 # n3 = 5
    P = rng.uniform(-1, 1, (5, 5, 5, 3))
    x1, x2, x3 = (uniform_samples(0, 1, 9), uniform_samples(-1, 1, 7),
                  uniform_samples(0, 2, 6))
    X = np.stack(np.meshgrid(x1, x2, x3, indexing="ij"), axis=-1)
    spl = NdBSpline((k1, k2, k3), P, (2, 3, 1))
    for i, (k, x) in enumerate([(k1, x1), (k2, x2), (k3, x3)], start=1):
        out[f"grid3_k{i}"], out[f"grid3_x{i}"] = k, x
    out["grid3_P"] = P
    out["grid3_eval_000"] = spl(X, nu=(0, 0, 0))
    out["grid3_eval_101"] = spl(X, nu=(1, 0, 1))

    # --- NURBS as a quotient of two B-splines ----------------------------
    k1 = clamped(np.linspace(0, 1, 5), 2)                 # n1 = 6
    k2 = clamped(np.linspace(0, 1, 4), 3)                 # n2 = 6
    P = rng.uniform(-1, 1, (6, 6, 3))
    W = rng.uniform(0.5, 2.0, (6, 6))
    x1, x2 = uniform_samples(0, 1, 10), uniform_samples(0, 1, 12)
    X = np.stack(np.meshgrid(x1, x2, indexing="ij"), axis=-1)
    num = NdBSpline((k1, k2), P * W[..., None], (2, 3))(X)
    den = NdBSpline((k1, k2), W, (2, 3))(X)
    out["nurbs2_k1"], out["nurbs2_k2"] = k1, k2
    out["nurbs2_x1"], out["nurbs2_x2"] = x1, x2
    out["nurbs2_P"], out["nurbs2_W"] = P, W
    out["nurbs2_eval"] = num / den[..., None]

    np.savez_compressed(OUT, **out)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
