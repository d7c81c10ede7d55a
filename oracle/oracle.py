"""CPU oracle for the SplineGrids.jl grid-evaluation path.

TEST INFRASTRUCTURE ONLY.  This module is the checker: ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it; the product package never does.  The numerics live in
``sg_oracle.c`` (built by ``oracle/Makefile`` into ``liboracle_sg.so``); this
file adds numpy plumbing and restates the host-side set-up of the reference.

Reference functions restated, and what pins each one (tests/test_oracle.py):

=========================  ==================================  ================================================
oracle function            reference                           pinned by
=========================  ==================================  ================================================
default_knot_values        src/knot_vector.jl:75-113           G3 (test/test_knot_vector.jl:26-36): exact
                                                               knots_all for (n=5, p=2) and affine extents
expand_knots               src/knot_vector.jl:28-37,           G3; multiplicity sums
                           src/util_kernels.jl:1-20
sample_points              src/spline_dimension.jl:121-128     closed form: endpoints exact, uniform spacing
sample_indices             src/util_kernels.jl:22-49           brute force (searchsorted 'right' + clamp) and
                                                               knots[i] <= t < knots[i+1] for interior t
basis                      src/spline_dimension.jl:160-215     scipy.interpolate.BSpline design matrices and
                                                               derivatives (fixtures in tests/golden/), partition
                                                               of unity, B >= 0, FD derivatives
eval_direct / eval_points  src/spline_grid.jl:119-183          G1 (test/test_spline_grid.jl:29-53), constant
                                                               control points => 1, scipy NdBSpline fixtures
                                                               (incl. mixed partials), dense Kronecker product
NURBS branch               src/spline_grid.jl:142-182          G4 unit circle (test/test_nurbs_grid.jl:25-52),
                                                               ratio of two B-spline evaluations (w*P and w)
adjoint_direct             src/adjoint.jl:1-83                 G2 index set (test/test_local_refinement.jl:
                                                               48-67), <Mx, y> = <x, M^T y>, dense M^T
adjoint_rational           none: the reference rejects the     the transpose identity <R(P), e> = <P, R^T e>
                           NURBS adjoint (src/adjoint.jl:57)   against eval_direct with weights; unit weights
                                                               = adjoint_direct
rational_derivs            none: the reference rejects NURBS   unit weights = B-spline derivatives, the NURBS
                           derivatives (src/validation.jl:     circle's R . dR/dt = 0, central finite
                           81-88); The NURBS Book eq. 4.8 /    differences of eval_direct with weights
                           Alg. A4.4 (Piegl & Tiller)
=========================  ==================================  ================================================

Every function above has at least one such check, so no oracle function is
"parity unpinned".
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_sg.so")
_lib = None

_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_f64p = ctypes.POINTER(ctypes.c_double)
_f32p = ctypes.POINTER(ctypes.c_float)


def build(force: bool = False) -> str:
    """Compile liboracle_sg.so with oracle/Makefile."""
    import subprocess

    if force and os.path.exists(_LIB_PATH):
        os.remove(_LIB_PATH)
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.sgo_expand_knots.restype = ctypes.c_int32
        L.sgo_max_threads.restype = ctypes.c_int32
        _lib = L
    return _lib


def max_threads() -> int:
    return int(lib().sgo_max_threads())


def _p(a: np.ndarray, t):
    return a.ctypes.data_as(t)


# ----------------------------------------------------------------------------
# Host-side set-up (restated independently of the product package)
# ----------------------------------------------------------------------------


def default_knot_values(n_basis: int, degree: int, extent=(0.0, 1.0),
                        dtype=np.float64) -> tuple[np.ndarray, np.ndarray]:
    """Clamped equispaced knot vector, src/knot_vector.jl:75-113 (:equispaced).

    Returns (knot_values, multiplicities)."""
    assert n_basis - degree >= 1
    k = n_basis - degree + 1
    lo, hi = float(extent[0]), float(extent[1])
    vals = np.array([lo + (hi - lo) * (i / (k - 1)) if 0 < i < k - 1 else
                     (lo if i == 0 else hi) for i in range(k)], dtype=np.float64)
    vals = vals.astype(dtype)
    mults = np.ones(k, dtype=np.int32)
    mults[0] = degree + 1
    mults[-1] = degree + 1
    return vals, mults


def expand_knots(values: np.ndarray, mults: np.ndarray) -> np.ndarray:
    values = np.ascontiguousarray(values, dtype=np.float64)
    mults = np.ascontiguousarray(mults, dtype=np.int32)
    out = np.zeros(int(mults.sum()), dtype=np.float64)
    n = lib().sgo_expand_knots(_p(values, _f64p), _p(mults, _i32p),
                               ctypes.c_int32(len(values)), _p(out, _f64p))
    assert n == len(out)
    return out


def sample_points(extent, m: int, dtype=np.float64) -> np.ndarray:
    """Uniform inclusive sample points, src/spline_dimension.jl:121-128."""
    lo, hi = float(extent[0]), float(extent[1])
    if m == 1:
        return np.array([lo], dtype=dtype)
    t = np.array([lo + (hi - lo) * (i / (m - 1)) for i in range(m)],
                 dtype=np.float64)
    t[0], t[-1] = lo, hi
    return t.astype(dtype)


def sample_indices(t: np.ndarray, knots_all: np.ndarray, degree: int) -> np.ndarray:
    t = np.ascontiguousarray(t, dtype=np.float64)
    knots_all = np.ascontiguousarray(knots_all, dtype=np.float64)
    out = np.zeros(len(t), dtype=np.int32)
    lib().sgo_sample_indices(_p(t, _f64p), ctypes.c_int32(len(t)),
                             _p(knots_all, _f64p), ctypes.c_int32(len(knots_all)),
                             ctypes.c_int32(degree), _p(out, _i32p))
    return out


def basis(knots_all: np.ndarray, t: np.ndarray, idx1: np.ndarray, degree: int,
          max_deriv: int) -> np.ndarray:
    """FP64 basis table, Fortran-ordered array of shape (m, p+1, D+1)."""
    t = np.ascontiguousarray(t, dtype=np.float64)
    knots_all = np.ascontiguousarray(knots_all, dtype=np.float64)
    idx1 = np.ascontiguousarray(idx1, dtype=np.int32)
    m = len(t)
    out = np.zeros((m, degree + 1, max_deriv + 1), dtype=np.float64, order="F")
    lib().sgo_basis(_p(knots_all, _f64p), _p(t, _f64p), _p(idx1, _i32p),
                    ctypes.c_int32(m), ctypes.c_int32(degree),
                    ctypes.c_int32(max_deriv), _p(out, _f64p))
    return out


class Dim:
    """One oracle dimension: knots, samples, spans and the FP64 table."""

    def __init__(self, knots_all, t, degree: int, max_deriv: int = 0):
        self.knots_all = np.asarray(knots_all, dtype=np.float64)
        self.t = np.asarray(t, dtype=np.float64)
        self.degree = int(degree)
        self.max_deriv = int(max_deriv)
        self.n = len(self.knots_all) - self.degree - 1
        self.m = len(self.t)
        self.idx1 = sample_indices(self.t, self.knots_all, self.degree)
        self.table = basis(self.knots_all, self.t, self.idx1, self.degree,
                           self.max_deriv)

    @classmethod
    def default(cls, n: int, degree: int, m: int, max_deriv: int = 0,
                extent=(0.0, 1.0), dtype=np.float64):
        """Reference defaults: clamped equispaced knots and uniform samples,
        both rounded to ``dtype`` (Tv) as the reference stores them."""
        vals, mults = default_knot_values(n, degree, extent, dtype)
        knots = expand_knots(vals.astype(np.float64), mults)
        t = sample_points((vals[0], vals[-1]), m, dtype).astype(np.float64)
        return cls(knots, t, degree, max_deriv)

    def dense(self, order: int = 0) -> np.ndarray:
        """Dense (m, n) basis matrix; src/spline_dimension.jl:251-272."""
        out = np.zeros((self.m, self.n))
        for j in range(self.m):
            base = self.idx1[j] - self.degree - 1
            for a in range(self.degree + 1):
                out[j, base + a] = self.table[j, a, order]
        return out


def _geom(dims: Sequence[Dim]):
    nin = len(dims)
    n = np.array([d.n for d in dims], dtype=np.int32)
    m = np.array([d.m for d in dims], dtype=np.int32)
    p = np.array([d.degree for d in dims], dtype=np.int32)
    D = np.array([d.max_deriv for d in dims], dtype=np.int32)
    return nin, n, m, p, D


def _ptr_array(arrs, ctype):
    arr_t = ctypes.POINTER(ctype) * len(arrs)
    return arr_t(*[a.ctypes.data_as(ctypes.POINTER(ctype)) for a in arrs])


def _check(dims, P, deriv, W):
    nin = len(dims)
    assert P.ndim == nin + 1 and P.shape[:nin] == tuple(d.n for d in dims)
    assert P.shape[-1] <= 64
    assert len(deriv) == nin
    for d, o in zip(dims, deriv):
        assert 0 <= o <= d.max_deriv
    if W is not None:
        assert W.shape == P.shape[:nin]


def eval_direct(dims: Sequence[Dim], P: np.ndarray, deriv=None,
                W: Optional[np.ndarray] = None, dtype=np.float64,
                last_range=None, nthreads: int = 1) -> np.ndarray:
    """Reference forward (src/spline_grid.jl:119-183) in Tv = ``dtype``.

    P is (n_1, ..., n_Nin, Nout) in any numpy order; the result is a Fortran
    array (m_1, ..., m_Nin, Nout), or only the slab ``last_range`` of the
    slowest sample dimension."""
    nin, n, m, p, D = _geom(dims)
    deriv = [0] * nin if deriv is None else list(deriv)
    _check(dims, P, deriv, W)
    nout = P.shape[-1]
    dt = np.dtype(dtype)
    ctype = ctypes.c_double if dt == np.float64 else ctypes.c_float
    tabs = [np.asfortranarray(d.table, dtype=dt) for d in dims]
    spans = [np.ascontiguousarray(d.idx1) for d in dims]
    Pf = np.asfortranarray(P, dtype=dt)
    Wf = None if W is None else np.asfortranarray(W, dtype=dt)
    lb, le = (0, int(m[-1])) if last_range is None else last_range
    shape = tuple(int(x) for x in m[:-1]) + (le - lb, nout)
    out = np.zeros(shape, dtype=dt, order="F")
    fn = lib().sgo_eval_direct_f64 if dt == np.float64 else lib().sgo_eval_direct_f32
    fn(ctypes.c_int32(nin), ctypes.c_int32(nout), _p(n, _i32p), _p(m, _i32p),
       _p(p, _i32p), _p(D, _i32p), _ptr_array(tabs, ctype),
       _ptr_array(spans, ctypes.c_int32), _p(np.array(deriv, dtype=np.int32), _i32p),
       Pf.ctypes.data_as(ctypes.POINTER(ctype)),
       None if Wf is None else Wf.ctypes.data_as(ctypes.POINTER(ctype)),
       out.ctypes.data_as(ctypes.POINTER(ctype)), ctypes.c_int32(lb),
       ctypes.c_int32(le), ctypes.c_int32(nthreads))
    return out


def eval_points(dims: Sequence[Dim], P: np.ndarray, lin: np.ndarray,
                deriv=None, W: Optional[np.ndarray] = None,
                nthreads: int = 1) -> np.ndarray:
    """FP64 direct sum at the given column-major linear sample indices;
    returns (len(lin), Nout)."""
    nin, n, m, p, D = _geom(dims)
    deriv = [0] * nin if deriv is None else list(deriv)
    _check(dims, P, deriv, W)
    nout = P.shape[-1]
    lin = np.ascontiguousarray(lin, dtype=np.int64)
    tabs = [np.asfortranarray(d.table) for d in dims]
    spans = [np.ascontiguousarray(d.idx1) for d in dims]
    Pf = np.asfortranarray(P, dtype=np.float64)
    Wf = None if W is None else np.asfortranarray(W, dtype=np.float64)
    out = np.zeros((len(lin), nout), order="F")
    lib().sgo_eval_points_f64(
        ctypes.c_int32(nin), ctypes.c_int32(nout), _p(n, _i32p), _p(m, _i32p),
        _p(p, _i32p), _p(D, _i32p), _ptr_array(tabs, ctypes.c_double),
        _ptr_array(spans, ctypes.c_int32), _p(np.array(deriv, dtype=np.int32), _i32p),
        _p(Pf, _f64p), None if Wf is None else _p(Wf, _f64p), _p(lin, _i64p),
        ctypes.c_int64(len(lin)), _p(out, _f64p), ctypes.c_int32(nthreads))
    return out


def adjoint_direct(dims: Sequence[Dim], ebar: np.ndarray, deriv=None,
                   dtype=np.float64, nthreads: int = 1) -> np.ndarray:
    """Reference adjoint (src/adjoint.jl:1-83).  FP64: sequential, fixed
    order.  FP32: OpenMP atomics as the reference's threaded CPU backend."""
    nin, n, m, p, D = _geom(dims)
    deriv = [0] * nin if deriv is None else list(deriv)
    assert ebar.shape[:nin] == tuple(int(x) for x in m)
    nout = ebar.shape[-1]
    dt = np.dtype(dtype)
    ctype = ctypes.c_double if dt == np.float64 else ctypes.c_float
    tabs = [np.asfortranarray(d.table, dtype=dt) for d in dims]
    spans = [np.ascontiguousarray(d.idx1) for d in dims]
    ef = np.asfortranarray(ebar, dtype=dt)
    out = np.zeros(tuple(int(x) for x in n) + (nout,), dtype=dt, order="F")
    args = [ctypes.c_int32(nin), ctypes.c_int32(nout), _p(n, _i32p), _p(m, _i32p),
            _p(p, _i32p), _p(D, _i32p), _ptr_array(tabs, ctype),
            _ptr_array(spans, ctypes.c_int32),
            _p(np.array(deriv, dtype=np.int32), _i32p),
            ef.ctypes.data_as(ctypes.POINTER(ctype)),
            out.ctypes.data_as(ctypes.POINTER(ctype))]
    if dt == np.float64:
        lib().sgo_adjoint_direct_f64(*args)
    else:
        lib().sgo_adjoint_direct_f32(*args, ctypes.c_int32(nthreads))
    return out


def _binom(n: int, k: int) -> int:
    from math import comb
    return comb(n, k)


def rational_derivs(dims: Sequence[Dim], P: np.ndarray, W: np.ndarray, deriv,
                    nthreads: int = 1) -> np.ndarray:
    """Partial derivative ``deriv`` of the rational map R = (sum N w P) / (sum N w)
    of a NURBS grid, in FP64 (Piegl & Tiller, The NURBS Book, eq. 4.8 and
    Alg. A4.4, for any number of dimensions): with A = sum N w P and
    w(x) = sum N w, Leibniz's rule on A = w R gives
        R^(a) = (A^(a) - sum_{0 < b <= a} C(a, b) w^(b) R^(a-b)) / w,
    C(a, b) = prod_d binom(a_d, b_d).  A^(b) and w^(b) are B-spline derivative
    fields of (w P) and w (eval_direct).  The reference itself rejects NURBS
    derivatives (src/validation.jl:81-88)."""
    import itertools

    deriv = tuple(int(x) for x in deriv)
    nin = len(dims)
    Wf = np.asarray(W, dtype=np.float64)
    H = np.asarray(P, dtype=np.float64) * Wf[..., None]
    lower = sorted(itertools.product(*[range(a + 1) for a in deriv]), key=lambda b: (sum(b), b))
    A = {b: eval_direct(dims, H, deriv=b, nthreads=nthreads) for b in lower}
    w = {b: eval_direct(dims, Wf[..., None], deriv=b, nthreads=nthreads)[..., 0] for b in lower}
    R = {}
    for g in lower:
        acc = A[g].copy()
        for b in lower:
            if sum(b) == 0 or any(b[d] > g[d] for d in range(nin)):
                continue
            c = 1
            for d in range(nin):
                c *= _binom(g[d], b[d])
            rest = tuple(g[d] - b[d] for d in range(nin))
            acc -= c * w[b][..., None] * R[rest]
        R[g] = acc / w[(0,) * nin][..., None]
    return R[deriv]


def adjoint_rational(dims: Sequence[Dim], W: np.ndarray, ebar: np.ndarray,
                     nthreads: int = 1) -> np.ndarray:
    """Adjoint of the NURBS map in its control points, weights fixed, FP64:
    R = B (w P) / w(x) is linear in P, so R^T e = w * B^T (e / w(x)).  The
    reference's evaluate_adjoint! rejects NURBS (src/adjoint.jl:57)."""
    Wf = np.asarray(W, dtype=np.float64)
    wx = eval_direct(dims, Wf[..., None], nthreads=nthreads)[..., 0]
    q = adjoint_direct(dims, np.asarray(ebar, dtype=np.float64) / wx[..., None])
    return Wf[..., None] * q
