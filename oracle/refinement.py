"""CPU oracle for refinement matrices, knot insertion and THB control points.

TEST INFRASTRUCTURE ONLY, like the rest of ``oracle/``: the checker of the
HIP refinement kernels (splinegrids.jl_amd/csrc/sg_refine.hip).  The product
package never imports it.

Reference functions restated, and what pins each one (tests/test_oracle_refinement.py):

==========================  ===============================================  =============================================
oracle function             reference                                        pinned by
==========================  ===============================================  =============================================
RefMat.invalid_rows         validate_refinement_matrix_kernel                the reference's own tests (test/test_refinement_
                            (src/refinement_matrix.jl:134-181)               matrix.jl:61-75: "Invalid rows: [2]." and
                                                                             "[2, 3].")
RefMat.from_dense / dense   RefinementMatrix(::Matrix) (:476-504),           round trip dense -> RefMat -> dense
                            collect (:329-361)
product                     Base.:* with its two kernels (:184-327)          dense products of random banded matrices
                                                                             (test/test_refinement_matrix.jl:46-59),
                                                                             identity and diagonal cases (:17-44)
insert_knot_matrix          build_refinement_matrix_kernel + RefinementMatrix the spline is unchanged by knot insertion
                            (src/refinement.jl:3-88)                         (test/test_refinement.jl:29-40): the oracle
                                                                             evaluation before and after agree to 1e-13
insert_knot_values          insert_knot(::KnotVector, knot_new) (:107-133)   expanded knots equal the sorted union
refine_dimension            refine(::SplineDimension; knots_new) (:256-284)  geometry invariance (test/test_refinement.jl:
                                                                             42-58), product of single insertions
mult / mult_adjoint         mult! (src/refinement_matrix.jl:421-445),        dense mode products (numpy Kronecker form),
                            mult_adjoint! (src/adjoint.jl:127-152)           <R x, y> = <x, R^T y>
mult_adjoint_flags          the Flag scatter of deactivate_overwritten_      brute force on the dense structure
                            control_points! (src/control_points.jl:624-680)
Thb (evaluate, adjoint,     src/control_points.jl:319-349 (evaluate!),       the transpose identity of the whole THB map,
activate, deactivate)       src/adjoint.jl:172-205 (evaluate_adjoint!),      the reference's counts (test/test_local_
                            :470-680                                         refinement.jl:11-41: 36 -> 63 control points,
                                                                             geometry kept), G2 index set (:43-67)
==========================  ===============================================  =============================================
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

import oracle as O


# --------------------------------------------------------------------------
# refinement matrices (1-based storage, as the reference keeps them)
# --------------------------------------------------------------------------

@dataclass
class RefMat:
    m: int
    n: int
    row_pointer: np.ndarray   # int32 [m], 1-based
    column_start: np.ndarray  # int32 [m], 1-based
    nzval: np.ndarray         # float64 [nnz]

    @property
    def nnz(self) -> int:
        return len(self.nzval)

    def row_range(self, i: int):
        """1-based (first column, last column) of 1-based row i
        (get_column_range, src/refinement_matrix.jl:103-131)."""
        nxt = self.nnz + 1 if i == self.m else int(self.row_pointer[i])
        rp = int(self.row_pointer[i - 1])
        cs = int(self.column_start[i - 1])
        return cs, cs + (nxt - rp) - 1

    def invalid_rows(self) -> List[int]:
        """validate_refinement_matrix_kernel (src/refinement_matrix.jl:134-181),
        1-based rows that fail."""
        bad = []
        for i in range(1, self.m + 1):
            cs, ce = self.row_range(i)
            ok = ce >= cs                                     # :148
            if ok:
                ok = cs >= 1 and ce <= self.n                 # :151-153
            if i == 1:
                ok = cs == 1 and int(self.row_pointer[0]) == 1  # :159-161 (replaces the above)
            if ok and i > 1:
                csp, cep = self.row_range(i - 1)
                ok = csp <= cs <= cep + 1                     # :165-171
                if ok:
                    ok = ce >= cep                            # :176-178
            if not ok:
                bad.append(i)
        return bad

    def dense(self) -> np.ndarray:
        """collect (src/refinement_matrix.jl:329-361)."""
        out = np.zeros((self.m, self.n))
        for i in range(1, self.m + 1):
            cs, ce = self.row_range(i)
            rp = int(self.row_pointer[i - 1])
            for k, j in enumerate(range(cs, ce + 1)):
                out[i - 1, j - 1] = self.nzval[rp - 1 + k]
        return out

    @classmethod
    def from_dense(cls, A: np.ndarray) -> "RefMat":
        """RefinementMatrix(::Matrix) (src/refinement_matrix.jl:476-504),
        without the validation (invalid_rows reports it)."""
        rp, cs, nz = [], [], []
        pointer = 1
        for row in np.asarray(A, dtype=np.float64):
            nzi = np.flatnonzero(row)
            rp.append(pointer)
            if len(nzi) == 0:
                cs.append(0)
            else:
                cs.append(int(nzi[0]) + 1)
                nz.extend(row[nzi[0]:nzi[-1] + 1])
                pointer += int(nzi[-1] - nzi[0] + 1)
        return cls(A.shape[0], A.shape[1], np.array(rp, np.int32), np.array(cs, np.int32),
                   np.array(nz, np.float64))

    @classmethod
    def eye(cls, n: int) -> "RefMat":
        """rmeye (src/refinement_matrix.jl:463-474)."""
        r = np.arange(1, n + 1, dtype=np.int32)
        return cls(n, n, r, r.copy(), np.ones(n))


def product(A: RefMat, B: RefMat) -> RefMat:
    """C = A * B (src/refinement_matrix.jl:273-327): per row of A, the number
    of non-zeros of C and the first column by scanning columns of B from the
    right (refinement_matrix_mul_nonzeros_kernel, :232-271), then the values
    summed over k in A's row and j in C's row (:184-230)."""
    assert A.n == B.m
    n_nonzero = np.zeros(A.m, np.int32)
    col_start = np.zeros(A.m, np.int32)
    for i in range(1, A.m + 1):
        acs, ace = A.row_range(i)
        cnt, start = 0, 0
        for j in range(B.n, 0, -1):
            for k in range(acs, ace + 1):
                bcs, bce = B.row_range(k)
                if bcs <= j <= bce:
                    cnt += 1
                    start = j
                    break
        n_nonzero[i - 1] = cnt
        col_start[i - 1] = start
    row_pointer = np.ones(A.m, np.int32)
    row_pointer[1:] += np.cumsum(n_nonzero[:-1])
    C = RefMat(A.m, B.n, row_pointer, col_start, np.zeros(int(n_nonzero.sum())))
    for i in range(1, A.m + 1):
        acs, ace = A.row_range(i)
        ccs, cce = C.row_range(i)
        pa = int(A.row_pointer[i - 1])
        for k in range(acs, ace + 1):
            bcs, bce = B.row_range(k)
            pc = int(C.row_pointer[i - 1])
            for j in range(ccs, cce + 1):
                if bcs <= j <= bce:
                    pb = int(B.row_pointer[k - 1]) + j - bcs
                    C.nzval[pc - 1] += A.nzval[pa - 1] * B.nzval[pb - 1]
                pc += 1
            pa += 1
    return C


# --------------------------------------------------------------------------
# knot insertion
# --------------------------------------------------------------------------

def insert_knot_matrix(knots_all: np.ndarray, degree: int, knot_span_index: int,
                       knot_new: float) -> RefMat:
    """build_refinement_matrix_kernel (src/refinement.jl:3-39) and its caller
    RefinementMatrix(spline_dimension, knot_span_index, knot_new) (:53-84):
    the (n + 1) x n matrix expressing the old basis in the new one;
    knot_span_index is 1-based with knots_all[k] <= knot_new < knots_all[k + 1]."""
    knots_all = np.asarray(knots_all, dtype=np.float64)
    n = len(knots_all) - degree - 1
    m = n + 1
    rp = np.zeros(m, np.int32)
    cs = np.zeros(m, np.int32)
    nz = np.ones(n + degree + 1)
    for i in range(1, m + 1):
        if i <= knot_span_index - degree:
            rp[i - 1] = i
            cs[i - 1] = i
        elif i <= knot_span_index:
            alpha = (knot_new - knots_all[i - 1]) / (knots_all[i + degree - 1] - knots_all[i - 1])
            p0 = 2 * i - knot_span_index + degree - 1
            rp[i - 1] = p0
            cs[i - 1] = i - 1
            nz[p0 - 1] = 1 - alpha
            nz[p0] = alpha
        else:
            rp[i - 1] = i + degree
            cs[i - 1] = i - 1
    return RefMat(m, n, rp, cs, nz)


def insert_knot_values(values: np.ndarray, mults: np.ndarray, knot_new: float):
    """insert_knot(::KnotVector, knot_new) (src/refinement.jl:107-133):
    (new values, new multiplicities, knot_span_index)."""
    values = np.asarray(values)
    mults = np.asarray(mults, dtype=np.int32)
    assert not np.any(values == knot_new)
    idx = int(np.searchsorted(values, knot_new, side="left"))  # searchsortedfirst - 1
    new_values = np.insert(values, idx, knot_new)
    new_mults = np.insert(mults, idx, 1).astype(np.int32)
    span = int(mults[:idx].sum())
    return new_values, new_mults, span


def refine_dimension(values, mults, degree: int, knots_new=None):
    """refine(::SplineDimension; knots_new) (src/refinement.jl:256-284): insert
    the knots one by one (default: the midpoints of the knot spans), the
    refinement matrix the product of the single insertions, newest first."""
    values = np.asarray(values, dtype=np.float64)
    mults = np.asarray(mults, dtype=np.int32)
    if knots_new is None:
        knots_new = values[:-1] + np.diff(values) / 2
    n = int(mults.sum()) - degree - 1
    R = RefMat.eye(n)
    for k in np.asarray(knots_new, dtype=np.float64):
        knots_all = O.expand_knots(values, mults)
        values, mults, span = insert_knot_values(values, mults, float(k))
        R = product(insert_knot_matrix(knots_all, degree, span, float(k)), R)
    return values, mults, R


# --------------------------------------------------------------------------
# mult! / mult_adjoint! (C restatement in sg_oracle.c)
# --------------------------------------------------------------------------

class _SgoRefmat(ctypes.Structure):
    _fields_ = [("m", ctypes.c_int32), ("n", ctypes.c_int32), ("nnz", ctypes.c_int32),
                ("row_pointer", ctypes.POINTER(ctypes.c_int32)),
                ("column_start", ctypes.POINTER(ctypes.c_int32)),
                ("nzval", ctypes.POINTER(ctypes.c_double))]


def _mult_args(mats: Sequence[RefMat], dims: Sequence[int], shape_y, shape_b):
    nd = len(shape_y)
    keep = []
    arr = (_SgoRefmat * max(len(mats), 1))()
    for i, A in enumerate(mats):
        rp = np.ascontiguousarray(A.row_pointer, np.int32)
        cs = np.ascontiguousarray(A.column_start, np.int32)
        nz = np.ascontiguousarray(A.nzval, np.float64)
        keep += [rp, cs, nz]
        arr[i] = _SgoRefmat(A.m, A.n, A.nnz, rp.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                            cs.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                            nz.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    refmat_of = np.zeros(nd, np.int32)
    for i, d in enumerate(dims):  # dims 1-based, as the reference's dims_refinement
        refmat_of[d - 1] = i + 1
        assert shape_y[d - 1] == mats[i].m and shape_b[d - 1] == mats[i].n
    for d in range(nd):
        if refmat_of[d] == 0:
            assert shape_y[d] == shape_b[d]
    sy = np.array(shape_y, np.int32)
    sb = np.array(shape_b, np.int32)
    keep += [sy, sb, refmat_of]
    i32 = ctypes.POINTER(ctypes.c_int32)
    return keep, (ctypes.c_int32(nd), sy.ctypes.data_as(i32), sb.ctypes.data_as(i32), arr,
                  refmat_of.ctypes.data_as(i32))


def mult(mats: Sequence[RefMat], dims: Sequence[int], B: np.ndarray) -> np.ndarray:
    """mult!(Y, As, B, dims_refinement) (src/refinement_matrix.jl:421-445);
    B any numpy array (the last axis the channel), the result Fortran-ordered."""
    shape_b = B.shape
    shape_y = list(shape_b)
    for A, d in zip(mats, dims):
        shape_y[d - 1] = A.m
    Bf = np.asfortranarray(B, dtype=np.float64)
    Y = np.zeros(shape_y, order="F")
    keep, args = _mult_args(mats, dims, shape_y, shape_b)
    O.lib().sgo_refmat_mult(*args, Bf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                            Y.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    del keep
    return Y


def mult_adjoint(mats: Sequence[RefMat], dims: Sequence[int], Y: np.ndarray) -> np.ndarray:
    """mult_adjoint!(B, As, Y, dims_refinement) (src/adjoint.jl:127-152)."""
    shape_y = Y.shape
    shape_b = list(shape_y)
    for A, d in zip(mats, dims):
        shape_b[d - 1] = A.n
    Yf = np.asfortranarray(Y, dtype=np.float64)
    B = np.zeros(shape_b, order="F")
    keep, args = _mult_args(mats, dims, shape_y, shape_b)
    O.lib().sgo_refmat_mult_adjoint(*args, Yf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                    B.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    del keep
    return B


def mult_adjoint_flags(mats: Sequence[RefMat], dims: Sequence[int], Y: np.ndarray) -> np.ndarray:
    """The Flag scatter of deactivate_overwritten_control_points!
    (src/control_points.jl:624-680): uint8 0/1 in and out."""
    shape_y = Y.shape
    shape_b = list(shape_y)
    for A, d in zip(mats, dims):
        shape_b[d - 1] = A.n
    Yf = np.asfortranarray(Y, dtype=np.uint8)
    B = np.zeros(shape_b, dtype=np.uint8, order="F")
    keep, args = _mult_args(mats, dims, shape_y, shape_b)
    u8 = ctypes.POINTER(ctypes.c_uint8)
    O.lib().sgo_refmat_mult_adjoint_flags(*args, Yf.ctypes.data_as(u8), B.ctypes.data_as(u8))
    del keep
    return B


def overwrite(cp: np.ndarray, indices: np.ndarray, values: np.ndarray) -> None:
    """local_refinement_kernel (src/control_points.jl:296-311) in place on a
    Fortran-ordered float64 cp; indices [n, Nin] 1-based, values [n, Nout]."""
    assert cp.flags.f_contiguous and cp.dtype == np.float64
    nin, nout = cp.ndim - 1, cp.shape[-1]
    idx = np.asfortranarray(indices, dtype=np.int32)
    val = np.asfortranarray(values, dtype=np.float64)
    shape = np.array(cp.shape[:-1], np.int32)
    O.lib().sgo_local_refinement_overwrite(
        ctypes.c_int32(nin), ctypes.c_int32(nout), shape.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
        ctypes.c_int64(len(idx)), idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
        val.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
        cp.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))


def gather(cp: np.ndarray, indices: np.ndarray) -> np.ndarray:
    """local_refinement_adjoint_kernel (src/adjoint.jl:154-170): returns the
    values at the indices and zeroes them in cp (Fortran float64, in place)."""
    assert cp.flags.f_contiguous and cp.dtype == np.float64
    nin, nout = cp.ndim - 1, cp.shape[-1]
    idx = np.asfortranarray(indices, dtype=np.int32)
    vals = np.zeros((len(idx), nout), order="F")
    shape = np.array(cp.shape[:-1], np.int32)
    O.lib().sgo_local_refinement_gather(
        ctypes.c_int32(nin), ctypes.c_int32(nout), shape.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
        ctypes.c_int64(len(idx)), idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
        vals.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
        cp.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return vals


# --------------------------------------------------------------------------
# THB control points (LocallyRefinedControlPoints)
# --------------------------------------------------------------------------

@dataclass
class Level:
    """One LocalRefinement (src/control_points.jl:60-113): the matrices that
    map the previous level to this one (empty for the base level), and the
    control points this level overwrites."""
    dims: List[int]
    mats: List[RefMat]
    indices: np.ndarray  # int32 [k, Nin], 1-based
    values: np.ndarray   # float64 [k, Nout]


@dataclass
class Thb:
    """LocallyRefinedControlPoints (src/control_points.jl:151-185): the dense
    arrays of every level (level 0 the base) and the local refinements."""
    arrays: List[np.ndarray]
    levels: List[Level] = field(default_factory=list)

    def n_control_points(self) -> int:
        """get_n_control_points (src/control_points.jl:279-287)."""
        return sum(len(lv.indices) for lv in self.levels)

    def flat_values(self) -> np.ndarray:
        """The vcat of every level's values (get_control_point_view, :187-195)."""
        return np.concatenate([lv.values for lv in self.levels], axis=0)

    def set_flat_values(self, M: np.ndarray) -> None:
        """copyto!(control_points, M) (:211-213)."""
        k = 0
        for lv in self.levels:
            n = len(lv.values)
            lv.values = np.array(M[k:k + n], dtype=np.float64)
            k += n

    def evaluate(self) -> np.ndarray:
        """evaluate!(::LocallyRefinedControlPoints) (src/control_points.jl:319-349):
        each level is the previous one times its refinement matrices, then its
        own control points are overwritten; returns the last level."""
        for i, lv in enumerate(self.levels):
            if i > 0:
                self.arrays[i] = mult(lv.mats, lv.dims, self.arrays[i - 1])
            if len(lv.indices):
                overwrite(self.arrays[i], lv.indices, lv.values)
        return self.arrays[-1]

    def evaluate_adjoint(self, last: np.ndarray) -> np.ndarray:
        """evaluate_adjoint!(::LocallyRefinedControlPoints) (src/adjoint.jl:174-205)
        with the last level's array given: levels in reverse, each gathers
        (and zeroes) its own control points, then maps back through the
        transposed refinement matrices.  Returns the flat values."""
        self.arrays[-1] = np.array(last, dtype=np.float64, order="F", copy=True)
        for i in range(len(self.levels) - 1, -1, -1):
            lv = self.levels[i]
            if len(lv.indices):
                lv.values = gather(self.arrays[i], lv.indices)
            if i > 0:
                self.arrays[i - 1] = mult_adjoint(lv.mats, lv.dims, self.arrays[i])
        return self.flat_values()

    def activate(self, indices: np.ndarray, level: Optional[int] = None) -> None:
        """activate_local_refinement! (src/control_points.jl:470-519): the union
        of the level's indices and the new ones (first occurrence order, as
        unique(...; dims = 1)), new values read from that level's array."""
        level = len(self.levels) - 1 if level is None else level
        lv = self.levels[level]
        allidx = np.concatenate([lv.indices.reshape(-1, indices.shape[1]), indices], axis=0)
        seen, keep = set(), []
        for r in allidx:
            t = tuple(int(x) for x in r)
            if t not in seen:
                seen.add(t)
                keep.append(t)
        new_idx = np.array(keep, dtype=np.int32).reshape(-1, indices.shape[1])
        arr = self.arrays[level]
        vals = np.zeros((len(new_idx), arr.shape[-1]))
        vals[:len(lv.values)] = lv.values
        for i in range(len(lv.values), len(new_idx)):
            vals[i] = arr[tuple(int(x) - 1 for x in new_idx[i])]
        lv.indices, lv.values = new_idx, vals

    def deactivate_overwritten(self) -> None:
        """deactivate_overwritten_control_points! (src/control_points.jl:584-680)
        for every level but the last, in reverse."""
        for level in range(len(self.levels) - 2, -1, -1):
            lv, nxt = self.levels[level], self.levels[level + 1]
            shape_next = self.arrays[level + 1].shape[:-1] + (1,)
            flags_next = np.ones(shape_next, dtype=np.uint8, order="F")
            for r in nxt.indices:
                flags_next[tuple(int(x) - 1 for x in r) + (0,)] = 0
            flags = mult_adjoint_flags(nxt.mats, nxt.dims, flags_next)
            alive = np.array([bool(flags[tuple(int(x) - 1 for x in r) + (0,)])
                              for r in lv.indices], dtype=bool)
            lv.indices = lv.indices[alive]
            lv.values = lv.values[alive]
