"""RefinementMatrix, rmeye, mult! and mult_adjoint!: src/refinement_matrix.jl
and src/adjoint.jl:85-152.

A refinement matrix is stored as the reference stores it: m x n with
consecutive non-zeros in every row and column, 1-based ``row_pointer[m]`` and
``column_start[m]`` (int32) and the non-zeros of every row in ``nzval``, all
device tensors.  Indices into a matrix (``A[i, j]``) are 1-based, as in the
reference, since they address this stored format.  Validation, products and
the mode products with arrays run in libsplinegrids_hip.so.
"""
from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np
import torch

from . import _lib
from .utils import current_stream_handle, dtype_code, np_float, torch_float


def _int_tensor(x, device) -> torch.Tensor:
    if isinstance(x, torch.Tensor):
        return x.to(device=device, dtype=torch.int32).contiguous()
    return torch.from_numpy(np.ascontiguousarray(np.asarray(x), dtype=np.int32)).to(device)


class RefinementMatrix:
    """``RefinementMatrix(m, n, row_pointer, column_start, nzval)``
    (src/refinement_matrix.jl:19-57) or ``RefinementMatrix(A)`` for a dense
    2-D host array (:476-504).

    The constructor checks that row_pointer is sorted and validates every
    row on the GPU (validate_refinement_matrix_kernel, :134-181), raising
    ``ValueError("Invalid rows: [i, ...].")`` with the 1-based invalid rows,
    the reference's message.  ``device`` places host inputs (default: the
    device of nzval when it is a tensor, else "cuda")."""

    def __init__(self, *args, device=None):
        if len(args) == 1:
            m, n, rp, cs, nz = _from_dense(np.asarray(args[0]))
        elif len(args) == 5:
            m, n, rp, cs, nz = args
        else:
            raise TypeError("RefinementMatrix(m, n, row_pointer, column_start, nzval) "
                            "or RefinementMatrix(A)")
        if device is None:
            device = nz.device if isinstance(nz, torch.Tensor) else "cuda"
        device = torch.device(device)
        if isinstance(nz, torch.Tensor):
            nzval = nz.to(device).contiguous()
        else:
            nz = np.asarray(nz)
            if not np.issubdtype(nz.dtype, np.floating):
                nz = nz.astype(np.float64)
            nzval = torch.from_numpy(np.ascontiguousarray(nz)).to(device)
        if nzval.dtype not in (torch.float32, torch.float64):
            raise TypeError(f"unsupported value type {nzval.dtype}")
        row_pointer = _int_tensor(rp, device)
        column_start = _int_tensor(cs, device)
        m, n = int(m), int(n)
        # src/refinement_matrix.jl:31-32
        assert len(row_pointer) == len(column_start) == m
        rp_host = row_pointer.cpu().numpy()
        assert np.array_equal(rp_host, np.sort(rp_host))
        self.m, self.n = m, n
        self.row_pointer, self.column_start, self.nzval = row_pointer, column_start, nzval
        invalid = self.invalid_rows()
        if invalid:
            raise ValueError(f"Invalid rows: {invalid}.")

    # ---- sizes and comparison (src/refinement_matrix.jl:59-70) ----
    @property
    def shape(self):
        return (self.m, self.n)

    def size(self):
        return (self.m, self.n)

    def __len__(self) -> int:
        return self.m * self.n

    @property
    def dtype(self):
        return self.nzval.dtype

    @property
    def device(self):
        return self.nzval.device

    @property
    def nnz(self) -> int:
        return int(self.nzval.numel())

    def __eq__(self, other) -> bool:
        if not isinstance(other, RefinementMatrix):
            return NotImplemented
        return (self.shape == other.shape and self.dtype == other.dtype
                and torch.equal(self.row_pointer.cpu(), other.row_pointer.cpu())
                and torch.equal(self.column_start.cpu(), other.column_start.cpu())
                and torch.equal(self.nzval.cpu(), other.nzval.cpu()))

    __hash__ = None

    def desc(self) -> _lib.RefmatDesc:
        return _lib.RefmatDesc(self.m, self.n, self.nnz, self.row_pointer.data_ptr(),
                               self.column_start.data_ptr(), self.nzval.data_ptr())

    def invalid_rows(self) -> list:
        """The 1-based rows that fail validate_refinement_matrix_kernel
        (:134-181): no non-zero, past the last column, or a start or end
        before the previous row's."""
        if self.m == 0:
            return []
        L = _lib.lib()
        flags = torch.zeros(self.m, dtype=torch.uint8, device=self.device)
        d = self.desc()
        with torch.cuda.device(self.device):
            _lib.check(L.sg_refinement_validate(ctypes.byref(d), flags.data_ptr(),
                                                current_stream_handle(self.device)))
        f = flags.cpu().numpy()
        return [int(i) + 1 for i in np.flatnonzero(f == 0)]

    def _host(self):
        return (self.row_pointer.cpu().numpy(), self.column_start.cpu().numpy(),
                self.nzval.cpu().numpy())

    def __getitem__(self, ij):
        """A[i, j], 1-based (src/refinement_matrix.jl:72-86)."""
        i, j = (int(x) for x in ij)
        if not (1 <= i <= self.m and 1 <= j <= self.n):
            raise IndexError(f"Index ({i}, {j}) out of bounds for refinement matrix of size "
                             f"({self.m}, {self.n}).")
        rp, cs, nz = self._host()
        start = int(cs[i - 1])
        nxt = int(rp[i]) if i < self.m else len(nz) + 1
        end = start + nxt - int(rp[i - 1]) - 1
        if start <= j <= end:
            return nz[int(rp[i - 1]) + j - start - 1].item()
        return nz.dtype.type(0).item()

    def collect(self) -> np.ndarray:
        """The dense matrix (Base.collect, :329-361), as a host array."""
        rp, cs, nz = self._host()
        out = np.zeros((self.m, self.n), dtype=nz.dtype)
        for i in range(self.m):
            nxt = int(rp[i + 1]) if i + 1 < self.m else len(nz) + 1
            k = nxt - int(rp[i])
            out[i, cs[i] - 1:cs[i] - 1 + k] = nz[rp[i] - 1:rp[i] - 1 + k]
        return out

    def to(self, device=None, dtype=None) -> "RefinementMatrix":
        """Adapt.adapt (:88-101): the same matrix on another device or with
        another value type."""
        device = self.device if device is None else torch.device(device)
        dtype = self.dtype if dtype is None else torch_float(dtype)
        if device == self.device and dtype == self.dtype:
            return self
        return RefinementMatrix(self.m, self.n, self.row_pointer.to(device),
                                self.column_start.to(device), self.nzval.to(device, dtype))

    def __matmul__(self, other: "RefinementMatrix") -> "RefinementMatrix":
        """C = A B (Base.:*, :273-327): the non-zero extents of C's rows, then
        its values, on the GPU."""
        A, B = self, other
        if A.n != B.m:
            raise ValueError("DimensionMismatch: Inner dimensions must match")
        if A.dtype != B.dtype or A.device != B.device:
            raise TypeError("both refinement matrices must share value type and device")
        L = _lib.lib()
        dev = A.device
        n_nonzero = torch.empty(A.m, dtype=torch.int32, device=dev)
        column_start = torch.empty(A.m, dtype=torch.int32, device=dev)
        da, db = A.desc(), B.desc()
        with torch.cuda.device(dev):
            stream = current_stream_handle(dev)
            _lib.check(L.sg_refinement_product_rows(ctypes.byref(da), ctypes.byref(db),
                                                    n_nonzero.data_ptr(),
                                                    column_start.data_ptr(), stream))
            row_pointer = torch.ones(A.m, dtype=torch.int32, device=dev)
            if A.m > 1:
                row_pointer[1:] += torch.cumsum(n_nonzero[:-1], 0, dtype=torch.int32)
            nnz = int(n_nonzero.sum().item())
            nzval = torch.zeros(nnz, dtype=A.dtype, device=dev)
            dc = _lib.RefmatDesc(A.m, B.n, nnz, row_pointer.data_ptr(), column_start.data_ptr(),
                                 nzval.data_ptr())
            _lib.check(L.sg_refinement_product_values(ctypes.byref(da), ctypes.byref(db),
                                                      ctypes.byref(dc), dtype_code(A.dtype),
                                                      stream))
        return RefinementMatrix(A.m, B.n, row_pointer, column_start, nzval)

    __mul__ = __matmul__

    def __repr__(self):
        return f"RefinementMatrix({self.m}x{self.n}, {self.nnz} non-zeros, {self.dtype})"


def _from_dense(A: np.ndarray):
    """The stored form of a dense matrix (:476-504): per row the span from the
    first to the last non-zero; a zero row gets column_start 0, which fails
    validation as in the reference."""
    assert A.ndim == 2
    rp, cs, nz = [], [], []
    pointer = 1
    for row in A:
        nzi = np.flatnonzero(row)
        rp.append(pointer)
        if len(nzi) == 0:
            cs.append(0)
        else:
            cs.append(int(nzi[0]) + 1)
            nz.extend(row[nzi[0]:nzi[-1] + 1])
            pointer += int(nzi[-1] - nzi[0] + 1)
    dt = A.dtype if np.issubdtype(A.dtype, np.floating) else np.float64
    return A.shape[0], A.shape[1], rp, cs, np.asarray(nz, dtype=dt)


def rmeye(n: int, *, device="cuda", float_type=np.float32, int_type=np.int32) -> RefinementMatrix:
    """The n x n identity (src/refinement_matrix.jl:463-474)."""
    assert np.dtype(int_type) == np.int32, "refinement matrices use int32 indices"
    idx = np.arange(1, n + 1, dtype=np.int32)
    return RefinementMatrix(n, n, idx, idx.copy(), np.ones(n, dtype=np_float(float_type)),
                            device=device)


def _mult_call(Y: torch.Tensor, mats: Sequence[RefinementMatrix], B: torch.Tensor,
               dims: Sequence[int]):
    mats, dims = list(mats), [int(d) for d in dims]
    # validate_mult_input (src/refinement_matrix.jl:393-419)
    assert len(mats) == len(dims), "one refinement dimension per refinement matrix"
    assert Y.ndim == B.ndim
    for A, d in zip(mats, dims):
        assert 1 <= d < Y.ndim, "the last dimension (the channel) is never refined"
        assert (Y.shape[d - 1], B.shape[d - 1]) == A.shape, \
            "refinement matrix size does not match the arrays"
    descs = (_lib.RefmatDesc * max(len(mats), 1))()
    for k, A in enumerate(mats):
        descs[k] = A.desc()
    return (Y.ndim, _lib.i32_array(Y.shape), _lib.i32_array(B.shape), len(mats), descs,
            _lib.i32_array(dims))


def _dtype_of(t: torch.Tensor) -> int:
    if t.dtype == torch.uint8:
        return _lib.SG_U8
    return dtype_code(t.dtype)


def mult_(Y: torch.Tensor, refinement_matrices: Sequence[RefinementMatrix], B: torch.Tensor,
          dims_refinement: Sequence[int]) -> None:
    """mult!(Y, As, B, dims_refinement) (src/refinement_matrix.jl:421-445):
    Y = B times refinement_matrices[k] along (1-based) dimension
    dims_refinement[k]; Y is overwritten.  Column-major device tensors."""
    from .utils import is_colmajor

    assert is_colmajor(Y) and is_colmajor(B), "arrays must be column-major"
    assert Y.dtype == B.dtype and all(A.dtype == Y.dtype for A in refinement_matrices)
    args = _mult_call(Y, refinement_matrices, B, dims_refinement)
    L = _lib.lib()
    with torch.cuda.device(Y.device):
        _lib.check(L.sg_refinement_mult(*args, B.data_ptr(), Y.data_ptr(), dtype_code(Y.dtype),
                                        current_stream_handle(Y.device)))


def mult_adjoint_(B: torch.Tensor, refinement_matrices: Sequence[RefinementMatrix],
                  Y: torch.Tensor, dims_refinement: Sequence[int], *,
                  accumulate: bool = False) -> None:
    """mult_adjoint!(B, As, Y, dims_refinement) (src/adjoint.jl:127-152):
    B = M^T Y, overwritten (or added to with ``accumulate=True``), without
    atomics and in a fixed summation order.  uint8 arrays (0/1) give the
    structural form that deactivate_overwritten_control_points! uses."""
    from .utils import is_colmajor

    assert is_colmajor(Y) and is_colmajor(B), "arrays must be column-major"
    assert Y.dtype == B.dtype
    if Y.dtype != torch.uint8:
        assert all(A.dtype == Y.dtype for A in refinement_matrices)
    args = _mult_call(Y, refinement_matrices, B, dims_refinement)
    L = _lib.lib()
    with torch.cuda.device(Y.device):
        _lib.check(L.sg_refinement_mult_adjoint(*args, Y.data_ptr(), B.data_ptr(), _dtype_of(Y),
                                                int(bool(accumulate)),
                                                current_stream_handle(Y.device)))


__all__ = ["RefinementMatrix", "rmeye", "mult_", "mult_adjoint_"]
