"""Small helpers: value types, column-major device tensors, streams.

Arrays follow the reference's Julia layout: column-major, first index
fastest.  A column-major tensor of shape (a, b, c) is a permuted view of a
contiguous (c, b, a) tensor, so ``t[i, j, k]`` indexes like Julia's
``t[i+1, j+1, k+1]`` and ``t.data_ptr()`` is the column-major buffer the
C-ABI expects.
"""
from __future__ import annotations

import numpy as np
import torch

_NP_FLOAT = {np.float32: np.float32, np.float64: np.float64,
             torch.float32: np.float32, torch.float64: np.float64,
             "float32": np.float32, "float64": np.float64}
_NP_INT = {np.int32: np.int32, np.int64: np.int64,
           torch.int32: np.int32, torch.int64: np.int64}


def np_float(t):
    if isinstance(t, np.dtype):
        t = t.type
    if t not in _NP_FLOAT:
        raise TypeError(f"unsupported float type {t}")
    return _NP_FLOAT[t]


def np_int(t):
    if isinstance(t, np.dtype):
        t = t.type
    if t not in _NP_INT:
        raise TypeError(f"unsupported int type {t}")
    return _NP_INT[t]


def torch_float(t):
    return torch.float32 if np_float(t) is np.float32 else torch.float64


def dtype_code(dtype) -> int:
    from ._lib import SG_F32, SG_F64

    if dtype == torch.float32:
        return SG_F32
    if dtype == torch.float64:
        return SG_F64
    raise TypeError(f"unsupported value type {dtype}; use float32 or float64")


def colmajor_empty(shape, dtype, device) -> torch.Tensor:
    shape = tuple(int(s) for s in shape)
    base = torch.empty(shape[::-1], dtype=dtype, device=device)
    return base.permute(*range(len(shape) - 1, -1, -1))


def colmajor_zeros(shape, dtype, device) -> torch.Tensor:
    t = colmajor_empty(shape, dtype, device)
    t.zero_()
    return t


def colmajor(t: torch.Tensor) -> torch.Tensor:
    """A column-major copy (or t itself when it already is)."""
    if is_colmajor(t):
        return t
    out = colmajor_empty(t.shape, t.dtype, t.device)
    out.copy_(t)
    return out


def is_colmajor(t: torch.Tensor) -> bool:
    if t.numel() == 0:
        return True
    expect = 1
    for size, stride in zip(t.shape, t.stride()):
        if size != 1 and stride != expect:
            return False
        expect *= size
    return True


def from_numpy_colmajor(a: np.ndarray, dtype, device) -> torch.Tensor:
    """Device tensor with a's values in Julia layout."""
    a = np.asarray(a)
    base = torch.from_numpy(np.ascontiguousarray(a.transpose()))
    return base.to(device=device, dtype=dtype).permute(*range(a.ndim - 1, -1, -1))


def to_numpy_colmajor(t: torch.Tensor) -> np.ndarray:
    """Host copy as a Fortran-ordered numpy array with the same indexing."""
    return np.asfortranarray(t.detach().cpu().numpy())


def current_stream_handle(device) -> int:
    return torch.cuda.current_stream(device).cuda_stream
