"""SplineGridBatch: many FP32 grids evaluated by one persistent launch.

The reference evaluates one grid per evaluate! call; a caller with many
small grids loops over them.  A batch builds every grid's tables once and
evaluates all of them with one persistent kernel fed by a device work queue
(sg_batch_* in include/splinegrids_hip.h).  evaluate_ launches only kernels,
so it can be captured in a torch.cuda.CUDAGraph and replayed.
"""
from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np
import torch

from . import _lib
from .control_points import obtain
from .spline_grid import SplineGrid
from .utils import current_stream_handle, is_colmajor


class SplineGridBatch:
    def __init__(self, grids: Sequence[SplineGrid]):
        self.grids = list(grids)
        if not self.grids:
            raise ValueError("a batch needs at least one grid")
        dev = self.grids[0].device
        for g in self.grids:
            if g.dtype != torch.float32:
                raise TypeError("batches take FP32 grids")
            if g.device != dev:
                raise ValueError("all grids of a batch must be on one device")
        self.device = dev
        L = _lib.lib()
        self._lib = L
        keep, descs = [], (_lib.GridDesc * len(self.grids))()
        index = dev.index if dev.index is not None else torch.cuda.current_device()
        for i, g in enumerate(self.grids):
            dd = (_lib.DimDesc * g.Nin)()
            for d, s in enumerate(g.spline_dimensions):
                knots = np.ascontiguousarray(s.host_knots(), dtype=np.float64)
                t = np.ascontiguousarray(s.host_sample_points(), dtype=np.float64)
                keep += [knots, t]
                dd[d] = _lib.DimDesc(s.degree, s.n_basis_functions, s.n_sample_points, 0,
                                     knots.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                     t.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
            keep.append(dd)
            descs[i] = _lib.GridDesc(g.Nin, g.Nout, int(g.is_nurbs), index, dd)
        handle = ctypes.c_void_p()
        with torch.cuda.device(dev):
            _lib.check(L.sg_batch_create(descs, len(self.grids), current_stream_handle(dev),
                                         ctypes.byref(handle)))
        self.handle = handle
        self.set_arrays()

    def set_arrays(self) -> None:
        """Point the batch at every grid's current control points, weights and eval."""
        n = len(self.grids)
        P = (ctypes.c_void_p * n)()
        W = (ctypes.c_void_p * n)()
        E = (ctypes.c_void_p * n)()
        for i, g in enumerate(self.grids):
            cp = obtain(g.control_points)
            for name, t in (("control_points", cp), ("eval", g.eval)):
                if not is_colmajor(t):
                    raise ValueError(f"grid {i}: {name} must be column-major")
            P[i] = cp.data_ptr() or None
            W[i] = g.weights.data_ptr() if g.weights is not None else None
            E[i] = g.eval.data_ptr() or None
        _lib.check(self._lib.sg_batch_set_arrays(self.handle, P, W, E))

    def evaluate_(self, tier: str = "native") -> None:
        """evaluate! of every grid (derivative order 0), one persistent launch."""
        with torch.cuda.device(self.device):
            _lib.check(self._lib.sg_batch_evaluate(self.handle, _lib.TIERS[tier],
                                                   current_stream_handle(self.device)))

    def evaluate_each_(self) -> None:
        """The per-grid launch loop a batch replaces (for comparison)."""
        with torch.cuda.device(self.device):
            _lib.check(self._lib.sg_batch_evaluate_each(self.handle,
                                                        current_stream_handle(self.device)))

    def autotune(self, reps: int = 3) -> dict:
        with torch.cuda.device(self.device):
            _lib.check(self._lib.sg_batch_autotune(self.handle, reps,
                                                   current_stream_handle(self.device)))
        return self.info()

    def info(self) -> dict:
        nt, ns = ctypes.c_int64(), ctypes.c_int64()
        npers, rows, bpc = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        _lib.check(self._lib.sg_batch_info(self.handle, ctypes.byref(nt), ctypes.byref(ns),
                                           ctypes.byref(npers), ctypes.byref(rows),
                                           ctypes.byref(bpc)))
        return {"tiles": nt.value, "samples": ns.value, "persistent_grids": npers.value,
                "grids": len(self.grids), "rows_per_wave": rows.value,
                "blocks_per_cu": bpc.value}

    def describe(self) -> str:
        """Tiling, LDS per workgroup and resident workgroups per CU, one line."""
        buf = ctypes.create_string_buffer(1024)
        _lib.check(self._lib.sg_batch_describe(self.handle, buf, len(buf)))
        return buf.value.decode()

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                self._lib.sg_batch_destroy(h)
            except Exception:
                pass
            self.handle = None
