"""ctypes binding of libsplinegrids_hip.so (include/splinegrids_hip.h).

The library is the only compute path: if it is missing or no HIP device is
visible, every entry point raises.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsplinegrids_hip.so")

SG_OK, SG_ERR_INVALID, SG_ERR_HIP, SG_ERR_UNSUPPORTED, SG_ERR_NO_MEMORY = 0, 1, 2, 3, 4
SG_F32, SG_F64, SG_U8 = 0, 1, 2
TIERS = {"native": 0, "split_f16": 1, "bf16": 2}
MAX_NIN = 8
MAX_FIELDS = 4

# every symbol include/splinegrids_hip.h declares
EXPORTS = (
    "sg_last_error", "sg_version", "sg_dimension_evaluate", "sg_dimension_evaluate_noalloc",
    "sg_plan_create",
    "sg_plan_destroy", "sg_plan_basis", "sg_plan_device_bytes", "sg_plan_describe", "sg_evaluate",
    "sg_evaluate_fields", "sg_evaluate_adjoint", "sg_stream_synchronize",
    "sg_batch_create", "sg_batch_destroy", "sg_batch_set_arrays", "sg_batch_evaluate",
    "sg_batch_evaluate_each", "sg_batch_autotune", "sg_batch_info", "sg_batch_describe",
    "sg_rational_workspace_bytes", "sg_evaluate_rational_fields",
    "sg_rational_adjoint_workspace_bytes", "sg_evaluate_rational_adjoint",
    "sg_refinement_insert_knot", "sg_refinement_product_rows", "sg_refinement_product_values",
    "sg_refinement_validate", "sg_refinement_mult", "sg_refinement_mult_adjoint",
    "sg_local_refinement_overwrite", "sg_local_refinement_gather",
)


class SplineGridsError(RuntimeError):
    """A C-ABI call returned a non-zero status; the message is sg_last_error()."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


class DimDesc(ctypes.Structure):
    _fields_ = [
        ("degree", ctypes.c_int32),
        ("n_basis", ctypes.c_int32),
        ("n_samples", ctypes.c_int32),
        ("max_derivative_order", ctypes.c_int32),
        ("knots_all", ctypes.POINTER(ctypes.c_double)),
        ("sample_points", ctypes.POINTER(ctypes.c_double)),
    ]


class RefmatDesc(ctypes.Structure):
    """sg_refmat: a refinement matrix's sizes and device arrays."""
    _fields_ = [
        ("m", ctypes.c_int32),
        ("n", ctypes.c_int32),
        ("nnz", ctypes.c_int32),
        ("row_pointer", ctypes.c_void_p),
        ("column_start", ctypes.c_void_p),
        ("nzval", ctypes.c_void_p),
    ]


class GridDesc(ctypes.Structure):
    _fields_ = [
        ("n_in", ctypes.c_int32),
        ("n_out", ctypes.c_int32),
        ("is_nurbs", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("dims", ctypes.POINTER(DimDesc)),
    ]


_lib = None


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Open the shared library and declare every signature (no GPU needed)."""
    if not os.path.exists(path):
        raise SplineGridsError(
            SG_ERR_UNSUPPORTED,
            f"{path} is not built; run __graft_entry__.build() first")
    L = ctypes.CDLL(path)
    i32, i64, vp = ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p
    i32p = ctypes.POINTER(ctypes.c_int32)
    f64p = ctypes.POINTER(ctypes.c_double)
    sig = {
        "sg_last_error": (ctypes.c_char_p, []),
        "sg_version": (ctypes.c_int, [i32p, i32p, i32p]),
        "sg_dimension_evaluate": (ctypes.c_int, [vp, i32, vp, i32, i32, i32, vp, vp, i32, vp]),
        "sg_dimension_evaluate_noalloc": (ctypes.c_int, [vp, i32, vp, i32, i32, i32, vp, vp, i32,
                                                         vp]),
        "sg_plan_create": (ctypes.c_int, [ctypes.POINTER(GridDesc), vp, ctypes.POINTER(vp)]),
        "sg_plan_destroy": (ctypes.c_int, [vp]),
        "sg_plan_basis": (ctypes.c_int, [vp, i32, f64p, i32p]),
        "sg_plan_device_bytes": (ctypes.c_int, [vp, ctypes.POINTER(i64)]),
        "sg_plan_describe": (ctypes.c_int, [vp, ctypes.c_char_p, i32]),
        "sg_evaluate": (ctypes.c_int, [vp, i32p, vp, vp, vp, i32, i32, vp]),
        "sg_evaluate_fields": (ctypes.c_int, [vp, i32, i32p, vp, ctypes.POINTER(vp), i32, i32, vp]),
        "sg_evaluate_adjoint": (ctypes.c_int, [vp, i32p, vp, vp, i32, i32, i32, vp]),
        "sg_stream_synchronize": (ctypes.c_int, [vp]),
        "sg_batch_create": (ctypes.c_int, [ctypes.POINTER(GridDesc), i32, vp, ctypes.POINTER(vp)]),
        "sg_batch_destroy": (ctypes.c_int, [vp]),
        "sg_batch_set_arrays": (ctypes.c_int, [vp, ctypes.POINTER(vp), ctypes.POINTER(vp),
                                               ctypes.POINTER(vp)]),
        "sg_batch_evaluate": (ctypes.c_int, [vp, i32, vp]),
        "sg_batch_evaluate_each": (ctypes.c_int, [vp, vp]),
        "sg_batch_autotune": (ctypes.c_int, [vp, i32, vp]),
        "sg_batch_info": (ctypes.c_int, [vp, ctypes.POINTER(i64), ctypes.POINTER(i64), i32p,
                                         i32p, i32p]),
        "sg_batch_describe": (ctypes.c_int, [vp, ctypes.c_char_p, i32]),
        "sg_rational_workspace_bytes": (ctypes.c_int, [vp, i32, i32p, i32, ctypes.POINTER(i64)]),
        "sg_evaluate_rational_fields": (ctypes.c_int, [vp, i32, i32p, vp, vp, ctypes.POINTER(vp),
                                                       vp, i64, i32, vp]),
        "sg_rational_adjoint_workspace_bytes": (ctypes.c_int, [vp, i32, ctypes.POINTER(i64)]),
        "sg_evaluate_rational_adjoint": (ctypes.c_int, [vp, vp, vp, vp, vp, i64, i32, i32, vp]),
        "sg_refinement_insert_knot": (ctypes.c_int, [vp, i32, i32, i32, ctypes.c_double, vp, vp,
                                                     vp, i32, vp]),
        "sg_refinement_product_rows": (ctypes.c_int, [ctypes.POINTER(RefmatDesc),
                                                      ctypes.POINTER(RefmatDesc), vp, vp, vp]),
        "sg_refinement_product_values": (ctypes.c_int, [ctypes.POINTER(RefmatDesc),
                                                        ctypes.POINTER(RefmatDesc),
                                                        ctypes.POINTER(RefmatDesc), i32, vp]),
        "sg_refinement_validate": (ctypes.c_int, [ctypes.POINTER(RefmatDesc), vp, vp]),
        "sg_refinement_mult": (ctypes.c_int, [i32, i32p, i32p, i32, ctypes.POINTER(RefmatDesc),
                                              i32p, vp, vp, i32, vp]),
        "sg_refinement_mult_adjoint": (ctypes.c_int, [i32, i32p, i32p, i32,
                                                      ctypes.POINTER(RefmatDesc), i32p, vp, vp,
                                                      i32, i32, vp]),
        "sg_local_refinement_overwrite": (ctypes.c_int, [i32, i32, i32p, i64, vp, vp, vp, i32,
                                                         vp]),
        "sg_local_refinement_gather": (ctypes.c_int, [i32, i32, i32p, i64, vp, vp, vp, i32, i32,
                                                       vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


def lib() -> ctypes.CDLL:
    """The loaded library; raises if it is not built or no GPU is visible."""
    global _lib
    if _lib is None:
        import torch

        if not torch.cuda.is_available():
            raise SplineGridsError(SG_ERR_UNSUPPORTED,
                                   "no HIP device is visible; this library runs on the GPU only")
        _lib = load()
    return _lib


def check(rc: int) -> None:
    if rc != SG_OK:
        msg = lib().sg_last_error().decode("utf-8", "replace")
        if rc == SG_ERR_INVALID:
            raise SplineGridsError(rc, msg)
        raise SplineGridsError(rc, f"splinegrids_hip error {rc}: {msg}")


def i32_array(values) -> ctypes.Array:
    vals = [int(v) for v in values]
    return (ctypes.c_int32 * max(len(vals), 1))(*vals)
