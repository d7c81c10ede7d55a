"""MI355X (gfx950) grid evaluation for SplineGrids.jl's hot path.

The API mirrors the reference: KnotVector, SplineDimension, SplineGrid,
NURBSGrid, evaluate! (``evaluate_``) and evaluate_adjoint!
(``evaluate_adjoint_``), refinement matrices, knot insertion and locally
refined (THB) control points; the trailing underscore marks in-place calls,
as Julia's ``!`` does.  Arrays are torch tensors on a HIP device in Julia's
column-major layout.  All arithmetic runs in libsplinegrids_hip.so.
"""
from ._lib import SplineGridsError, lib
from .knot_vector import KnotVector
from .spline_dimension import SplineDimension, decompress
from .control_points import (DefaultControlPoints, LocalRefinement,
                             LocallyRefinedControlPoints, get_n_control_points, obtain)
from .spline_grid import (NURBSGrid, Plan, SplineGrid, evaluate_, evaluate_adjoint_,
                          evaluate_fields_, evaluate_rational_adjoint_,
                          evaluate_rational_fields_, reshape_colmajor, vec_colmajor)
from .validation import validate_partial_derivatives, validate_spline_grid
from .autograd import SplineEval, spline_eval
from .linear_map import LinearMap
from .batch import SplineGridBatch
from .refinement_matrix import RefinementMatrix, mult_, mult_adjoint_, rmeye
from .refinement import (activate_local_control_point_range_, activate_local_refinement_,
                         add_default_local_refinement, deactivate_overwritten_control_points_,
                         error_informed_local_refinement_, insert_knot, refine)
from .utils import colmajor, colmajor_empty, colmajor_zeros, from_numpy_colmajor, to_numpy_colmajor

__version__ = "0.1.0"

__all__ = [
    "SplineGridsError", "lib", "KnotVector", "SplineDimension", "decompress",
    "DefaultControlPoints", "get_n_control_points", "obtain", "NURBSGrid", "Plan",
    "SplineGrid", "evaluate_", "evaluate_adjoint_", "evaluate_fields_",
    "evaluate_rational_fields_", "evaluate_rational_adjoint_",
    "reshape_colmajor", "vec_colmajor", "validate_partial_derivatives",
    "validate_spline_grid", "colmajor", "colmajor_empty", "colmajor_zeros",
    "from_numpy_colmajor", "to_numpy_colmajor", "SplineEval", "spline_eval", "LinearMap",
    "SplineGridBatch", "RefinementMatrix", "rmeye", "mult_", "mult_adjoint_", "insert_knot",
    "refine", "LocalRefinement", "LocallyRefinedControlPoints", "add_default_local_refinement",
    "activate_local_refinement_", "activate_local_control_point_range_",
    "error_informed_local_refinement_", "deactivate_overwritten_control_points_",
]
