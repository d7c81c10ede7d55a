# hip_hooks.jl -- three generic functions for the HIP package extensions.
#
# To be included from src/SplineGrids.jl after spline_grid.jl and adjoint.jl
# (`include("hip_hooks.jl")`).  They own no methods here: the extension
# SplineGridsHIPExt (ext/SplineGridsHIPExt.jl) adds the methods for grids whose
# arrays live on an AMD GPU, and the AD extensions attach their rules to
# these functions.  The rules need positional arguments (Enzyme does not
# differentiate keyword arguments of a rule), which is why evaluate! on a HIP
# grid forwards to hip_evaluate! instead of carrying the rule itself.

"""
    hip_evaluate!(spline_grid, derivative_order, control_points, eval)

Positional form of `evaluate!(spline_grid; derivative_order, control_points, eval)`
for grids whose arrays live on an AMD GPU (methods in `SplineGridsHIPExt`).
`control_points` and `eval` are dense device arrays; `eval` is overwritten.
"""
function hip_evaluate! end

"""
    hip_evaluate_adjoint!(spline_grid, derivative_order, eval, control_points; accumulate = false)

Positional form of `evaluate_adjoint!` for HIP grids: `control_points = Mᵀ eval`
(overwritten, or added to with `accumulate = true`).  For NURBS grids this is
the adjoint of the rational map in its control points with the weights fixed
(derivative order 0), which `evaluate_adjoint!` itself does not offer.
"""
function hip_evaluate_adjoint! end

"""
    spline_eval(spline_grid, control_points; derivative_order) -> eval

Non-mutating evaluation into a new array of the shape of `spline_grid.eval`;
the form a ChainRulesCore `rrule` can be attached to.
"""
function spline_eval end
