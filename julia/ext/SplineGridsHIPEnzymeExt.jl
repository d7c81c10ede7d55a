# SplineGridsHIPEnzymeExt -- Enzyme rules for the HIP grid evaluation.
#
# Loaded when AMDGPU and EnzymeCore are (Project.toml: SplineGridsHIPEnzymeExt =
# ["AMDGPU", "EnzymeCore"]).  Enzyme cannot see through the ccall into
# libsplinegrids_hip.so, so the rules are written by hand on the positional
# primitive hip_evaluate!(grid, derivative_order, control_points, eval) that
# evaluate!(::HIPGrid; ...) forwards to.  The keyword method itself stays
# plain Julia, which Enzyme differentiates down to the primitive; that is how
# the reference's reverse-mode test (test/test_EnzymeExt.jl:24-58:
# evaluate!(spline_grid; control_points = reshape(flat, ...)) with a Duplicated
# grid) reaches these rules.
#
# eval = M(grid, derivative_order) * control_points is linear in
# control_points (for NURBS grids with the weights fixed: R(P) = B(wP)/B(w)),
# so the rules are exact:
#   forward:  d_eval = M * d_control_points            (sg_evaluate on the shadows)
#   reverse:  d_control_points += Mᵀ * d_eval            (sg_evaluate_adjoint, accumulate = 1;
#                                                       the NURBS adjoint for NURBS grids)
#             d_eval = 0                                 (eval was overwritten, so its
#                                                       previous value has no influence)
# Basis tables, knots and weights are constants: their shadows are not
# accumulated (the reference's Enzyme-through-kernel path would also
# accumulate into the shadow basis tables).
module SplineGridsHIPEnzymeExt

using SplineGrids
using SplineGrids: SplineGrid, hip_evaluate!, hip_evaluate_adjoint!
using AMDGPU
using EnzymeCore
using EnzymeCore: Annotation, Const, Duplicated, BatchDuplicated, DuplicatedNoNeed,
                  BatchDuplicatedNoNeed
using EnzymeCore.EnzymeRules

const HIPGrid = SplineGrid{Nin, Nout, Tv, Ti, S, C, W, <:ROCArray, HasWeights} where {
    Nin, Nout, Tv, Ti, S, C, W, HasWeights}

# the shadows of an annotated argument as a tuple (empty for Const)
shadows(x::Const) = ()
shadows(x::Union{Duplicated, DuplicatedNoNeed}) = (x.dval,)
shadows(x::Union{BatchDuplicated, BatchDuplicatedNoNeed}) = x.dval

# ---- forward mode ---------------------------------------------------------

function EnzymeRules.forward(config::EnzymeRules.FwdConfig,
        func::Const{typeof(hip_evaluate!)}, ::Type{RT}, grid::Annotation{<:HIPGrid},
        derivative_order::Const, control_points::Annotation, eval::Annotation) where {RT}
    hip_evaluate!(grid.val, derivative_order.val, control_points.val, eval.val)
    dcp = shadows(control_points)
    for (i, dev) in enumerate(shadows(eval))
        if isempty(dcp)
            fill!(dev, 0)  # constant control points: the tangent of eval is zero
        else
            hip_evaluate!(grid.val, derivative_order.val, dcp[i], dev)
        end
    end
    return nothing  # hip_evaluate! returns nothing
end

# ---- reverse mode ---------------------------------------------------------

function EnzymeRules.augmented_primal(config::EnzymeRules.RevConfig,
        func::Const{typeof(hip_evaluate!)}, ::Type{RT}, grid::Annotation{<:HIPGrid},
        derivative_order::Const, control_points::Annotation, eval::Annotation) where {RT}
    hip_evaluate!(grid.val, derivative_order.val, control_points.val, eval.val)
    return EnzymeRules.AugmentedReturn(nothing, nothing, nothing)
end

function EnzymeRules.reverse(config::EnzymeRules.RevConfig,
        func::Const{typeof(hip_evaluate!)}, ::Type{RT}, tape, grid::Annotation{<:HIPGrid},
        derivative_order::Const, control_points::Annotation, eval::Annotation) where {RT}
    dcp = shadows(control_points)
    for (i, dev) in enumerate(shadows(eval))
        if !isempty(dcp)
            hip_evaluate_adjoint!(grid.val, derivative_order.val, dev, dcp[i];
                accumulate = true)
        end
        fill!(dev, 0)
    end
    return (nothing, nothing, nothing, nothing)
end

# ---- shadows of a HIP grid ------------------------------------------------

# As the reference's make_zero! (ext/SplineGridsEnzymeExt.jl:5-11) for grids
# on an AMD GPU: the shadow eval and the shadow basis tables are zeroed (on
# the device, without a host round trip).  The plan that the HIP extension
# keeps for a grid holds no differentiable state.
function EnzymeCore.make_zero!(spline_grid::HIPGrid)::Nothing
    fill!(spline_grid.eval, 0)
    for spline_dimension in spline_grid.spline_dimensions
        fill!(spline_dimension.eval, 0)
        fill!(spline_dimension.eval_prev, 0)
    end
    return nothing
end

end # module SplineGridsHIPEnzymeExt
