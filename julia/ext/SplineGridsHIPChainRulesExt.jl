# SplineGridsHIPChainRulesExt -- a ChainRulesCore rule for HIP grids.
#
# Loaded when AMDGPU and ChainRulesCore are (Project.toml:
# SplineGridsHIPChainRulesExt = ["AMDGPU", "ChainRulesCore"]).  ChainRules-based
# AD (Zygote and the like) cannot differentiate in-place evaluate!, so the rule
# is on the non-mutating spline_eval(grid, control_points; derivative_order).
# The map is linear in control_points (NURBS: weights fixed), so the pullback
# is exactly the adjoint: control_points_bar = Mᵀ eval_bar.  Basis tables,
# knots and weights get no tangent.
module SplineGridsHIPChainRulesExt

using SplineGrids
using SplineGrids: SplineGrid, spline_eval, hip_evaluate_adjoint!
using AMDGPU
using ChainRulesCore

const HIPGrid = SplineGrid{Nin, Nout, Tv, Ti, S, C, W, <:ROCArray, HasWeights} where {
    Nin, Nout, Tv, Ti, S, C, W, HasWeights}

# a cotangent on the device in the grid's value type (Fill-like cotangents of
# sum() and host arrays are materialised once)
device_cotangent(ȳ::ROCArray{Tv}, ::Type{Tv}) where {Tv} = ȳ
device_cotangent(ȳ, ::Type{Tv}) where {Tv} = ROCArray(collect(Tv, ȳ))

function ChainRulesCore.rrule(::typeof(spline_eval), grid::HIPGrid{Nin, Nout, Tv},
        control_points::ROCArray{Tv};
        derivative_order::NTuple{Nin, <:Integer} = ntuple(_ -> 0, Nin)) where {Nin, Nout, Tv}
    y = spline_eval(grid, control_points; derivative_order)
    function spline_eval_pullback(ȳ)
        ȳ = unthunk(ȳ)
        ȳ isa AbstractZero && return (NoTangent(), NoTangent(), ZeroTangent())
        p̄ = similar(control_points)
        hip_evaluate_adjoint!(grid, derivative_order, device_cotangent(ȳ, Tv), p̄)
        return (NoTangent(), NoTangent(), p̄)
    end
    return y, spline_eval_pullback
end

end # module SplineGridsHIPChainRulesExt
