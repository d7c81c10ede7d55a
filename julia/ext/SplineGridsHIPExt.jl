# SplineGridsHIPExt -- SplineGrids.jl's grid-evaluation hot path on an AMD GPU
# through libsplinegrids_hip.so (C-ABI: include/splinegrids_hip.h).
#
# Loaded when AMDGPU is (Project.toml: AMDGPU under [weakdeps],
# SplineGridsHIPExt = "AMDGPU" under [extensions], as the reference does for
# its own extensions, Project.toml:16-24).  The methods below are added for
# grids and dimensions whose arrays are ROCArrays, selected by the array type
# parameters of the reference's own types:
#
#   SplineDimension{Tv, Ti, K, S, I, E}          E = typeof(eval)  (src/spline_dimension.jl:19-26)
#   SplineGrid{Nin, Nout, Tv, Ti, S, C, W, E, HasWeights}
#                                                E = typeof(eval)  (src/spline_grid.jl:19-29)
#
# so they are more specific than the reference's methods and never replace
# them: grids on other backends keep the KernelAbstractions kernels.
#
# What replaces what:
#   evaluate!(::SplineDimension)        -> sg_dimension_evaluate   (src/spline_dimension.jl:231-242,
#                                                                    src/utils.jl:19-29)
#   evaluate!(::SplineGrid; ...)        -> sg_evaluate             (src/spline_grid.jl:200-230)
#   evaluate_adjoint!(::SplineGrid; ...) -> sg_evaluate_adjoint    (src/adjoint.jl:52-83)
#   RefinementMatrix * RefinementMatrix  -> sg_refinement_product_rows / _values
#                                                                   (src/refinement_matrix.jl:273-327)
#   mult!(Y, As, B, dims)               -> sg_refinement_mult      (src/refinement_matrix.jl:421-445)
#   mult_adjoint!(B, As, Y, dims)       -> sg_refinement_mult_adjoint (src/adjoint.jl:127-152;
#                                          no atomics; Flag arrays take the structural form)
# for refinement matrices and arrays on ROCArrays; evaluate! and
# evaluate_adjoint! of LocallyRefinedControlPoints then run on these
# (their overwrite / gather kernels are the reference's own KernelAbstractions
# kernels, which run on the AMDGPU backend as they are).
# Validation, size asserts and their messages stay the reference's
# (validate_partial_derivatives, the @asserts), before any ccall.
module SplineGridsHIPExt

using SplineGrids
using SplineGrids: AbstractControlPointArray, obtain, is_nurbs,
                   validate_partial_derivatives, hip_evaluate!, hip_evaluate_adjoint!,
                   spline_eval, RefinementMatrix, Flag, validate_mult_input
using AMDGPU
using Libdl

# ---------------------------------------------------------------------------
# the library
# ---------------------------------------------------------------------------

const SYMBOLS = (:sg_last_error, :sg_dimension_evaluate, :sg_plan_create, :sg_plan_destroy,
    :sg_evaluate, :sg_evaluate_fields, :sg_evaluate_adjoint,
    :sg_rational_adjoint_workspace_bytes, :sg_evaluate_rational_adjoint,
    :sg_rational_workspace_bytes, :sg_evaluate_rational_fields, :sg_batch_create,
    :sg_batch_destroy, :sg_batch_set_arrays, :sg_batch_evaluate, :sg_batch_autotune,
    :sg_refinement_product_rows, :sg_refinement_product_values, :sg_refinement_mult,
    :sg_refinement_mult_adjoint)

# Function pointers, resolved once in __init__ and only read afterwards (so
# finalizers may use them without taking a lock).
const FN = Dict{Symbol, Ptr{Cvoid}}()

function __init__()
    path = get(ENV, "SPLINEGRIDS_HIP_LIB", "libsplinegrids_hip.so")
    handle = Libdl.dlopen(path)
    for s in SYMBOLS
        FN[s] = Libdl.dlsym(handle, s)
    end
    return nothing
end

"""A non-zero status of a C-ABI call other than SG_ERR_INVALID (which is
thrown as an ArgumentError): HIP failures, unsupported inputs, no memory."""
struct SplineGridsHIPError <: Exception
    code::Cint
    msg::String
end
Base.showerror(io::IO, e::SplineGridsHIPError) = print(io, "splinegrids_hip error ", e.code,
    ": ", e.msg)

function check(rc::Cint)
    rc == 0 && return nothing
    msg = unsafe_string(ccall(FN[:sg_last_error], Cstring, ()))
    rc == 1 && throw(ArgumentError(msg))
    throw(SplineGridsHIPError(rc, msg))
end

# sg_dim_desc / sg_grid_desc (same field order, sizes and alignment)
struct SgDimDesc
    degree::Int32
    n_basis::Int32
    n_samples::Int32
    max_derivative_order::Int32
    knots_all::Ptr{Float64}
    sample_points::Ptr{Float64}
end

struct SgGridDesc
    n_in::Int32
    n_out::Int32
    is_nurbs::Int32
    device::Int32
    dims::Ptr{SgDimDesc}
end

dtype_code(::Type{Float32}) = Int32(0)
dtype_code(::Type{Float64}) = Int32(1)
dtype_code(::Type{T}) where {T} = throw(ArgumentError("HIP grids take Float32 or Float64 values, got $T"))

stream_handle() = AMDGPU.stream().stream                 # hipStream_t of the task's stream
device_ordinal() = Int32(AMDGPU.device_id(AMDGPU.device()) - 1)  # HIP ordinals are 0-based
devptr(x::ROCArray) = reinterpret(Ptr{Cvoid}, pointer(x))

"""Arithmetic of the contraction for every call of this extension:
`:native` (FP32 or FP64 FMA), `:split_f16` (hi+lo FP16 on MFMA, FP32 accumulate)
or `:bf16` (BF16 on MFMA); the MFMA tiers take Float32 grids."""
const TIER = Ref{Int32}(0)
function set_tier!(t::Symbol)
    TIER[] = Int32(Dict(:native => 0, :split_f16 => 1, :bf16 => 2)[t])
    return nothing
end

const HIPDimension = SplineDimension{Tv, Ti, K, S, I, <:ROCArray} where {Tv, Ti, K, S, I}
const HIPGrid{Nin, Nout, Tv, HasWeights} =
    SplineGrid{Nin, Nout, Tv, Ti, S, C, W, <:ROCArray, HasWeights} where {Ti, S, C, W}

dense(x::ROCArray) = x
dense(x) = throw(ArgumentError("HIP grids take dense ROCArrays, got $(typeof(x))"))

# ---------------------------------------------------------------------------
# dimensions: span search + Cox-de Boor in FP64 on the GPU
# ---------------------------------------------------------------------------

# Re-evaluations of a dimension, keyed by its basis table (a mutable array:
# the dimension itself is immutable).  Plans remember the count they were
# built with and are rebuilt when it changes (edited sample points or knots).
const DIM_VERSION = WeakKeyDict{Any, Int}()
dim_version(d) = get(DIM_VERSION, d.eval, 0)

function SplineGrids.evaluate!(spline_dimension::HIPDimension)::Nothing
    (; degree, max_derivative_order, knot_vector, sample_points, sample_indices, eval) =
        spline_dimension
    Tv = eltype(eval)
    knots = knot_vector.knots_all
    idx = eltype(sample_indices) == Int32 ? sample_indices : similar(sample_indices, Int32)
    GC.@preserve knots sample_points idx eval begin
        check(ccall(FN[:sg_dimension_evaluate], Cint,
            (Ptr{Cvoid}, Int32, Ptr{Cvoid}, Int32, Int32, Int32, Ptr{Cvoid}, Ptr{Cvoid},
                Int32, Ptr{Cvoid}),
            devptr(knots), length(knots), devptr(sample_points), length(sample_points),
            degree, max_derivative_order, devptr(idx), devptr(eval), dtype_code(Tv),
            stream_handle()))
    end
    idx === sample_indices || copyto!(sample_indices, idx)
    AMDGPU.synchronize()  # the reference synchronizes after every launch
    DIM_VERSION[eval] = dim_version(spline_dimension) + 1
    return nothing
end

# ---------------------------------------------------------------------------
# plans: FP64 tables, tiling and the adjoint workspace of one grid
# ---------------------------------------------------------------------------

mutable struct Plan
    ptr::Ptr{Cvoid}
    key::Vector{Tuple{UInt, Int}}        # (objectid, version) of each dimension's table
    rational_ws::Union{Nothing, ROCVector{UInt8}}  # NURBS adjoint / derivative workspace
end

function destroy!(p::Plan)
    if p.ptr != C_NULL
        ccall(FN[:sg_plan_destroy], Cint, (Ptr{Cvoid},), p.ptr)
        p.ptr = C_NULL
    end
    return nothing
end

# Keyed by the grid's own eval array: a SplineGrid is immutable (no finalizer
# can be attached to it) but its eval field is a mutable ROCArray that lives
# exactly as long as the grid.
const PLANS = WeakKeyDict{Any, Plan}()

function plan(grid::HIPGrid{Nin, Nout})::Plan where {Nin, Nout}
    dims = grid.spline_dimensions
    key = [(objectid(d.eval), dim_version(d)) for d in dims]
    p = get(PLANS, grid.eval, nothing)
    p !== nothing && p.key == key && p.ptr != C_NULL && return p
    p === nothing || destroy!(p)
    knots = [Vector{Float64}(Array(d.knot_vector.knots_all)) for d in dims]
    samples = [Vector{Float64}(Array(d.sample_points)) for d in dims]
    out = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve knots samples begin
        descs = [SgDimDesc(d.degree, length(knots[i]) - d.degree - 1, length(samples[i]),
                     d.max_derivative_order, pointer(knots[i]), pointer(samples[i]))
                 for (i, d) in enumerate(dims)]
        GC.@preserve descs begin
            desc = Ref(SgGridDesc(Nin, Nout, is_nurbs(grid), device_ordinal(), pointer(descs)))
            check(ccall(FN[:sg_plan_create], Cint,
                (Ptr{SgGridDesc}, Ptr{Cvoid}, Ptr{Ptr{Cvoid}}), desc, stream_handle(), out))
        end
    end
    p = Plan(out[], key, nothing)
    finalizer(destroy!, p)
    PLANS[grid.eval] = p
    return p
end

# ---------------------------------------------------------------------------
# evaluate! / evaluate_adjoint! (the reference's signatures and checks)
# ---------------------------------------------------------------------------

function SplineGrids.evaluate!(spline_grid::HIPGrid{Nin, Nout, Tv, HasWeights};
        derivative_order::NTuple{Nin, <:Integer} = ntuple(_ -> 0, Nin),
        control_points::AbstractControlPointArray{Nin, Nout, Tv} = spline_grid.control_points,
        eval::AbstractArray = spline_grid.eval
)::Nothing where {Nin, Nout, Tv, HasWeights}
    validate_partial_derivatives(spline_grid, derivative_order)
    @assert size(control_points) == size(spline_grid.control_points)
    @assert size(eval) == size(spline_grid.eval)
    hip_evaluate!(spline_grid, derivative_order, dense(obtain(control_points)), dense(eval))
    AMDGPU.synchronize()
    return nothing
end

function SplineGrids.evaluate_adjoint!(spline_grid::HIPGrid{Nin, Nout, Tv, false};
        derivative_order::NTuple{Nin, <:Integer} = ntuple(_ -> 0, Nin),
        control_points::AbstractControlPointArray{Nin, Nout, Tv} = spline_grid.control_points,
        eval::AbstractArray = spline_grid.eval
)::Nothing where {Nin, Nout, Tv}
    @assert !is_nurbs(spline_grid) "Adjoint evaluation not supported for NURBS."
    validate_partial_derivatives(spline_grid.spline_dimensions, derivative_order)
    control_points = obtain(control_points)
    @assert size(control_points) == size(spline_grid.control_points)
    @assert size(eval) == size(spline_grid.eval)
    hip_evaluate_adjoint!(spline_grid, derivative_order, dense(eval), dense(control_points))
    AMDGPU.synchronize()
    return nothing
end

# The positional primitives (what the AD rules are attached to).  Stream-
# ordered on the task's stream; no synchronisation of their own.

function SplineGrids.hip_evaluate!(grid::HIPGrid{Nin, Nout, Tv}, derivative_order,
        control_points::ROCArray{Tv}, eval::ROCArray{Tv})::Nothing where {Nin, Nout, Tv}
    order = Int32[derivative_order...]
    W = grid.weights
    GC.@preserve grid control_points eval order W begin
        check(ccall(FN[:sg_evaluate], Cint,
            (Ptr{Cvoid}, Ptr{Int32}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int32, Int32,
                Ptr{Cvoid}),
            plan(grid).ptr, order, devptr(control_points),
            W === nothing ? C_NULL : devptr(dense(W)), devptr(eval), dtype_code(Tv),
            TIER[], stream_handle()))
    end
    return nothing
end

function SplineGrids.hip_evaluate_adjoint!(grid::HIPGrid{Nin, Nout, Tv}, derivative_order,
        eval::ROCArray{Tv}, control_points::ROCArray{Tv};
        accumulate::Bool = false)::Nothing where {Nin, Nout, Tv}
    p = plan(grid)
    if is_nurbs(grid)
        # the adjoint of the rational map, weights fixed (order 0 only)
        all(iszero, derivative_order) ||
            throw(ArgumentError("the NURBS adjoint takes derivative order 0"))
        nbytes = Ref{Int64}(0)
        check(ccall(FN[:sg_rational_adjoint_workspace_bytes], Cint,
            (Ptr{Cvoid}, Int32, Ptr{Int64}), p.ptr, dtype_code(Tv), nbytes))
        if p.rational_ws === nothing || length(p.rational_ws) < nbytes[]
            p.rational_ws = ROCVector{UInt8}(undef, max(nbytes[], 1))
        end
        ws = p.rational_ws
        W = dense(grid.weights)
        GC.@preserve grid eval control_points ws W begin
            check(ccall(FN[:sg_evaluate_rational_adjoint], Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Int32,
                    Int32, Ptr{Cvoid}),
                p.ptr, devptr(W), devptr(eval), devptr(control_points), devptr(ws),
                length(ws), dtype_code(Tv), Int32(accumulate), stream_handle()))
        end
        return nothing
    end
    order = Int32[derivative_order...]
    GC.@preserve grid eval control_points order begin
        check(ccall(FN[:sg_evaluate_adjoint], Cint,
            (Ptr{Cvoid}, Ptr{Int32}, Ptr{Cvoid}, Ptr{Cvoid}, Int32, Int32, Int32, Ptr{Cvoid}),
            p.ptr, order, devptr(eval), devptr(control_points), dtype_code(Tv), TIER[],
            Int32(accumulate), stream_handle()))
    end
    return nothing
end

function SplineGrids.spline_eval(grid::HIPGrid{Nin, Nout, Tv}, control_points::ROCArray{Tv};
        derivative_order::NTuple{Nin, <:Integer} = ntuple(_ -> 0, Nin)) where {Nin, Nout, Tv}
    validate_partial_derivatives(grid, derivative_order)
    @assert size(control_points) == size(grid.control_points)
    eval = similar(grid.eval)
    hip_evaluate!(grid, derivative_order, control_points, eval)
    return eval
end

# ---------------------------------------------------------------------------
# many grids in one persistent launch (no counterpart in the reference)
# ---------------------------------------------------------------------------

"""
    Batch(grids)

Every grid (Float32, on the current AMD GPU) evaluated with derivative order 0
by one persistent launch: `evaluate!(batch)`.  `autotune!(batch)` picks the
tiling on the current arrays; do it before capturing the launch in a graph.
"""
mutable struct Batch
    ptr::Ptr{Cvoid}
    grids::Vector{Any}
end

function destroy!(b::Batch)
    if b.ptr != C_NULL
        ccall(FN[:sg_batch_destroy], Cint, (Ptr{Cvoid},), b.ptr)
        b.ptr = C_NULL
    end
    return nothing
end

function Batch(grids::AbstractVector)
    keep = Any[]
    gdescs = SgGridDesc[]
    for g in grids
        g isa HIPGrid || throw(ArgumentError("a batch takes grids on an AMD GPU"))
        eltype(g.eval) == Float32 || throw(ArgumentError("a batch takes Float32 grids"))
        dims = g.spline_dimensions
        knots = [Vector{Float64}(Array(d.knot_vector.knots_all)) for d in dims]
        samples = [Vector{Float64}(Array(d.sample_points)) for d in dims]
        descs = [SgDimDesc(d.degree, length(knots[i]) - d.degree - 1, length(samples[i]), 0,
                     pointer(knots[i]), pointer(samples[i])) for (i, d) in enumerate(dims)]
        push!(keep, (knots, samples, descs))
        push!(gdescs, SgGridDesc(length(dims), size(obtain(g.control_points))[end],
            is_nurbs(g), device_ordinal(), pointer(descs)))
    end
    out = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve keep gdescs begin
        check(ccall(FN[:sg_batch_create], Cint,
            (Ptr{SgGridDesc}, Int32, Ptr{Cvoid}, Ptr{Ptr{Cvoid}}),
            gdescs, length(gdescs), stream_handle(), out))
    end
    b = Batch(out[], collect(Any, grids))
    finalizer(destroy!, b)
    set_arrays!(b)
    return b
end

function set_arrays!(b::Batch)
    P = Ptr{Cvoid}[devptr(dense(obtain(g.control_points))) for g in b.grids]
    W = Ptr{Cvoid}[g.weights === nothing ? C_NULL : devptr(dense(g.weights)) for g in b.grids]
    E = Ptr{Cvoid}[devptr(g.eval) for g in b.grids]
    check(ccall(FN[:sg_batch_set_arrays], Cint,
        (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}), b.ptr, P, W, E))
    return nothing
end

function SplineGrids.evaluate!(b::Batch)::Nothing
    GC.@preserve b begin
        check(ccall(FN[:sg_batch_evaluate], Cint, (Ptr{Cvoid}, Int32, Ptr{Cvoid}),
            b.ptr, TIER[], stream_handle()))
    end
    AMDGPU.synchronize()
    return nothing
end

function autotune!(b::Batch; reps::Integer = 3)
    check(ccall(FN[:sg_batch_autotune], Cint, (Ptr{Cvoid}, Int32, Ptr{Cvoid}),
        b.ptr, reps, stream_handle()))
    return nothing
end

end # module SplineGridsHIPExt

# ---------------------------------------------------------------------------
# refinement matrices and THB levels (src/refinement_matrix.jl, src/adjoint.jl:85-152)
# ---------------------------------------------------------------------------

# sg_refmat (same field order, sizes and alignment)
struct SgRefmat
    m::Int32
    n::Int32
    nnz::Int32
    row_pointer::Ptr{Int32}
    column_start::Ptr{Int32}
    nzval::Ptr{Cvoid}
end

const HIPRefinementMatrix{Tv} =
    RefinementMatrix{Tv, Int32, <:ROCVector{Int32}, <:ROCVector{Tv}} where {Tv}

sg_refmat(A) = SgRefmat(Int32(A.m), Int32(A.n), Int32(length(A.nzval)),
    reinterpret(Ptr{Int32}, pointer(A.row_pointer)),
    reinterpret(Ptr{Int32}, pointer(A.column_start)), devptr(A.nzval))

function Base.:*(A::HIPRefinementMatrix{Tv}, B::HIPRefinementMatrix{Tv}) where {Tv}
    A.n == B.m || throw(DimensionMismatch("Inner dimensions must match"))
    n_nonzero = ROCVector{Int32}(undef, A.m)
    column_start = ROCVector{Int32}(undef, A.m)
    dA, dB = Ref(sg_refmat(A)), Ref(sg_refmat(B))
    GC.@preserve A B n_nonzero column_start begin
        check(ccall(FN[:sg_refinement_product_rows], Cint,
            (Ptr{SgRefmat}, Ptr{SgRefmat}, Ptr{Int32}, Ptr{Int32}, Ptr{Cvoid}),
            dA, dB, reinterpret(Ptr{Int32}, pointer(n_nonzero)),
            reinterpret(Ptr{Int32}, pointer(column_start)), stream_handle()))
    end
    row_pointer = similar(n_nonzero)
    row_pointer[1:1] .= 1
    A.m > 1 && (row_pointer[2:end] .= 1 .+ cumsum(n_nonzero[1:(end - 1)]))
    nzval = AMDGPU.zeros(Tv, Int(sum(n_nonzero)))
    C = (m = A.m, n = B.n, row_pointer = row_pointer, column_start = column_start,
        nzval = nzval)
    dC = Ref(sg_refmat(C))
    GC.@preserve A B C begin
        check(ccall(FN[:sg_refinement_product_values], Cint,
            (Ptr{SgRefmat}, Ptr{SgRefmat}, Ptr{SgRefmat}, Int32, Ptr{Cvoid}),
            dA, dB, dC, dtype_code(Tv), stream_handle()))
    end
    # the reference's constructor validates the rows (src/refinement_matrix.jl:30-49)
    return RefinementMatrix(A.m, B.n, row_pointer, column_start, nzval)
end

function mult_args(Y, As, B, dims_refinement)
    shape_y = Int32[size(Y)...]
    shape_b = Int32[size(B)...]
    mats = SgRefmat[sg_refmat(A) for A in As]
    dims = Int32[dims_refinement...]
    return shape_y, shape_b, mats, dims
end

function SplineGrids.mult!(Y::ROCArray{Tv}, As::NTuple{N, HIPRefinementMatrix{Tv}},
        B::ROCArray{Tv}, dims_refinement::NTuple{N, <:Integer})::Nothing where {N, Tv}
    validate_mult_input(Y, As, B, dims_refinement)
    shape_y, shape_b, mats, dims = mult_args(Y, As, B, dims_refinement)
    GC.@preserve As Y B shape_y shape_b mats dims begin
        check(ccall(FN[:sg_refinement_mult], Cint,
            (Int32, Ptr{Int32}, Ptr{Int32}, Int32, Ptr{SgRefmat}, Ptr{Int32}, Ptr{Cvoid},
                Ptr{Cvoid}, Int32, Ptr{Cvoid}),
            Int32(ndims(Y)), shape_y, shape_b, Int32(N), mats, dims, devptr(B), devptr(Y),
            dtype_code(Tv), stream_handle()))
    end
    return nothing
end

# B = M^T Y, overwritten as in the reference (src/adjoint.jl:135); Flag arrays
# (deactivate_overwritten_control_points!) take the structural uint8 form
function SplineGrids.mult_adjoint!(B::ROCArray{T}, As::NTuple{N, <:HIPRefinementMatrix},
        Y::ROCArray{T}, dims_refinement::NTuple{N, <:Integer})::Nothing where {N, T}
    validate_mult_input(Y, As, B, dims_refinement)
    code = T === Flag ? Int32(2) : dtype_code(T)
    shape_y, shape_b, mats, dims = mult_args(Y, As, B, dims_refinement)
    GC.@preserve As Y B shape_y shape_b mats dims begin
        check(ccall(FN[:sg_refinement_mult_adjoint], Cint,
            (Int32, Ptr{Int32}, Ptr{Int32}, Int32, Ptr{SgRefmat}, Ptr{Int32}, Ptr{Cvoid},
                Ptr{Cvoid}, Int32, Int32, Ptr{Cvoid}),
            Int32(ndims(Y)), shape_y, shape_b, Int32(N), mats, dims, devptr(Y), devptr(B),
            code, Int32(0), stream_handle()))
    end
    return nothing
end
