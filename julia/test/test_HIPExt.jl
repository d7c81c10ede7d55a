# The reference's hot-path tests on an AMD GPU through SplineGridsHIPExt.
# Run with SPLINEGRIDS_HIP_LIB pointing at libsplinegrids_hip.so, after the
# extension files and hip_hooks.jl are in place (see the integration notes).
# Mirrors test/test_spline_grid.jl, test/test_nurbs_grid.jl,
# test/test_EnzymeExt.jl and test/test_LinearMapsExt.jl:14-32.
using SplineGrids
using AMDGPU
using Enzyme
using Random
using Test

const HIPExt = Base.get_extension(SplineGrids, :SplineGridsHIPExt)
backend = AMDGPU.ROCBackend()

@testset "HIP: the extension is the one evaluating" begin
    dims = SplineDimension.((5, 6), (3, 2), (7, 9); backend)
    grid = SplineGrid(dims, 2)
    @test grid isa HIPExt.HIPGrid
    @test which(evaluate!, (typeof(grid),)).module === HIPExt
end

@testset "HIP: G1 golden (test/test_spline_grid.jl:29-53) against the CPU path" begin
    Random.seed!(1)
    for Tv in (Float32, Float64)
        cpu = SplineGrid(SplineDimension.((5, 6), (3, 2), (7, 9); float_type = Tv), 2)
        gpu = SplineGrid(SplineDimension.((5, 6), (3, 2), (7, 9); float_type = Tv, backend), 2)
        P = rand(Tv, 5, 6, 2)
        copyto!(cpu.control_points, P)
        copyto!(gpu.control_points, ROCArray(P))
        evaluate!(cpu)
        evaluate!(gpu)
        @test Array(gpu.eval) ≈ cpu.eval rtol = (Tv == Float32 ? 1e-6 : 1e-12)
        for d in 1:2
            @test Array(gpu.spline_dimensions[d].sample_indices) == cpu.spline_dimensions[d].sample_indices
        end
    end
end

@testset "HIP: derivative validation messages unchanged" begin
    dims = SplineDimension.((5, 6), (3, 2), (7, 9); max_derivative_order = 1, backend)
    grid = SplineGrid(dims, 2)
    @test_throws ErrorException evaluate!(grid; derivative_order = (2, 1))
    nurbs = NURBSGrid(SplineDimension(4, 2, 100; max_derivative_order = 1, backend), 1)
    @test_throws ErrorException evaluate!(nurbs; derivative_order = (1,))
end

@testset "HIP: adjoint is the transpose (test/test_LinearMapsExt.jl:14-32)" begin
    dims = SplineDimension.((10, 9, 8), (3, 2, 2), (40, 33, 21); float_type = Float64, backend)
    grid = SplineGrid(dims, 2)
    x = ROCArray(randn(10, 9, 8, 2))
    y = ROCArray(randn(40, 33, 21, 2))
    Mx = similar(grid.eval)
    MTy = similar(x)
    evaluate!(grid; control_points = x, eval = Mx)
    evaluate_adjoint!(grid; eval = y, control_points = MTy)
    @test sum(Array(Mx) .* Array(y)) ≈ sum(Array(x) .* Array(MTy)) rtol = 1e-12
end

@testset "HIP: Enzyme reverse mode (test/test_EnzymeExt.jl:24-58)" begin
    Random.seed!(1)
    dims = SplineDimension.((10, 10), (2, 2), (50, 50); backend)
    grid = SplineGrid(dims, 2)
    function loss(control_points_flat, spline_grid)
        evaluate!(spline_grid;
            control_points = reshape(control_points_flat, size(spline_grid.control_points)))
        return sum(spline_grid.eval)
    end
    flat = ROCArray(rand(Float32, length(grid.control_points)))
    dflat = Duplicated(flat, make_zero(flat))
    dgrid = Duplicated(grid, make_zero(grid))
    autodiff(Reverse, loss, Active, dflat, dgrid)
    # d/dP sum(M P) = Mᵀ 1
    ones_eval = ROCArray(ones(Float32, size(grid.eval)))
    want = similar(grid.control_points.control_points)
    evaluate_adjoint!(grid; eval = ones_eval, control_points = want)
    @test Array(dflat.dval) ≈ vec(Array(want)) rtol = 1e-5
    make_zero!(dgrid.dval)
    @test all(iszero, Array(dgrid.dval.eval))
end
