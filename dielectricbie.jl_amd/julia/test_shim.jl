# test_shim.jl -- DielectricBIEHip against the reference's own CPU operators.  Needs julia with
# DielectricBIE.jl (and its FMM3D dependency) on the load path, a gfx950 GPU, and DBIE_LIB pointing at
# libdbie.so.  The reference's far field is FMM3D at fmm_tol; libdbie's is the exact direct sum, so
# agreement is stated at a tolerance above fmm_tol.
using Test, LinearAlgebra
using DielectricBIE
include(joinpath(@__DIR__, "DielectricBIEHip.jl"))
const H = DielectricBIEHip

@testset "DielectricBIEHip" begin
    itf = single_dielectric_box3d(1.2, 0.8, 0.6, 4, 0.2, 4.0, 1.0)
    N = num_points(itf)
    op = H.batched_lhs_dielectric_box3d_fmm3d_corrected(itf, 1e-12, 1e-9, 8)
    ref = DielectricBIE.batched_lhs_dielectric_box3d_fmm3d_corrected(itf, 1e-12, 1e-9, 8)
    @test size(op) == (N, N) && eltype(op) == Float64

    x = collect(range(0.1, 1.0, length = N))
    X = hcat(x, 2 .* x .- 1)
    Y = similar(X)
    mul!(Y, op, X)
    Yr = similar(X)
    mul!(Yr, ref, X)
    @test norm(Y - Yr) / norm(Yr) < 1e-9

    src = [0.1 -0.2; 0.1 0.05; 0.1 0.1]
    F = H.rhs_point_charges(op, src, [1.0, -0.5], 4.0)
    @test norm(F[:, 1] - rhs_dielectric_box3d(itf, PointSource((0.1, 0.1, 0.1), 1.0), 4.0)) <= 1e-13 * norm(F[:, 1])
    Σ, st = H.solve_block(op, F; rtol = 1e-12)
    @test st.solved == 1
    R = similar(F)
    mul!(R, ref, Σ)
    @test norm(F - R) / norm(F) < 1e-9

    Σm, stm = H.solve_block(op, F; rtol = 1e-11, precision = :fp16x3)
    @test stm.solved == 1
    @test norm(Σm - Σ) / norm(Σ) < 1e-8

    D = H.laplace3d_D_fmm3d_corrected(itf, 1e-12, 1e-9, 8)
    Dr = DielectricBIE.laplace3d_D_fmm3d_corrected(itf, 1e-12, 1e-9, 8)
    @test norm(D * x - Dr * x) / norm(Dr * x) < 1e-9
end
