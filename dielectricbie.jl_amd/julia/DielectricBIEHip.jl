# DielectricBIEHip.jl -- the DielectricBIE.jl side of libdbie (include/dbie.h): the reference's
# operator protocol and solver entry points reach the gfx950 library through ccall.
#
# A maintainer adds this file next to src/solver/multi_rhs.jl of DielectricBIE.jl and sets
# DBIE_LIB to the built dielectricbie.jl_amd/dbie/libdbie.so.  Struct layouts mirror the C structs
# of include/dbie.h field for field (checked against the header by tests/test_boundary.py).
#
#   op = DielectricBIEHip.batched_lhs_dielectric_box3d_fmm3d_corrected(itf, 1e-9, 1e-9, 8)
#   Krylov.block_gmres(op, F)                         # the reference's operator protocol, unchanged
#   Σ, stats = DielectricBIEHip.solve_block(op, F)    # GMRES on the device (no host round trips)
module DielectricBIEHip

using LinearAlgebra
import DielectricBIE: DielectricInterface, FlatPanel, num_points

const lib = get(ENV, "DBIE_LIB", joinpath(@__DIR__, "..", "dbie", "libdbie.so"))

# ---- option and statistics structs (include/dbie.h) ------------------------------------------
struct NearOpts
    up_tol::Float64
    max_order::Cint
    correct_edges::Cint
    adaptive_atol::Float64
    adaptive_rtol::Float64
    adaptive_n_GL::Cint
    adaptive_max_depth::Cint
    build_corrections::Cint
    n_threads::Cint
    device_build::Cint
end

struct SolveOpts
    rtol::Float64
    atol::Float64
    itmax::Cint
    restart::Cint
    precondition::Cint
    precision::Cint
    inner_rtol::Float64
    max_refine::Cint
    check_residual::Cint
    krylov::Cint
    recycle::Cint
    verify::Cint
end

mutable struct SolveStats
    niter::Cint
    nrefine::Cint
    solved::Cint
    max_rel_residual::Float64
    time_ms::Float64
end

const DBIE_FP64 = Cint(0)
const DBIE_FP16X3 = Cint(4)
const DBIE_OP_DT = Cint(0)
const DBIE_OP_D = Cint(2)
const PRECISIONS = Dict(:fp64 => Cint(0), :fp32 => Cint(1), :bf16 => Cint(2), :bf16x3 => Cint(3), :fp16x3 => Cint(4))

# status codes -> the reference's exception types
function check(st::Cint)
    st == 0 && return nothing
    buf = Vector{UInt8}(undef, 1024)
    ccall((:dbie_last_error, lib), Cint, (Ptr{UInt8}, Csize_t), buf, length(buf))
    msg = unsafe_string(pointer(buf))
    st == 2 && throw(DimensionMismatch(msg))
    st == 1 && throw(ArgumentError(msg))
    error("libdbie status $st: $msg")
end

# ---- geometry: the reference's own panels and quadrature, as given ---------------------------
# DielectricInterface(panels, eps_in, eps_out) (src/core/panels.jl:5-22,64-68): corners (lb, rb, rt, lt),
# normals, edge flags, and the panels' points, weights, gl_xs, gl_ws and bary_weights.
function hip_geometry(itf::DielectricInterface{FlatPanel{Float64, 3}, Float64})
    panels = itf.panels
    P = length(panels)
    nq = panels[1].n_quad
    all(p -> p.n_quad == nq, panels) || throw(ArgumentError("libdbie takes one n_quad for all panels"))
    corners = Float64[c[k] for p in panels for c in p.corners for k in 1:3]
    normals = Float64[p.normal[k] for p in panels for k in 1:3]
    is_edge = Cint[p.is_edge for p in panels]
    points = Float64[x[k] for p in panels for x in p.points for k in 1:3]
    weights = Float64[w for p in panels for w in p.weights]
    p1 = panels[1]
    g = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:dbie_geometry_from_panel_nodes, lib), Cint,
                (Int64, Cint, Ptr{Float64}, Ptr{Float64}, Ptr{Cint}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                 Ptr{Float64}, Ptr{Float64}, Ptr{Ptr{Cvoid}}),
                P, nq, corners, normals, is_edge, points, weights, p1.gl_xs, p1.gl_ws, p1.bary_weights, g))
    return g[]
end

# ---- the operator: drop-in for BatchedDielectricOperator (src/solver/multi_rhs.jl:46-141) ---------
mutable struct HipDielectricOperator
    ptr::Ptr{Cvoid}
    n::Int
end

function batched_lhs_dielectric_box3d_fmm3d_corrected(itf, fmm_tol, up_tol, max_order;
        correct_edges = true, adaptive_atol = up_tol, adaptive_rtol = sqrt(eps(Float64)),
        adaptive_n_GL = 0, adaptive_max_depth = 20)
    g = hip_geometry(itf)
    o = Ref(NearOpts(up_tol, max_order, correct_edges, adaptive_atol, adaptive_rtol,
                     adaptive_n_GL, adaptive_max_depth, 1, 0, 1))  # corrections built on the GPU
    op = Ref{Ptr{Cvoid}}(C_NULL)
    st = ccall((:dbie_operator_create, lib), Cint,
               (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ref{NearOpts}, Ptr{Ptr{Cvoid}}),
               g, Vector{Float64}(itf.eps_in), Vector{Float64}(itf.eps_out), o, op)
    ccall((:dbie_geom_destroy, lib), Cvoid, (Ptr{Cvoid},), g)
    check(st)
    h = HipDielectricOperator(op[], num_points(itf))
    finalizer(x -> ccall((:dbie_operator_destroy, lib), Cvoid, (Ptr{Cvoid},), x.ptr), h)
    return h
end

Base.size(op::HipDielectricOperator) = (op.n, op.n)
Base.size(op::HipDielectricOperator, d::Integer) = d <= 2 ? op.n : 1
Base.eltype(::HipDielectricOperator) = Float64

function LinearAlgebra.mul!(Y::StridedMatrix{Float64}, op::HipDielectricOperator, X::StridedMatrix{Float64})
    size(X, 1) == op.n || throw(DimensionMismatch("X has $(size(X, 1)) rows, operator has $(op.n)"))
    size(Y) == size(X) || throw(DimensionMismatch("Y and X differ in size"))
    check(ccall((:dbie_apply, lib), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Cint),
                op.ptr, size(X, 2), X, stride(X, 2), Y, stride(Y, 2), DBIE_FP64))
    return Y
end
LinearAlgebra.mul!(y::StridedVector{Float64}, op::HipDielectricOperator, x::StridedVector{Float64}) =
    (mul!(reshape(y, :, 1), op, reshape(x, :, 1)); y)
Base.:*(op::HipDielectricOperator, x::AbstractVecOrMat{Float64}) = mul!(similar(x), op, x)

# ---- layer potentials with near corrections (src/kernel/laplace3d_near.jl:395-445) ---------------
struct HipKernelMap
    op::HipDielectricOperator
    kernel::Cint
end
Base.size(A::HipKernelMap) = size(A.op)
function Base.:*(A::HipKernelMap, X::AbstractVecOrMat{Float64})
    Xm = Matrix{Float64}(reshape(X, size(X, 1), :))
    Y = similar(Xm)
    check(ccall((:dbie_apply_kernel_corrected, lib), Cint,
                (Ptr{Cvoid}, Cint, Cint, Ptr{Float64}, Int64, Ptr{Float64}, Int64),
                A.op.ptr, A.kernel, size(Xm, 2), Xm, stride(Xm, 2), Y, stride(Y, 2)))
    return X isa AbstractVector ? vec(Y) : Y
end
laplace3d_DT_fmm3d_corrected(itf, fmm_tol, up_tol, max_order; kw...) =
    HipKernelMap(batched_lhs_dielectric_box3d_fmm3d_corrected(itf, fmm_tol, up_tol, max_order; kw...), DBIE_OP_DT)
laplace3d_D_fmm3d_corrected(itf, fmm_tol, up_tol, max_order; kw...) =
    HipKernelMap(batched_lhs_dielectric_box3d_fmm3d_corrected(itf, fmm_tol, up_tol, max_order; kw...), DBIE_OP_D)

# ---- right-hand side and solves ---------------------------------------------------------------
# one column per point charge (rhs_dielectric_box3d(itf, PointSource, eps_src),
# src/solver/dielectric_box3d.jl:95-104); src: 3 x K positions, q: K charges
function rhs_point_charges(op::HipDielectricOperator, src::Matrix{Float64}, q::Vector{Float64}, eps_src::Float64)
    size(src, 1) == 3 && size(src, 2) == length(q) || throw(DimensionMismatch("src must be 3 x K"))
    F = Matrix{Float64}(undef, op.n, length(q))
    check(ccall((:dbie_rhs_points, lib), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Float64}, Ptr{Float64}, Float64, Ptr{Float64}, Int64),
                op.ptr, length(q), src, q, eps_src, F, op.n))
    return F
end

# solve_dielectric_box3d_block on the device (src/solver/multi_rhs.jl:270-286)
function solve_block(op::HipDielectricOperator, F::Matrix{Float64}; rtol = 1e-10, atol = 0.0,
                     itmax = 500,
                     precondition = true,  # true / 1 diagonal scaling, 2 near field, 3 near field in two sweeps, 0 none
                     precision = :fp64, inner_rtol = 1e-5, max_refine = 10,
                     krylov = 1,   # 1 = block GMRES as Krylov.block_gmres, 0 = per column
                     recycle = 1,  # mixed precision, per column: later refinements deflate the first space
                     verify = 1)   # mixed precision: 1 = end on an FP64 residual, 0 = on the Krylov estimate
    size(F, 1) == op.n || throw(DimensionMismatch("F has $(size(F, 1)) rows, operator has $(op.n)"))
    Σ = similar(F)
    o = Ref(SolveOpts(rtol, atol, itmax, 0, precondition, PRECISIONS[precision], inner_rtol, max_refine, 1,
                      krylov, recycle, verify))
    s = Ref(SolveStats(0, 0, 0, 0.0, 0.0))
    st = ccall((:dbie_solve_block, lib), Cint,
               (Ptr{Cvoid}, Cint, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ref{SolveOpts}, Ref{SolveStats}),
               op.ptr, size(F, 2), F, stride(F, 2), Σ, stride(Σ, 2), o, s)
    st == 5 || check(st)   # 5 = not converged: stats.solved == 0, as Krylov.jl reports it
    return Σ, s[]
end

# solve_lu (src/utils/linear_algebra.jl:1) on the GPU
function solve_lu(A::Matrix{Float64}, b::AbstractVecOrMat{Float64})
    n = size(A, 1)
    size(A, 2) == n && size(b, 1) == n || throw(DimensionMismatch("A must be n x n and b n x k"))
    B = Matrix{Float64}(reshape(b, n, :))
    check(ccall((:dbie_solve_lu, lib), Cint,
                (Int64, Ptr{Float64}, Int64, Cint, Ptr{Float64}, Int64),
                n, A, stride(A, 2), size(B, 2), B, stride(B, 2)))
    return b isa AbstractVector ? vec(B) : B
end

end # module
