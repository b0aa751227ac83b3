#
# DiffEqOperatorsMI355X.jl -- the Julia side of the drop-in: `ccall` bindings of libdop.so
# (include/dop.h) and the methods that route DiffEqOperators.jl's application calls on
# device arrays to it.  Constructors, `L*Q`, `L1 + L2` and `compose` stay the package's
# own; only `*` / `mul!` on a device array change.
#
# Not executed in this repository's pipeline (no Julia toolchain in the image); the same
# C ABI is exercised from Python (tests/) with identical descriptors.
#
module DiffEqOperatorsMI355X

using DiffEqOperators, LinearAlgebra
import DiffEqOperators: DerivativeOperator, GhostDerivativeOperator, DiffEqOperatorCombination,
                        AtomicBC, AffineBC, PeriodicBC, MultiDimDirectionalBC, ComposedMultiDimBC,
                        use_winding, diff_axis

const libdop = get(ENV, "DOP_LIBRARY", joinpath(@__DIR__, "..", "lib", "libdop.so"))

# ---- mirrors of the C structs (include/dop.h) -----------------------------------------
struct DopOperator
    axis::Int32; wind::Int32; derivative_order::Int32; approximation_order::Int32
    len::Int32; stencil_length::Int32; boundary_stencil_length::Int32
    boundary_point_count::Int32; offside::Int32; uniform_grid::Int32; coeff_kind::Int32
    reserved::Int32
    stencil_coefs::Ptr{Float64}; low_boundary_coefs::Ptr{Float64}
    high_boundary_coefs::Ptr{Float64}; coefficients::Ptr{Float64}
end

struct DopAffineBC
    present::Int32; n_l::Int32; n_r::Int32; kind::Int32
    a_l::NTuple{8, Float64}; b_l::Float64; a_r::NTuple{8, Float64}; b_r::Float64
end

struct DopDesc
    ndim::Int32; storage::Int32; accum::Int32; input::Int32; semantics::Int32; n_ops::Int32
    size::NTuple{3, Int64}; padded::NTuple{3, Int32}; transverse_view::Int32
    ops::Ptr{DopOperator}; bc::NTuple{3, DopAffineBC}
    line_bc::NTuple{3, Ptr{DopAffineBC}}   # per-line BCs of an axis (C_NULL: bc[a] on all lines)
end

const DOP_INPUT_GHOST, DOP_INPUT_PADDED = 0, 1
const DOP_SEM_LINES, DOP_SEM_BPV = 0, 1
storage_code(::Type{Float64}) = Int32(0)
storage_code(::Type{Float32}) = Int32(1)

function check(rc)
    rc == 0 && return
    msg = unsafe_string(ccall((:dop_last_error, libdop), Cstring, ()))
    rc == -1 && throw(AssertionError(msg))
    rc == -2 && throw(DimensionMismatch(msg))
    error("libdop: $msg ($rc)")
end

# Flatten the operator's own tables (row-major, as dop.h documents) and keep them alive.
function flat_tables(L::DerivativeOperator)
    st = L.stencil_coefs
    flat(v) = collect(Float64, Iterators.flatten(v))
    if use_winding(L) && !(L.dx isa Number)      # [upwind; downwind] matrices of SVectors
        s = vcat(flat(st[1, :]), flat(st[2, :]))
        lo = vcat(flat(L.low_boundary_coefs[1, :]), flat(L.low_boundary_coefs[2, :]))
        hi = vcat(flat(L.high_boundary_coefs[1, :]), flat(L.high_boundary_coefs[2, :]))
    else
        s = eltype(st) <: AbstractVector ? flat(st) : collect(Float64, st)
        lo = flat(L.low_boundary_coefs)
        hi = flat(L.high_boundary_coefs)
    end
    (s, lo, hi, collect(Float64, L.coefficients))
end

coeff_kind(L) = L.coeff_func isa Number ? (L.coeff_func >= 0 ? Int32(1) : Int32(2)) : Int32(0)

function dop_operator(L::DerivativeOperator, keep)
    s, lo, hi, c = flat_tables(L)
    push!(keep, s, lo, hi, c)
    DopOperator(diff_axis(L) - 1, use_winding(L), L.derivative_order, L.approximation_order,
                L.len, L.stencil_length, L.boundary_stencil_length, L.boundary_point_count,
                L.offside, L.dx isa Number, coeff_kind(L), 0,
                pointer(s), pointer(lo), pointer(hi), pointer(c))
end

pad8(v) = ntuple(i -> i <= length(v) ? Float64(v[i]) : 0.0, 8)
const DOP_BC_AFFINE, DOP_BC_PERIODIC, DOP_BC_PERIODIC_ND = 0, 1, 2
dop_bc(q::AffineBC) = DopAffineBC(1, length(q.a_l), length(q.a_r), DOP_BC_AFFINE, pad8(q.a_l), q.b_l,
                                  pad8(q.a_r), q.b_r)
# kind: DOP_BC_PERIODIC for Q*u on a line (bc_operators.jl:192), DOP_BC_PERIODIC_ND for the
# N-D padding of an all-periodic MultiDimDirectionalBC / ComposedMultiDimBC
# (multi_dim_bc_operators.jl:221-252)
dop_bc(q::PeriodicBC, kind) = DopAffineBC(1, 0, 0, kind, pad8(()), 0.0, pad8(()), 0.0)
dop_bc(q::AffineBC, kind) = dop_bc(q)
const NO_BC = DopAffineBC(0, 0, 0, 0, pad8(()), 0.0, pad8(()), 0.0)

# per-axis (atomic BC, kind) of a Q applied to an array of size s
axis_bcs(Q::AtomicBC, s) = Dict(1 => (Q, length(s) == 1 ? DOP_BC_PERIODIC : DOP_BC_PERIODIC_ND))
axis_bcs(Q::MultiDimDirectionalBC{T, B, D}, s) where {T, B, D} =
    Dict(D => (first(Q.BCs), DOP_BC_PERIODIC_ND))
axis_bcs(Q::ComposedMultiDimBC{T, B}, s) where {T, B} =
    Dict(d => (first(Q.BCs[d]), B <: PeriodicBC ? DOP_BC_PERIODIC_ND : DOP_BC_PERIODIC)
         for d in 1:length(Q.BCs))

# per-line BCs: an axis whose MultiDimDirectionalBC holds different atomic BCs on its lines
# (MultiDimBC{dim}(BC::Array), multi_dim_bc_operators.jl:93-95) passes all of them, in the
# array's column-major order, as line_bc (affine BCs only)
line_bcs(Q::MultiDimDirectionalBC{T, B, D}) where {T, B, D} =
    all(b -> b === first(Q.BCs), Q.BCs) ? Dict() : Dict(D => [dop_bc(b) for b in vec(Q.BCs)])
line_bcs(Q::ComposedMultiDimBC) = merge(Dict(), (line_bcs(q) for q in Q.BCs)...)
line_bcs(Q) = Dict()

mutable struct Plan
    handle::Ptr{Cvoid}
    keep::Vector{Any}
end

function Plan(ops::Vector, Q, sz::NTuple{N, Int}, ::Type{T}; input = DOP_INPUT_GHOST,
              padded = (0, 0, 0), transverse_view = 0) where {N, T}
    keep = Any[]
    dops = [dop_operator(L, keep) for L in ops]
    push!(keep, dops)
    bcs = Q === nothing ? Dict() : axis_bcs(Q, sz)
    sem = (N == 1 && Q isa AtomicBC) ? DOP_SEM_BPV : DOP_SEM_LINES
    lines = Q === nothing ? Dict() : line_bcs(Q)
    push!(keep, lines)
    desc = DopDesc(N, storage_code(T), 0, input, sem, length(dops),
                   ntuple(i -> i <= N ? sz[i] : 1, 3), Int32.(padded), transverse_view,
                   pointer(dops), ntuple(i -> haskey(bcs, i) ? dop_bc(bcs[i]...) : NO_BC, 3),
                   ntuple(i -> haskey(lines, i) ? pointer(lines[i]) : Ptr{DopAffineBC}(C_NULL), 3))
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:dop_plan_create, libdop), Cint, (Ref{DopDesc}, Ref{Ptr{Cvoid}}), desc, h))
    p = Plan(h[], keep)
    finalizer(p -> ccall((:dop_plan_destroy, libdop), Cvoid, (Ptr{Cvoid},), p.handle), p)
    p
end

# `du` / `u` are device arrays (e.g. AMDGPU.ROCArray); pointer() yields device addresses and
# the current HIP stream is passed through (C_NULL: default stream).
apply!(p::Plan, du, u; accumulate = false, stream = C_NULL) =
    check(ccall((:dop_apply, libdop), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Cvoid}),
                p.handle, pointer(du), pointer(u), accumulate, stream))

# ---- the methods that route the reference's entry points -------------------------------
# A device array type is supplied by the GPU package in use; `DeviceArray` stands for it.
const DeviceArray = AbstractArray   # narrowed by the loading package (e.g. ROCArray)

const PLAN_CACHE = IdDict{Any, Plan}()

ghost_terms(A::GhostDerivativeOperator) = ([A.L], A.Q)
function ghost_terms(A::DiffEqOperatorCombination)
    all(op -> op isa GhostDerivativeOperator, A.ops) || return nothing
    ([op.L for op in A.ops], A.ops[1].Q)
end

function cached_plan(A, u)
    get!(PLAN_CACHE, (A, size(u), eltype(u))) do
        ops, Q = ghost_terms(A)
        Plan(ops, Q, size(u), eltype(u))
    end
end

# mul!(out, A::GhostDerivativeOperator, u)  (ghost_derivative_operator.jl:15-24)
function gpu_mul!(out, A::Union{GhostDerivativeOperator, DiffEqOperatorCombination}, u)
    apply!(cached_plan(A, u), out, u)
    out
end

# *(A::GhostDerivativeOperator, u)  (ghost_derivative_operator.jl:26-37)
gpu_mul(A, u) = gpu_mul!(similar(u), A, u)

# ---- concretization and A \ u --------------------------------------------------------------
# sparse(A, s) of L*Q (concretization.jl:866-913) as device CSR arrays (0-based columns):
# `alloc(T, n)` allocates a device vector (e.g. `(T, n) -> ROCVector{T}(undef, n)`).
function gpu_sparse(A, s::NTuple{N, Int}, alloc; stream = C_NULL) where {N}
    p = get!(PLAN_CACHE, (A, s, Float64)) do
        ops, Q = ghost_terms(A)
        Plan(ops, Q, s, Float64)
    end
    n = prod(s)
    rowptr = alloc(Int64, n + 1)
    nnz, kl, ku = Ref{Int64}(0), Ref{Int64}(0), Ref{Int64}(0)
    check(ccall((:dop_plan_csr_rows, libdop), Cint,
                (Ptr{Cvoid}, Ptr{Int64}, Ref{Int64}, Ref{Int64}, Ref{Int64}, Ptr{Cvoid}),
                p.handle, pointer(rowptr), nnz, kl, ku, stream))
    col, val, b = alloc(Int64, nnz[]), alloc(Float64, nnz[]), alloc(Float64, n)
    check(ccall((:dop_plan_csr_fill, libdop), Cint,
                (Ptr{Cvoid}, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}, Ptr{Float64}, Ptr{Cvoid}),
                p.handle, pointer(rowptr), pointer(col), pointer(val), pointer(b), stream))
    (rowptr, col, val, b)
end

const SOLVER_CACHE = IdDict{Any, Ptr{Cvoid}}()

# \(A::GhostDerivativeOperator, u)  (ghost_derivative_operator.jl:39-58); u::Float64 device array
function gpu_ldiv(A, u; stream = C_NULL)
    h = get!(SOLVER_CACHE, (A, size(u))) do
        ops, Q = ghost_terms(A)
        p = Plan(ops, Q, size(u), Float64)
        r = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:dop_solver_create, libdop), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ref{Ptr{Cvoid}}),
                    p.handle, stream, r))
        r[]
    end
    x = similar(u)
    check(ccall((:dop_solver_solve, libdop), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cint, Ptr{Cvoid}),
                h, pointer(x), pointer(u), 1, stream))
    x
end

end # module
