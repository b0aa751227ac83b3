#
# DiffEqOperatorsMI355X.jl -- the Julia side of the drop-in: `ccall` bindings of libdop.so
# (include/dop.h) and the methods that route DiffEqOperators.jl's application calls on
# device arrays to it.  Constructors, `L*Q`, `L1 + L2` and `compose` stay the package's
# own; only `*` / `mul!` on a device array change.
#
# Not executed in this repository's pipeline: the image has no Julia toolchain, so this file is
# untested.  The same C ABI, with the same descriptors and the same grouping of terms by BC and
# plan reuse by structure, is exercised from the Python mirror (tests/).
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
# a PeriodicBC on one line wraps it (bc_operators.jl:192)
dop_bc(q::PeriodicBC) = DopAffineBC(1, 0, 0, DOP_BC_PERIODIC, pad8(()), 0.0, pad8(()), 0.0)
# kind: DOP_BC_PERIODIC for Q*u on a line (bc_operators.jl:192), DOP_BC_PERIODIC_ND for the
# N-D padding of an all-periodic MultiDimDirectionalBC / ComposedMultiDimBC
# (multi_dim_bc_operators.jl:221-252)
dop_bc(q::PeriodicBC, kind) = DopAffineBC(1, 0, 0, kind, pad8(()), 0.0, pad8(()), 0.0)
dop_bc(q::AffineBC, kind) = dop_bc(q)
const NO_BC = DopAffineBC(0, 0, 0, 0, pad8(()), 0.0, pad8(()), 0.0)
# the axis entry of an axis whose lines carry their own BCs (line_bc): affine kind, no taps
const LINES_BC = DopAffineBC(1, 0, 0, DOP_BC_AFFINE, pad8(()), 0.0, pad8(()), 0.0)

# per-axis (atomic BC, kind) of a Q applied to an array of size s.  The element type B of the
# BC arrays decides the method: B <: PeriodicBC takes the N-D periodic padding (first and last
# planes, multi_dim_bc_operators.jl:221-228, 244-252), any other B pads line by line
# (slice_rmul, :29-39), where a PeriodicBC line wraps.
axis_bcs(Q::AtomicBC, s) = Dict(1 => (Q, length(s) == 1 ? DOP_BC_PERIODIC : DOP_BC_PERIODIC_ND))
axis_bcs(Q::MultiDimDirectionalBC{T, B, D}, s) where {T, B, D} =
    Dict(D => (first(Q.BCs), B <: PeriodicBC ? DOP_BC_PERIODIC_ND : DOP_BC_PERIODIC))
axis_bcs(Q::ComposedMultiDimBC{T, B}, s) where {T, B} =
    Dict(d => (first(Q.BCs[d]), B <: PeriodicBC ? DOP_BC_PERIODIC_ND : DOP_BC_PERIODIC)
         for d in 1:length(Q.BCs))

# per-line BCs: an axis padded line by line whose lines do not all share one BC
# (MultiDimBC{dim}(BC::Array), multi_dim_bc_operators.jl:93-95) passes every line's atomic BC,
# affine or periodic, in the array's column-major order, as line_bc
uniform_lines(A) = all(b -> b === first(A), A)
line_bcs(Q::MultiDimDirectionalBC{T, B, D}) where {T, B, D} =
    (B <: PeriodicBC || uniform_lines(Q.BCs)) ? Dict() :
    Dict(D => [dop_bc(b) for b in vec(Q.BCs)])
line_bcs(Q::ComposedMultiDimBC{T, B}) where {T, B} =
    B <: PeriodicBC ? Dict() :
    Dict(d => [dop_bc(b) for b in vec(Q.BCs[d])] for d in 1:length(Q.BCs) if !uniform_lines(Q.BCs[d]))
line_bcs(Q) = Dict()

mutable struct Plan
    handle::Ptr{Cvoid}
    keep::Vector{Any}
end

# The descriptor of operators `ops` with the per-axis BCs of Q (the values and the arrays they
# point to stay alive in `keep`).
function descriptor(ops::Vector, Q, sz::NTuple{N, Int}, ::Type{T}; input = DOP_INPUT_GHOST,
                    padded = (0, 0, 0), transverse_view = 0) where {N, T}
    keep = Any[]
    dops = [dop_operator(L, keep) for L in ops]
    push!(keep, dops)
    bcs = Q === nothing ? Dict() : axis_bcs(Q, sz)
    sem = (N == 1 && Q isa AtomicBC) ? DOP_SEM_BPV : DOP_SEM_LINES
    lines = Q === nothing ? Dict() : line_bcs(Q)
    push!(keep, lines)
    axis_entry(i) = haskey(lines, i) ? LINES_BC : haskey(bcs, i) ? dop_bc(bcs[i]...) : NO_BC
    desc = DopDesc(N, storage_code(T), 0, input, sem, length(dops),
                   ntuple(i -> i <= N ? sz[i] : 1, 3), Int32.(padded), transverse_view,
                   pointer(dops), ntuple(axis_entry, 3),
                   ntuple(i -> haskey(lines, i) ? pointer(lines[i]) : Ptr{DopAffineBC}(C_NULL), 3))
    (desc, keep)
end

function Plan(ops::Vector, Q, sz::NTuple{N, Int}, ::Type{T}; kw...) where {N, T}
    desc, keep = descriptor(ops, Q, sz, T; kw...)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve keep check(ccall((:dop_plan_create, libdop), Cint, (Ref{DopDesc}, Ref{Ptr{Cvoid}}),
                                  desc, h))
    p = Plan(h[], keep)
    finalizer(p -> ccall((:dop_plan_destroy, libdop), Cvoid, (Ptr{Cvoid},), p.handle), p)
    p
end

# New values of the same structure, in place (dop_plan_update); false if the structure differs.
function update!(p::Plan, ops::Vector, Q, sz::NTuple{N, Int}, ::Type{T}; stream = current_stream(),
                 kw...) where {N, T}
    desc, keep = descriptor(ops, Q, sz, T; kw...)
    rc = GC.@preserve keep ccall((:dop_plan_update, libdop), Cint,
                                 (Ptr{Cvoid}, Ref{DopDesc}, Ptr{Cvoid}), p.handle, desc, stream)
    rc == -1 && return false
    check(rc)
    p.keep = keep
    true
end

# The HIP stream of the calling task: AMDGPU.jl's task-local stream when AMDGPU is loaded,
# else the default stream (C_NULL), so the library's work is ordered with the caller's.
current_stream() =
    isdefined(Main, :AMDGPU) ? Base.invokelatest(() -> Main.AMDGPU.stream().stream) : C_NULL

# `du` / `u` are device arrays (e.g. AMDGPU.ROCArray); pointer() yields device addresses.
apply!(p::Plan, du, u; accumulate = false, stream = current_stream()) =
    check(ccall((:dop_apply, libdop), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Cvoid}),
                p.handle, pointer(du), pointer(u), accumulate, stream))

# ---- the methods that route the reference's entry points -------------------------------
# A device array type is supplied by the GPU package in use; `DeviceArray` stands for it.
const DeviceArray = AbstractArray   # narrowed by the loading package (e.g. ROCArray)

# ---- plans of L*Q and sums of them ---------------------------------------------------------
# (L, Q) terms of A, split into groups whose BCs agree on every axis they pad: one plan per
# group, the groups' results summed (composite_operators.jl:64-65 sums the terms' results).
ghost_terms(A::GhostDerivativeOperator) = [(A.L, A.Q)]
function ghost_terms(A::DiffEqOperatorCombination)
    all(op -> op isa GhostDerivativeOperator, A.ops) || return nothing
    [(op.L, op.Q) for op in A.ops]
end

function ghost_groups(terms, sz)
    groups = Vector{Tuple{Vector{Any}, Any}}()     # (operators, Q)
    for (L, Q) in terms
        ax = diff_axis(L)
        g = findfirst(gr -> gr[2] === Q || axis_bcs(gr[2], sz) == axis_bcs(Q, sz) &&
                                           line_bcs(gr[2]) == line_bcs(Q), groups)
        g === nothing ? push!(groups, (Any[L], Q)) : push!(groups[g][1], L)
    end
    groups
end

# Plans are keyed by structure -- the operators' kinds and sizes and the BCs' kinds and tap
# counts -- not by the operator object, so a method-of-lines RHS that builds a new BC at every
# call (heat_equation.jl:38-43, KdV.jl:26-30) reuses one plan and only replaces its values
# (dop_plan_update).  At most PLAN_CACHE_MAX structures are kept (least recently used dropped).
const PLAN_CACHE = Dict{Any, Tuple{Plan, Any}}()   # structure => (plan, values bound to it)
const PLAN_ORDER = Any[]
const PLAN_CACHE_MAX = 64

op_structure(L) = (diff_axis(L), use_winding(L), L.derivative_order, L.approximation_order,
                   L.len, L.stencil_length, L.boundary_stencil_length, L.boundary_point_count,
                   L.offside, L.dx isa Number, coeff_kind(L))
bc_structure(q::PeriodicBC) = (:periodic,)
bc_structure(q::AffineBC) = (:affine, length(q.a_l), length(q.a_r))
q_structure(Q, sz) = (sort!([(d, bc_structure(b), k) for (d, (b, k)) in axis_bcs(Q, sz)]),
                      sort!([(d, map(bc_structure, v)) for (d, v) in line_bcs_atoms(Q)]))
line_bcs_atoms(Q::MultiDimDirectionalBC{T, B, D}) where {T, B, D} =
    (B <: PeriodicBC || uniform_lines(Q.BCs)) ? Dict() : Dict(D => vec(Q.BCs))
line_bcs_atoms(Q::ComposedMultiDimBC{T, B}) where {T, B} =
    B <: PeriodicBC ? Dict() :
    Dict(d => vec(Q.BCs[d]) for d in 1:length(Q.BCs) if !uniform_lines(Q.BCs[d]))
line_bcs_atoms(Q) = Dict()

function bound_plan(ops, Q, sz, T)
    key = (map(op_structure, ops), q_structure(Q, sz), sz, T)
    vals = (ops, Q)
    hit = get(PLAN_CACHE, key, nothing)
    if hit !== nothing
        p, bound = hit
        filter!(k -> k != key, PLAN_ORDER)
        push!(PLAN_ORDER, key)
        bound === vals && return p
        if update!(p, collect(ops), Q, sz, T)
            PLAN_CACHE[key] = (p, vals)
            return p
        end
    end
    p = Plan(collect(ops), Q, sz, T)
    PLAN_CACHE[key] = (p, vals)
    filter!(k -> k != key, PLAN_ORDER)
    push!(PLAN_ORDER, key)
    while length(PLAN_ORDER) > PLAN_CACHE_MAX
        delete!(PLAN_CACHE, popfirst!(PLAN_ORDER))
    end
    p
end

# mul!(out, A::GhostDerivativeOperator, u)  (ghost_derivative_operator.jl:15-24) and the sum of
# such terms (composite_operators.jl:76-84): the first group overwrites, the others accumulate.
function gpu_mul!(out, A::Union{GhostDerivativeOperator, DiffEqOperatorCombination}, u;
                  stream = current_stream())
    terms = ghost_terms(A)
    terms === nothing && throw(ArgumentError("gpu_mul! needs L*Q or a sum of L*Q terms"))
    for (i, (ops, Q)) in enumerate(ghost_groups(terms, size(u)))
        apply!(bound_plan(ops, Q, size(u), eltype(u)), out, u; accumulate = i > 1, stream = stream)
    end
    out
end

# *(A::GhostDerivativeOperator, u)  (ghost_derivative_operator.jl:26-37)
gpu_mul(A, u) = gpu_mul!(similar(u), A, u)

# ---- concretization and A \ u --------------------------------------------------------------
# sparse(A, s) of L*Q (concretization.jl:866-913) as device CSR arrays (0-based columns):
# `alloc(T, n)` allocates a device vector (e.g. `(T, n) -> ROCVector{T}(undef, n)`).
function gpu_sparse(A, s::NTuple{N, Int}, alloc; stream = current_stream()) where {N}
    groups = ghost_groups(ghost_terms(A), s)
    length(groups) == 1 ||
        throw(ArgumentError("gpu_sparse needs the terms of A to share one BC per axis"))
    p = bound_plan(groups[1][1], groups[1][2], s, Float64)
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

# Factorizations keyed by the operator object (a factorization depends on every value of A),
# at most SOLVER_CACHE_MAX of them: the oldest is destroyed when a new one is made.
mutable struct Solver
    handle::Ptr{Cvoid}
end
const SOLVER_CACHE = IdDict{Any, Solver}()
const SOLVER_ORDER = Any[]
const SOLVER_CACHE_MAX = 16

# \(A::GhostDerivativeOperator, u)  (ghost_derivative_operator.jl:39-58); u::Float64 device array
function gpu_ldiv(A, u; stream = current_stream())
    key = (A, size(u))
    sol = get(SOLVER_CACHE, key, nothing)
    if sol === nothing
        groups = ghost_groups(ghost_terms(A), size(u))
        length(groups) == 1 ||
            throw(ArgumentError("A \\ u needs the terms of A to share one BC per axis"))
        p = Plan(collect(groups[1][1]), groups[1][2], size(u), Float64)
        r = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:dop_solver_create, libdop), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ref{Ptr{Cvoid}}),
                    p.handle, stream, r))
        sol = Solver(r[])
        finalizer(x -> ccall((:dop_solver_destroy, libdop), Cvoid, (Ptr{Cvoid},), x.handle), sol)
        SOLVER_CACHE[key] = sol
        push!(SOLVER_ORDER, key)
        while length(SOLVER_ORDER) > SOLVER_CACHE_MAX
            delete!(SOLVER_CACHE, popfirst!(SOLVER_ORDER))
        end
    end
    h = sol.handle
    x = similar(u)
    check(ccall((:dop_solver_solve, libdop), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cint, Ptr{Cvoid}),
                h, pointer(x), pointer(u), 1, stream))
    x
end

end # module
