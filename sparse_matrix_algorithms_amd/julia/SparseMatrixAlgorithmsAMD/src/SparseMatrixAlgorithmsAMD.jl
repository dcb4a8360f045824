"""
    SparseMatrixAlgorithmsAMD

Julia front end of `libsmalgo.so` (MI355X / gfx950).  Keeps the signatures of
`polfed.jl` and `chebyshev_subspace_iteration.jl`:

    polfed(A::SparseMatrixCSC{Float64,Int}, nev, degree; kwargs...)      # new sparse method
    polfed(A::MatrixFreeRealSymOp, nev, degree; kwargs...)                # polfed.jl:1128
    polfed(matvec!, n, nev, degree; matblock!=nothing, kwargs...)         # polfed.jl:1183
    chebyshev_poly_real_symm_sparse(A::SparseMatrixCSC, emin, emax; ...)  # chebyshev_subspace_iteration.jl:667
    chebyshev_poly_real_symm_sparse(matvec!, n, emin, emax; matblock, ...)# chebyshev_subspace_iteration.jl:771

The start block is drawn exactly as the reference does (`Random.seed!(seed);
randn(n,s)`, polfed.jl:1133) and passed to the library, so the GPU run starts
from the reference's own vectors.  For a symmetric matrix the CSC arrays are
the CSR arrays of A, uploaded with index base 1.
"""
module SparseMatrixAlgorithmsAMD

using LinearAlgebra
using Random
using SparseArrays

export polfed, polfed_slices, chebyshev_poly_real_symm_sparse, MatrixFreeRealSymOp, gershgorin_bounds
export save_csr, load_csr, subspace_rr
export jackson_kernel, cheb_delta_coeffs, cheb_window_coeffs, cheb_smooth_window_coeffs
export apply_poly_block!, POLFEDWorkspace, Chebyshev_workspace

const libsmalgo = get(ENV, "SMALGO_LIB", joinpath(@__DIR__, "..", "..", "..", "libsmalgo.so"))

const WINDOWS = Dict(:step => Cint(0), :delta => Cint(1), :smooth => Cint(2))

# mirrors smalgo_polfed_params (include/smalgo.h)
mutable struct PolfedParams
    s::Cint; has_bounds::Cint; smin::Cdouble; smax::Cdouble; sigma::Cdouble
    emin::Cdouble; emax::Cdouble; tol::Cdouble; overextract::Cint; refine::Cint
    window_type::Cint; jackson::Cint; mmax::Cint; rtol::Cdouble; reorth_passes::Cint
    extra_keep::Cint; seed::UInt64; chunk::Cint; show_progress::Cint; precision::Cint
    refine_guard::Cint; refine_degree::Cint; refine_maxiter::Cint; refine_tol::Cdouble
    refine_guard_degree::Cint
    PolfedParams() = new()
end

# mirrors smalgo_slice (batched slices; no reference counterpart)
struct SmalgoSlice
    emin::Cdouble; emax::Cdouble; sigma::Cdouble; nev::Cint; s::Cint; degree::Cint; mmax::Cint
end

# mirrors smalgo_cheb_params
mutable struct ChebParams
    has_bounds::Cint; smin::Cdouble; smax::Cdouble; sigma::Cdouble; m0::Cint; degree::Cint
    maxiter::Cint; tol::Cdouble; res_gate::Cdouble; jackson::Cint; debug::Cint; two_hit::Cint
    seed::UInt64; window_type::Cint
    ChebParams() = new()
end

struct MatrixFreeRealSymOp{F,G}
    n::Int
    matvec!::F
    matblock!::Union{Nothing,G}
end
MatrixFreeRealSymOp(n::Int, matvec!::F) where {F} = MatrixFreeRealSymOp{F,Nothing}(n, matvec!, nothing)
MatrixFreeRealSymOp(n::Int, matvec!::F, matblock!::G) where {F,G} = MatrixFreeRealSymOp{F,G}(n, matvec!, matblock!)

# ---- context --------------------------------------------------------------
mutable struct Context
    h::Ptr{Cvoid}
end
const _ctx = Ref{Union{Nothing,Context}}(nothing)
function context()
    if _ctx[] === nothing
        h = Ref{Ptr{Cvoid}}(C_NULL)
        rc = ccall((:smalgo_ctx_create, libsmalgo), Cint, (Cint, Ptr{Ptr{Cvoid}}), 0, h)
        rc == 0 || error("smalgo_ctx_create failed with status $rc")
        c = Context(h[])
        finalizer(c -> ccall((:smalgo_ctx_destroy, libsmalgo), Cvoid, (Ptr{Cvoid},), c.h), c)
        _ctx[] = c
    end
    return _ctx[]
end

function check(rc::Cint)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:smalgo_last_error, libsmalgo), Cstring, (Ptr{Cvoid},), context().h))
    rc == -1 ? throw(ArgumentError(msg)) : error(msg)
end

# ---- matrices ---------------------------------------------------------------
mutable struct DeviceMatrix
    h::Ptr{Cvoid}
    n::Int
end
function DeviceMatrix(A::SparseMatrixCSC{Float64,Int})
    n = size(A, 1)
    size(A, 2) == n || throw(ArgumentError("A must be square"))
    h = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve A begin
        check(ccall((:smalgo_csr_upload, libsmalgo), Cint,
                    (Ptr{Cvoid}, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Cdouble}, Cint, Ptr{Ptr{Cvoid}}),
                    context().h, n, A.colptr, A.rowval, A.nzval, 1, h))
    end
    d = DeviceMatrix(h[], n)
    finalizer(d -> ccall((:smalgo_mat_free, libsmalgo), Cvoid, (Ptr{Cvoid},), d.h), d)
    return d
end

function gershgorin_bounds(A::SparseMatrixCSC{Float64,Int})
    d = DeviceMatrix(A)
    lo, hi = Ref(0.0), Ref(0.0)
    check(ccall((:smalgo_gershgorin_bounds, libsmalgo), Cint, (Ptr{Cvoid}, Ref{Cdouble}, Ref{Cdouble}),
                d.h, lo, hi))
    return lo[], hi[]
end

# ---- host callbacks (matrix-free operators) ---------------------------------
# int (*)(void* user, int64 n, int64 b, const double* X, int64 ldx, double* Y, int64 ldy)
function _matblock_tramp(user::Ptr{Cvoid}, n::Int64, b::Int64, X::Ptr{Cdouble}, ldx::Int64,
                         Y::Ptr{Cdouble}, ldy::Int64)::Cint
    # `user` points at a Ref box: the operator itself is immutable, and
    # pointer_from_objref needs a mutable object
    op = (unsafe_pointer_to_objref(user)::Base.RefValue{Any})[]::MatrixFreeRealSymOp
    try
        Xa = view(unsafe_wrap(Matrix{Float64}, X, (ldx, b)), 1:n, :)
        Ya = view(unsafe_wrap(Matrix{Float64}, Y, (ldy, b)), 1:n, :)
        if op.matblock! !== nothing
            op.matblock!(Ya, Xa)
        else
            for j in 1:b      # called on the calling thread only
                op.matvec!(view(Ya, :, j), view(Xa, :, j))
            end
        end
        return Cint(0)
    catch
        return Cint(1)
    end
end
const _tramp = Ref{Ptr{Cvoid}}(C_NULL)
function tramp()
    if _tramp[] == C_NULL
        _tramp[] = @cfunction(_matblock_tramp, Cint,
                              (Ptr{Cvoid}, Int64, Int64, Ptr{Cdouble}, Int64, Ptr{Cdouble}, Int64))
    end
    return _tramp[]
end

# ---- results ----------------------------------------------------------------
function take_result(h::Ptr{Cvoid})
    k = ccall((:smalgo_result_count, libsmalgo), Int64, (Ptr{Cvoid},), h)
    n = ccall((:smalgo_result_n, libsmalgo), Int64, (Ptr{Cvoid},), h)
    λ = zeros(k); res = zeros(k); X = zeros(n, k)
    rc = ccall((:smalgo_result_copy, libsmalgo), Cint,
               (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}, Int64, Ptr{Cdouble}), h, λ, X, n, res)
    ccall((:smalgo_result_free, libsmalgo), Cvoid, (Ptr{Cvoid},), h)
    check(rc)
    return λ, X, res
end

# ---- Chebyshev coefficients (polfed.jl:275-414), host code of the library ----------
argcheck(rc) = rc == 0 ? nothing : throw(ArgumentError("invalid coefficient arguments"))

"`jackson_kernel(K)` (polfed.jl:275-284): K+1 damping factors, `[1.0]` for K <= 0."
function jackson_kernel(K::Int)
    g = zeros(max(K, 0) + 1)
    argcheck(ccall((:smalgo_jackson_kernel, libsmalgo), Cint, (Cint, Ptr{Cdouble}), K, g))
    return g
end

"`cheb_delta_coeffs(σ̂, K; jackson)` (polfed.jl:299-313)."
function cheb_delta_coeffs(σ̂::Float64, K::Int; jackson::Bool=true)
    c = zeros(K + 1)
    argcheck(ccall((:smalgo_cheb_delta_coeffs, libsmalgo), Cint, (Cdouble, Cint, Cint, Ptr{Cdouble}),
                   σ̂, K, jackson, c))
    return c
end

"`cheb_window_coeffs(a, b, K; jackson)` (polfed.jl:329-347)."
function cheb_window_coeffs(a::Float64, b::Float64, K::Int; jackson::Bool=true)
    c = zeros(K + 1)
    argcheck(ccall((:smalgo_cheb_window_coeffs, libsmalgo), Cint,
                   (Cdouble, Cdouble, Cint, Cint, Ptr{Cdouble}), a, b, K, jackson, c))
    return c
end

"`cheb_smooth_window_coeffs(a, b, K; jackson, nsamp, δ)` (polfed.jl:364-388)."
function cheb_smooth_window_coeffs(a::Float64, b::Float64, K::Int; jackson::Bool=true,
                                   nsamp::Int=4096, δ::Float64=0.02)
    c = zeros(K + 1)
    argcheck(ccall((:smalgo_cheb_smooth_window_coeffs, libsmalgo), Cint,
                   (Cdouble, Cdouble, Cint, Cint, Cint, Cdouble, Ptr{Cdouble}),
                   a, b, K, jackson, nsamp, δ, c))
    return c
end

"`_build_coeffs(degree, smin, smax, sigma, emin, emax; window_type, jackson)` (polfed.jl:404-414)."
function _build_coeffs(degree, smin, smax, sigma, emin, emax; window_type::Symbol=:step,
                       jackson::Bool=true)
    haskey(WINDOWS, window_type) || error("Unknown window type: $window_type")
    c = zeros(Int(degree) + 1)
    argcheck(ccall((:smalgo_build_coeffs, libsmalgo), Cint,
                   (Cint, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cint, Cint, Ptr{Cdouble}),
                   degree, smin, smax, sigma, emin, emax, WINDOWS[window_type], jackson, c))
    return c
end

# ---- block filter apply_poly_block! (polfed.jl:464-545) ----------------------------
"""
    POLFEDWorkspace(n, s), Chebyshev_workspace(n, b)

Accepted for signature parity with `apply_poly_block!`; the recurrence's
temporaries live in device memory owned by the library.
"""
struct POLFEDWorkspace
    n::Int
    s::Int
end
const Chebyshev_workspace = POLFEDWorkspace

"""
    apply_poly_block!(Y, A, Q, coeffs, smin, smax, work) -> Y

`Y = p(Â) Q` with Â = (A − cI)/d and p = Σ coeffs[k+1] T_k (polfed.jl:464-545),
on the GPU in FP64.  `A` is a `SparseMatrixCSC` (uploaded on each call; keep a
`DeviceMatrix` to reuse it) or a `MatrixFreeRealSymOp`.
"""
function apply_poly_block!(Y::AbstractMatrix{Float64}, A::Union{SparseMatrixCSC{Float64,Int},DeviceMatrix},
                           Q::AbstractMatrix{Float64}, coeffs::AbstractVector{Float64},
                           smin::Float64, smax::Float64, work=nothing)
    dev = A isa DeviceMatrix ? A : DeviceMatrix(A)
    n, b = size(Q)
    (size(Y) == (n, b) && n == dev.n) || throw(DimensionMismatch("Y, Q and A sizes differ"))
    Qc = Matrix{Float64}(Q); Yc = Matrix{Float64}(undef, n, b); co = Vector{Float64}(coeffs)
    GC.@preserve Qc Yc co begin
        check(ccall((:smalgo_apply_poly_block, libsmalgo), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ptr{Cdouble}, Int64, Ptr{Cdouble}, Cint,
                     Cdouble, Cdouble, Cint, Ptr{Cdouble}, Int64),
                    context().h, dev.h, b, Qc, n, co, length(co) - 1, smin, smax, 0, Yc, n))
    end
    copyto!(Y, Yc)
    return Y
end

function apply_poly_block!(Y::AbstractMatrix{Float64}, A::MatrixFreeRealSymOp,
                           Q::AbstractMatrix{Float64}, coeffs::AbstractVector{Float64},
                           smin::Float64, smax::Float64, work=nothing)
    n, b = size(Q)
    (size(Y) == (n, b) && n == A.n) || throw(DimensionMismatch("Y, Q and A sizes differ"))
    Qc = Matrix{Float64}(Q); Yc = Matrix{Float64}(undef, n, b); co = Vector{Float64}(coeffs)
    box = Ref{Any}(A)
    GC.@preserve box Qc Yc co begin
        check(ccall((:smalgo_apply_poly_block_mf, libsmalgo), Cint,
                    (Ptr{Cvoid}, Int64, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ptr{Cdouble}, Int64,
                     Ptr{Cdouble}, Cint, Cdouble, Cdouble, Ptr{Cdouble}, Int64),
                    context().h, n, tramp(), pointer_from_objref(box), b, Qc, n, co,
                    length(co) - 1, smin, smax, Yc, n))
    end
    copyto!(Y, Yc)
    return Y
end

# ---- Rayleigh–Ritz tail for FEAST / Zolotarev (feast.jl:540-553, zoloterov.jl:1698-1706)
"""
    subspace_rr(A, Y) -> (λ, X, res, Q)

thin_orth(Y) → A·Q → Qᵀ(AQ) → eigen → X = Q·Xq → residuals on the GPU, for all
columns of Y in ascending λ; `Q` is thin_orth(Y).  The shifted solves that form
Y stay on the host.
"""
function subspace_rr(A::Union{SparseMatrixCSC{Float64,Int},DeviceMatrix}, Y::AbstractMatrix{Float64})
    dev = A isa DeviceMatrix ? A : DeviceMatrix(A)
    n, k = size(Y)
    n == dev.n || throw(DimensionMismatch("Y and A sizes differ"))
    Yc = Matrix{Float64}(Y); Qc = zeros(n, k)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve Yc Qc begin
        check(ccall((:smalgo_subspace_rr, libsmalgo), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Int64, Ptr{Cdouble},
                     Int64, Ptr{Cdouble}, Int64, Ptr{Ptr{Cvoid}}),
                    context().h, dev.h, C_NULL, C_NULL, n, k, Yc, n, Qc, n, h))
    end
    λ, X, res = take_result(h[])
    return λ, X, res, Qc
end

# ---- POLFED -------------------------------------------------------------------
function _polfed(dev, op, n::Int, nev::Int, degree::Int; s::Int=8, smin=nothing, smax=nothing,
                 sigma::Float64=0.0, emin::Float64=-Inf, emax::Float64=Inf, tol::Float64=1e-10,
                 threaded::Bool=true, overextract::Int=2, refine::Bool=true,
                 window_type::Symbol=:step, jackson::Bool=true, mmax::Union{Nothing,Int}=nothing,
                 rtol::Float64=1e-12, reorth_passes::Int=1, extra_keep::Int=8, seed::Int=1,
                 show_progress::Bool=true, chunk=64, precision::Symbol=:fp64)
    haskey(WINDOWS, window_type) || error("Unknown window type: $window_type")
    p = PolfedParams()
    ccall((:smalgo_polfed_default_params, libsmalgo), Cvoid, (Ref{PolfedParams},), p)
    p.s = s
    if smin !== nothing && smax !== nothing
        p.has_bounds = 1; p.smin = smin; p.smax = smax
    end
    p.sigma = sigma; p.emin = emin; p.emax = emax; p.tol = tol; p.overextract = overextract
    p.refine = refine; p.window_type = WINDOWS[window_type]; p.jackson = jackson
    p.mmax = mmax === nothing ? 0 : mmax; p.rtol = rtol; p.reorth_passes = reorth_passes
    p.extra_keep = extra_keep; p.seed = UInt64(seed); p.chunk = chunk
    p.show_progress = show_progress; p.precision = precision === :fp64 ? 0 : 1
    Random.seed!(seed)
    Q1 = randn(n, s)                       # polfed.jl:1133-1134
    h = Ref{Ptr{Cvoid}}(C_NULL)
    box = Ref{Any}(op)
    GC.@preserve box Q1 begin
        rc = ccall((:smalgo_polfed, libsmalgo), Cint,
                   (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Cint, Cint,
                    Ref{PolfedParams}, Ptr{Cdouble}, Ptr{Ptr{Cvoid}}),
                   context().h, dev === nothing ? C_NULL : dev.h,
                   op === nothing ? C_NULL : tramp(),
                   op === nothing ? C_NULL : pointer_from_objref(box), n, nev, degree, p, Q1, h)
    end
    check(rc)
    return take_result(h[])
end

polfed(A::SparseMatrixCSC{Float64,Int}, nev::Int, degree::Int; kwargs...) =
    _polfed(DeviceMatrix(A), nothing, size(A, 1), nev, degree; kwargs...)
polfed(A::MatrixFreeRealSymOp, nev::Int, degree::Int; kwargs...) =
    _polfed(nothing, A, A.n, nev, degree; kwargs...)
function polfed(matvec!::F, n::Int, nev::Int, degree::Int; matblock!::Union{Nothing,G}=nothing,
                kwargs...) where {F,G}
    Aop = isnothing(matblock!) ? MatrixFreeRealSymOp(n, matvec!) : MatrixFreeRealSymOp(n, matvec!, matblock!)
    return polfed(Aop, nev, degree; kwargs...)
end

"""
    polfed_slices(A, slices; tol, precision, use_graph, kwargs...)

Many `polfed` intervals of one sparse matrix in one batched device run.
`slices` is a vector of named tuples `(emin, emax, nev, degree)` with optional
`s`, `sigma` and `mmax`; slice `i` uses start-block seed `seed + i - 1`.
Returns a vector of `(λ, X, res)`, one per slice.
"""
function polfed_slices(A::SparseMatrixCSC{Float64,Int}, slices; s::Int=8, smin=nothing,
                       smax=nothing, tol::Float64=1e-10, overextract::Int=2, refine::Bool=true,
                       window_type::Symbol=:step, jackson::Bool=true, rtol::Float64=1e-12,
                       reorth_passes::Int=1, extra_keep::Int=8, seed::Int=1,
                       precision::Symbol=:fp64, use_graph::Bool=true)
    haskey(WINDOWS, window_type) || error("Unknown window type: $window_type")
    dev = DeviceMatrix(A)
    p = PolfedParams()
    ccall((:smalgo_polfed_default_params, libsmalgo), Cvoid, (Ref{PolfedParams},), p)
    p.s = s
    if smin !== nothing && smax !== nothing
        p.has_bounds = 1; p.smin = smin; p.smax = smax
    end
    p.tol = tol; p.overextract = overextract; p.refine = refine
    p.window_type = WINDOWS[window_type]; p.jackson = jackson; p.rtol = rtol
    p.reorth_passes = reorth_passes; p.extra_keep = extra_keep; p.seed = UInt64(seed)
    p.show_progress = 0; p.precision = precision === :fp64 ? 0 : 1
    sl = [SmalgoSlice(d.emin, d.emax, get(d, :sigma, (d.emin + d.emax) / 2), d.nev,
                      get(d, :s, s), d.degree, something(get(d, :mmax, nothing), 0)) for d in slices]
    hs = fill(C_NULL, length(sl))
    rc = ccall((:smalgo_polfed_slices, libsmalgo), Cint,
               (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{SmalgoSlice}, Ref{PolfedParams}, Cint,
                Ptr{Ptr{Cvoid}}),
               context().h, dev.h, length(sl), sl, p, use_graph, hs)
    check(rc)
    return [take_result(h) for h in hs]
end

# ---- on-disk fixtures (the format of include/smalgo.h, shared with Python) ----------
function ioerr(rc)
    rc == 0 && return nothing
    error(unsafe_string(ccall((:smalgo_io_last_error, libsmalgo), Cstring, ())))
end

"Write A (square) as a CSR fixture file; the CSC arrays of transpose(A) are A's CSR arrays."
function save_csr(path::AbstractString, A::SparseMatrixCSC{Float64,Int})
    T = SparseMatrixCSC(transpose(A))
    ioerr(ccall((:smalgo_csr_save, libsmalgo), Cint,
                (Cstring, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Cdouble}, Cint, Cint),
                path, size(A, 1), T.colptr, T.rowval, T.nzval, 1, issymmetric(A)))
end

"Read a CSR fixture file into a SparseMatrixCSC."
function load_csr(path::AbstractString)
    n = Ref{Int64}(0); nnz = Ref{Int64}(0); sym = Ref{Cint}(0)
    ioerr(ccall((:smalgo_csr_info, libsmalgo), Cint, (Cstring, Ref{Int64}, Ref{Int64}, Ref{Cint}),
                path, n, nnz, sym))
    rp = Vector{Int64}(undef, n[] + 1); ci = Vector{Int64}(undef, nnz[]); v = Vector{Float64}(undef, nnz[])
    ioerr(ccall((:smalgo_csr_load, libsmalgo), Cint, (Cstring, Ptr{Int64}, Ptr{Int64}, Ptr{Cdouble}),
                path, rp, ci, v))
    # CSR of A read as CSC is transpose(A); 0-based -> 1-based
    return SparseMatrixCSC(transpose(SparseMatrixCSC(n[], n[], rp .+ 1, ci .+ 1, v)))
end

# ---- Chebyshev subspace iteration -------------------------------------------------
function _cheb(dev, op, n::Int, emin::Float64, emax::Float64; smin=nothing, smax=nothing,
               σ::Float64=0.0, m0::Int=200, degree::Int=80, maxiter::Int=8, tol::Float64=1e-10,
               res_gate::Float64=1e-6, jackson::Bool=true, debug::Bool=false,
               two_hit::Bool=true, seed::Int=1, window_type::Symbol=:step)
    (emin < emax) || throw(ArgumentError("need emin<emax"))
    m0 > 0 || throw(ArgumentError("m0 must be positive"))
    degree >= 0 || throw(ArgumentError("degree must be nonnegative"))
    haskey(WINDOWS, window_type) || error("Unknown window type, possible are :step, :delta, and :smooth")
    p = ChebParams()
    ccall((:smalgo_cheb_default_params, libsmalgo), Cvoid, (Ref{ChebParams},), p)
    if smin !== nothing && smax !== nothing
        p.has_bounds = 1; p.smin = smin; p.smax = smax
    end
    p.sigma = σ; p.m0 = m0; p.degree = degree; p.maxiter = maxiter; p.tol = tol
    p.res_gate = res_gate; p.jackson = jackson; p.debug = debug; p.two_hit = two_hit
    p.seed = UInt64(seed); p.window_type = WINDOWS[window_type]
    Random.seed!(seed)
    Q0 = randn(n, min(m0, n))              # chebyshev_subspace_iteration.jl:701-702
    h = Ref{Ptr{Cvoid}}(C_NULL)
    box = Ref{Any}(op)
    GC.@preserve box Q0 begin
        rc = ccall((:smalgo_chebyshev, libsmalgo), Cint,
                   (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Cdouble, Cdouble,
                    Ref{ChebParams}, Ptr{Cdouble}, Ptr{Ptr{Cvoid}}),
                   context().h, dev === nothing ? C_NULL : dev.h,
                   op === nothing ? C_NULL : tramp(),
                   op === nothing ? C_NULL : pointer_from_objref(box), n, emin, emax, p, Q0, h)
    end
    check(rc)
    return take_result(h[])
end

function chebyshev_poly_real_symm_sparse(Ain::SparseMatrixCSC{Float64,Int}, emin::Float64,
                                         emax::Float64; kwargs...)
    issymmetric(Ain) || throw(ArgumentError("Chebyshev filtered solver requires a real symmetric matrix"))
    return _cheb(DeviceMatrix(Ain), nothing, size(Ain, 1), emin, emax; kwargs...)
end

function chebyshev_poly_real_symm_sparse(matvec!::F, n::Int, emin::Float64, emax::Float64;
                                         matblock::Union{G,Nothing}=nothing, kwargs...) where {F,G}
    op = isnothing(matblock) ? MatrixFreeRealSymOp(n, matvec!) : MatrixFreeRealSymOp(n, matvec!, matblock)
    return _cheb(nothing, op, n, emin, emax; kwargs...)
end

end # module
