"""
    FixedEffectsROCm

The `method = :rocm` backend of FixedEffects.jl (v3.1.0): the solver object
and the methods a backend defines (the same plugin points the CUDA and Metal
extensions use, ext/CUDAExt.jl:99-121, ext/MetalExt.jl:179-189), implemented
by `ccall`s into libfehip.so.  LSMR runs inside the library; Julia only
passes host vectors and reads back residuals, iteration counts and flags.

Load it after FixedEffects:

    using FixedEffects, FixedEffectsROCm
    solve_residuals!(y, fes; method = :rocm, double_precision = true)

The module adds methods to FixedEffects' functions; it does not change any
existing method.
"""
module FixedEffectsROCm

using FixedEffects
using FixedEffects: AbstractFixedEffectSolver, FixedEffect
using StatsBase: AbstractWeights, UnitWeights, uweights

const libfehip = get(ENV, "FEHIP_LIB", joinpath(@__DIR__, "..", "lib", "libfehip.so"))

const FEHIP_FP64 = Int32(0)
const FEHIP_FP32 = Int32(1)
const FEHIP_MIXED = Int32(2)

struct FehipError <: Exception
    code::Int
    msg::String
end
Base.showerror(io::IO, e::FehipError) = print(io, "libfehip error ", e.code, ": ", e.msg)

function check(rc::Integer)
    if rc != 0
        msg = unsafe_string(ccall((:fehip_last_error, libfehip), Cstring, ()))
        throw(FehipError(rc, msg))
    end
    return nothing
end

"""
    FixedEffectSolverROCm{T} <: AbstractFixedEffectSolver{T}

Owns a libfehip handle (refs, CSR bucketings, scales and caches in device
memory).  The handle is freed by a finalizer.
"""
mutable struct FixedEffectSolverROCm{T} <: AbstractFixedEffectSolver{T}
    handle::Ptr{Cvoid}
    fes::Vector{<:FixedEffect}
    weights::AbstractWeights
    N::Int
end

_weights_ptr(w::UnitWeights) = Ptr{Float64}(C_NULL)

function FixedEffects.AbstractFixedEffectSolver{T}(fes::Vector{<:FixedEffect}, weights::AbstractWeights,
                                                   ::Type{Val{:rocm}}; mixed::Bool = false) where {T}
    N = length(weights)
    G = length(fes)
    refs = [convert(Vector{Int64}, fe.refs) for fe in fes]
    nlev = Int64[fe.n for fe in fes]
    inter = [fe.interaction isa UnitWeights ? nothing : convert(Vector{Float64}, fe.interaction) for fe in fes]
    w = weights isa UnitWeights ? nothing : convert(Vector{Float64}, weights)
    precision = mixed ? FEHIP_MIXED : (T === Float64 ? FEHIP_FP64 : FEHIP_FP32)
    out = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve refs inter w begin
        refp = Ptr{Int64}[pointer(r) for r in refs]
        interp = Ptr{Float64}[z === nothing ? Ptr{Float64}(C_NULL) : pointer(z) for z in inter]
        wp = w === nothing ? Ptr{Float64}(C_NULL) : pointer(w)
        check(ccall((:fehip_create, libfehip), Cint,
                    (Ptr{Ptr{Cvoid}}, Int64, Int32, Ptr{Ptr{Int64}}, Ptr{Int64}, Ptr{Ptr{Float64}},
                     Ptr{Float64}, Int32),
                    out, N, G, refp, nlev, interp, wp, precision))
    end
    feM = FixedEffectSolverROCm{T}(out[], fes, weights, N)
    finalizer(feM) do s
        s.handle != C_NULL && ccall((:fehip_destroy, libfehip), Cvoid, (Ptr{Cvoid},), s.handle)
        s.handle = C_NULL
    end
    return feM
end

# update_weights!(feM, w)  (src/SolverCPU.jl:67-76)
function FixedEffects.update_weights!(feM::FixedEffectSolverROCm, weights::AbstractWeights)
    w = weights isa UnitWeights ? nothing : convert(Vector{Float64}, weights)
    GC.@preserve w begin
        check(ccall((:fehip_update_weights, libfehip), Cint, (Ptr{Cvoid}, Ptr{Float64}),
                    feM.handle, w === nothing ? Ptr{Float64}(C_NULL) : pointer(w)))
    end
    feM.weights = weights
    return feM
end

# inner solve_residuals!(r, feM; tol, maxiter)  (src/AbstractFixedEffectSolver.jl:49-68)
function FixedEffects.solve_residuals!(r::AbstractVector{<:Real}, feM::FixedEffectSolverROCm{T};
                                       tol::Real = sqrt(eps(T)), maxiter::Integer = 100_000) where {T}
    _, iters, convs = _solve_many!([r], feM, tol, maxiter)
    return r, iters[1], convs[1]
end

# multi-column solve_residuals!(xs, feM)  (src/AbstractFixedEffectSolver.jl:74-97):
# all columns go to the device in one batched solve instead of a loop.
function FixedEffects.solve_residuals!(xs, feM::FixedEffectSolverROCm{T}; progress_bar = true,
                                       tol::Real = sqrt(eps(T)), maxiter::Integer = 100_000) where {T}
    _, iters, convs = _solve_many!(collect(xs), feM, tol, maxiter)
    return xs, iters, convs
end

function _solve_many!(cols::Vector, feM::FixedEffectSolverROCm, tol, maxiter)
    K = length(cols)
    bufs = [c isa Vector{Float64} ? c : convert(Vector{Float64}, c) for c in cols]
    iters = zeros(Int64, K)
    convs = zeros(UInt8, K)
    GC.@preserve bufs begin
        ptrs = Ptr{Float64}[pointer(b) for b in bufs]
        check(ccall((:fehip_solve_residuals, libfehip), Cint,
                    (Ptr{Cvoid}, Ptr{Ptr{Float64}}, Int32, Float64, Int64, Ptr{Int64}, Ptr{UInt8}),
                    feM.handle, ptrs, K, tol, maxiter, iters, convs))
    end
    for (c, b) in zip(cols, bufs)
        c === b || copyto!(c, b)
    end
    return cols, Int.(iters), Bool.(convs)
end

# inner solve_coefficients!(r, feM; tol, maxiter)  (src/AbstractFixedEffectSolver.jl:147-160)
# The coefficients are collected into eltype(r) before normalize! (:158-159):
# for Float64 callers the solve, normalize! and full all run on the device
# (fehip_solve_coefficients_full); other element types take the raw
# coefficients and the reference's own normalize!/full in that type.
function FixedEffects.solve_coefficients!(r::AbstractVector, feM::FixedEffectSolverROCm{T};
                                          tol::Real = sqrt(eps(T)), maxiter::Integer = 100_000) where {T}
    y = convert(Vector{Float64}, r)
    iters = Ref{Int64}(0)
    conv = Ref{UInt8}(0)
    if eltype(r) === Float64
        obs = [zeros(Float64, length(y)) for _ in feM.fes]
        GC.@preserve obs begin
            optrs = Ptr{Float64}[pointer(o) for o in obs]
            check(ccall((:fehip_solve_coefficients_full, libfehip), Cint,
                        (Ptr{Cvoid}, Ptr{Float64}, Float64, Int64, Ptr{Ptr{Float64}}, Ptr{Ptr{Float64}}, Ptr{Int64},
                         Ptr{UInt8}),
                        feM.handle, y, tol, maxiter, C_NULL, optrs, iters, conv))
        end
        return obs, Int(iters[]), conv[] != 0
    end
    coefs = [zeros(Float64, fe.n) for fe in feM.fes]
    GC.@preserve coefs begin
        ptrs = Ptr{Float64}[pointer(c) for c in coefs]
        check(ccall((:fehip_solve_coefficients, libfehip), Cint,
                    (Ptr{Cvoid}, Ptr{Float64}, Float64, Int64, Ptr{Ptr{Float64}}, Ptr{Int64}, Ptr{UInt8}),
                    feM.handle, y, tol, maxiter, ptrs, iters, conv))
    end
    x = Vector{eltype(r)}[convert(Vector{eltype(r)}, c) for c in coefs]
    return FixedEffects.full(FixedEffects.normalize!(x, feM.fes), feM.fes), Int(iters[]), conv[] != 0
end

# Schedule autotuner (no reference counterpart): times candidate tile
# schedules on this solver's data for K columns and keeps the fastest.
function autotune!(feM::FixedEffectSolverROCm, K::Integer = 1; reps::Integer = 10)
    out = zeros(Float64, 6)
    check(ccall((:fehip_autotune, libfehip), Cint, (Ptr{Cvoid}, Int32, Int32, Ptr{Float64}),
                feM.handle, Int32(K), Int32(reps), out))
    return (tiles_s = Int(out[1]), tiles_c = Int(out[2]), fuse_dense = Int(out[3]),
            ms_per_iter = out[5], default_ms_per_iter = out[6])
end

# Many panels at once (multi-panel form of the multi-column solve_residuals!):
# xss[i] holds the columns of panel i (host vectors), residualized in place;
# the library stages them in device memory and runs all panels' LSMR
# iterations concurrently on the GPU (fehip_solve_residuals_panels_host).
function solve_residuals_panels!(xss::Vector, feMs::Vector{<:FixedEffectSolverROCm};
                                 tol::Real = 1e-8, maxiter::Integer = 10_000)
    length(xss) == length(feMs) || error("one column set per solver")
    P = length(feMs)
    Ks = Int32[length(xs) for xs in xss]
    cols = [[convert(Vector{Float64}, x) for x in xs] for xs in xss]
    iters = [zeros(Int64, k) for k in Ks]
    convs = [zeros(UInt8, k) for k in Ks]
    handles = [feM.handle for feM in feMs]
    GC.@preserve cols iters convs begin
        colptrs = [[pointer(c) for c in cs] for cs in cols]
        GC.@preserve colptrs begin
            cp_ = [pointer(v) for v in colptrs]
            ip = [pointer(a) for a in iters]
            vp = [pointer(a) for a in convs]
            check(ccall((:fehip_solve_residuals_panels_host, libfehip), Cint,
                        (Ptr{Ptr{Cvoid}}, Int32, Ptr{Ptr{Ptr{Float64}}}, Ptr{Int32}, Float64, Int64,
                         Ptr{Ptr{Int64}}, Ptr{Ptr{UInt8}}),
                        handles, Int32(P), cp_, Ks, tol, maxiter, ip, vp))
        end
    end
    for i in 1:P, (c, x) in enumerate(xss[i])
        copyto!(x, cols[i][c])
    end
    return xss, iters, [Bool.(c) for c in convs]
end

# FixedEffect(keys...; interaction) with the grouping on the device
# (src/FixedEffect.jl:18-21): integer key columns only; levels are numbered in
# ascending lexicographic key order.
function fixedeffect(keys::AbstractVector{<:Integer}...; interaction::AbstractVector = uweights(length(keys[1])))
    cols = [convert(Vector{Int64}, k) for k in keys]
    N = length(cols[1])
    all(length(c) == N for c in cols) || throw(DimensionMismatch("all key vectors must have the same length"))
    refs = Vector{Int64}(undef, N)
    ng = Ref{Int64}(0)
    GC.@preserve cols begin
        check(ccall((:fehip_group, libfehip), Cint, (Int64, Int32, Ptr{Ptr{Int64}}, Ptr{Int64}, Ptr{Int64}),
                    N, Int32(length(cols)), Ptr{Int64}[pointer(c) for c in cols], refs, ng))
    end
    return FixedEffect{typeof(refs), typeof(interaction)}(refs, interaction, Int(ng[]))
end

# LSMR algorithm of this solver's solves (no reference counterpart):
# :lsmr (the reference's) or :projected (the FE of most levels eliminated --
# the same residuals and normalized coefficients, fewer iterations).
function set_algorithm!(feM::FixedEffectSolverROCm, algorithm::Symbol)
    code = algorithm === :lsmr ? Int32(0) : algorithm === :projected ? Int32(1) :
           throw(ArgumentError("algorithm must be :lsmr or :projected"))
    check(ccall((:fehip_set_algorithm, libfehip), Cint, (Ptr{Cvoid}, Int32), feM.handle, code))
    return feM
end

end # module
