# Julia shim for libpeec_hip.so.  `include` this file after the reference's
# own files inside `module Solver`: every definition below replaces the
# reference function of the same name and arity (Julia dispatch picks the
# most recent method), forwarding to the C ABI declared in include/peec_hip.h.
#
# The shim keeps the reference's calling conventions: compute_Lp_Voxels
# returns `nothing` on a stop request, compute_row_P_sup returns NaN, and
# gmres_custom returns flag 99.  Other library errors raise ErrorException.

const LIBPEEC = get(ENV, "PEEC_HIP_LIB", joinpath(@__DIR__, "..", "solver_amd", "libpeec_hip.so"))
const PEEC_OK = Cint(0)
const PEEC_STOPPED = Cint(1)

struct CscF64
    colptr::Ptr{Int64}
    rowval::Ptr{Int64}
    nzval::Ptr{Float64}
    nrow::Int64
    ncol::Int64
end

struct CscC128
    colptr::Ptr{Int64}
    rowval::Ptr{Int64}
    nzval::Ptr{ComplexF64}
    nrow::Int64
    ncol::Int64
end

CscF64(A::SparseMatrixCSC{Float64,Int64}) = CscF64(pointer(A.colptr), pointer(A.rowval), pointer(A.nzval), size(A, 1), size(A, 2))
CscC128(A::SparseMatrixCSC{ComplexF64,Int64}) = CscC128(pointer(A.colptr), pointer(A.rowval), pointer(A.nzval), size(A, 1), size(A, 2))

peec_check(rc, what) = rc == PEEC_OK || rc == PEEC_STOPPED ? rc :
    error(what * ": " * unsafe_string(ccall((:peec_last_error, LIBPEEC), Cstring, ())))

# one context (one HIP stream) per simulation id, with its stop flag mirrored
# from is_stop_requested(id) (solver_start.jl:370-374)
const PEEC_CTX = Dict{Any,Ptr{Cvoid}}()
const PEEC_STOP = Dict{Any,Base.RefValue{Int32}}()
function peec_ctx(id)
    get!(PEEC_CTX, id) do
        h = Ref{Ptr{Cvoid}}(C_NULL)
        peec_check(ccall((:peec_ctx_create, LIBPEEC), Cint, (Cint, Ref{Ptr{Cvoid}}), 0, h), "peec_ctx_create")
        s = get!(PEEC_STOP, id, Ref{Int32}(0))
        ccall((:peec_set_stop_flag, LIBPEEC), Cint, (Ptr{Cvoid}, Ptr{Int32}), h[], Base.unsafe_convert(Ptr{Int32}, s))
        h[]
    end
end
peec_sync_stop!(id) = (PEEC_STOP[id][] = is_stop_requested(id) ? Int32(1) : Int32(0))

# compute_Lp_Voxels.jl:1-63 (hot-path call: M = 1, is_sym = false)
function compute_Lp_Voxels(centri_m, centri_n, sx, sy, sz, sx2, sy2, sz2, dc, is_sym, id)
    is_sym && error("is_sym = true is not on the FFT path")
    ctx = peec_ctx(id); peec_sync_stop!(id)
    c = Matrix{Float64}(centri_n); N = size(c, 1)
    out = Vector{Float64}(undef, N); src = Vector{Float64}(centri_m[1:3])
    rc = GC.@preserve c out src ccall((:peec_lp_row, LIBPEEC), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Int64, Float64, Float64, Float64, Float64, Float64, Float64, Cint, Ptr{Float64}),
        ctx, src, c, N, sx, sy, sz, sx2, sy2, sz2, dc, out)
    peec_check(rc, "peec_lp_row") == PEEC_STOPPED && return nothing
    return reshape(out, 1, N)
end

# compute_row_P_sup.jl:1-104
function compute_row_P_sup(centers, centers2, sx, sy, sz, caser, caser2, id)
    ctx = peec_ctx(id); peec_sync_stop!(id)
    c = Matrix{Float64}(centers2); N = size(c, 1)
    out = Vector{Float64}(undef, N); src = Vector{Float64}(vec(centers)[1:3])
    rc = GC.@preserve c out src ccall((:peec_p_row, LIBPEEC), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Int64, Float64, Float64, Float64, Cint, Cint, Ptr{Float64}),
        ctx, src, c, N, sx, sy, sz, caser, caser2, out)
    peec_check(rc, "peec_p_row") == PEEC_STOPPED && return NaN
    return reshape(out, N, 1)
end

# compute_diagonals.jl:136-155
function Lp_self(l, W, T)
    out = Ref{Float64}(0.0)
    peec_check(ccall((:peec_lp_self, LIBPEEC), Cint, (Float64, Float64, Float64, Ref{Float64}), l, W, T, out), "peec_lp_self")
    out[]
end

# gmres_custom.jl:1-331.  The operator, the frequency data and the LU factors
# are loaded into the context by FFT_solver_QS_S_type's replacement through
# peec_operator_set / peec_frequency_set / peec_lu_set_factors (see
# INTEGRATION.md); the FFT plans and spectra arguments are then unused.
function gmres_custom(b, restarted, tol, maxit, x, wk, incidence_selection, FFTCP, FFTCLp, DZ, Yle, expansions,
                      invZ, invP, lu, PLIVector, PVector, PLI2Vector, P2Vector, chi2Vector, id, chan, portIndex)
    restarted && error("restarted GMRES is dead code in the reference")
    ctx = peec_ctx(id); peec_sync_stop!(id)
    bb = Vector{ComplexF64}(vec(b)); n = length(bb)
    xx = isempty(x) ? zeros(ComplexF64, n) : Vector{ComplexF64}(vec(x))
    flag = Ref{Cint}(0); relres = Ref{Float64}(0.0); it = zeros(Int64, 2)
    resvec = zeros(Float64, min(maxit, n) + 1); nres = Ref{Int64}(0)
    rc = GC.@preserve bb xx it resvec ccall((:peec_gmres, LIBPEEC), Cint,
        (Ptr{Cvoid}, Ptr{ComplexF64}, Float64, Cint, Ptr{ComplexF64}, Ref{Cint}, Ref{Float64}, Ptr{Int64}, Ptr{Float64}, Ref{Int64}),
        ctx, bb, tol, maxit, xx, flag, relres, it, resvec, nres)
    peec_check(rc, "peec_gmres") == PEEC_STOPPED && return xx, 99, 0, it, 0
    return xx, Int(flag[]), relres[], it, resvec[1:nres[]]
end
