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
# peec_kernels_build / peec_operator_set once per geometry and
# peec_frequency_set / peec_lu_set_factors per frequency; the FFT plans and
# spectra arguments are then unused.
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

# ---------------------------------------------------------------------------
# Batched sweep and adjoint (no counterpart in the reference).  Frequency
# slots hold per-frequency DZ, Yle and device LU factors; right-hand side r of
# a batched call runs against slot r ÷ rhs_per_slot (0-based).

peec_frequency_slots(ctx, nslots) =
    peec_check(ccall((:peec_frequency_slots, LIBPEEC), Cint, (Ptr{Cvoid}, Cint), ctx, nslots), "peec_frequency_slots")

function peec_frequency_set_slot(ctx, slot, w, DZ::Vector{ComplexF64}, Yle::SparseMatrixCSC{ComplexF64,Int64})
    Y = CscC128(Yle)
    rc = GC.@preserve DZ Yle ccall((:peec_frequency_set_slot, LIBPEEC), Cint,
        (Ptr{Cvoid}, Cint, Float64, Ptr{ComplexF64}, Ref{CscC128}), ctx, slot, w, DZ, Y)
    peec_check(rc, "peec_frequency_set_slot")
    peec_check(ccall((:peec_lu_factor_slot, LIBPEEC), Cint, (Ptr{Cvoid}, Cint), ctx, slot), "peec_lu_factor_slot")
end

# tn = M^-1 [0; 0; is] (adjoint = false) or M^-H [0; 0; g] (adjoint = true)
function peec_precond_rhs_batched(ctx, IS::AbstractMatrix, rhs_per_slot, N; adjoint = false)
    n, nrhs = size(IS)
    TN = zeros(ComplexF64, N, nrhs)
    if adjoint
        G = Matrix{ComplexF64}(IS)
        rc = GC.@preserve G TN ccall((:peec_precond_rhs_batched_adjoint, LIBPEEC), Cint,
            (Ptr{Cvoid}, Cint, Cint, Ptr{ComplexF64}, Ptr{ComplexF64}), ctx, nrhs, rhs_per_slot, G, TN)
    else
        R = Matrix{Float64}(IS)
        rc = GC.@preserve R TN ccall((:peec_precond_rhs_batched, LIBPEEC), Cint,
            (Ptr{Cvoid}, Cint, Cint, Ptr{Float64}, Ptr{ComplexF64}), ctx, nrhs, rhs_per_slot, R, TN)
    end
    peec_check(rc, "peec_precond_rhs_batched")
    TN
end

# Restarted GMRES in FP64 iterative refinement for all columns of B at once
# (mixed = true: complex64 FFT convolutions inside the cycles).
function peec_gmres_batched(ctx, B::Matrix{ComplexF64}, rhs_per_slot, tol; X0 = zeros(ComplexF64, size(B)),
                            restart = 100, max_cycles = 20, mixed = false, inner_tol = 0.0, adjoint = false)
    N, nrhs = size(B)
    X = copy(X0)
    flag = zeros(Cint, nrhs); relres = zeros(Float64, nrhs); iters = zeros(Cint, nrhs); cycles = zeros(Cint, nrhs)
    rc = GC.@preserve B X flag relres iters cycles begin
        if adjoint
            ccall((:peec_gmres_batched_adjoint, LIBPEEC), Cint,
                (Ptr{Cvoid}, Cint, Cint, Ptr{ComplexF64}, Float64, Cint, Cint, Cint, Float64, Ptr{ComplexF64},
                 Ptr{Cint}, Ptr{Float64}, Ptr{Cint}, Ptr{Cint}),
                ctx, nrhs, rhs_per_slot, B, tol, restart, max_cycles, mixed, inner_tol, X, flag, relres, iters, cycles)
        else
            ccall((:peec_gmres_batched, LIBPEEC), Cint,
                (Ptr{Cvoid}, Cint, Cint, Ptr{ComplexF64}, Float64, Cint, Cint, Cint, Float64, Ptr{ComplexF64},
                 Ptr{Cint}, Ptr{Float64}, Ptr{Cint}, Ptr{Cint}),
                ctx, nrhs, rhs_per_slot, B, tol, restart, max_cycles, mixed, inner_tol, X, flag, relres, iters, cycles)
        end
    end
    peec_check(rc, "peec_gmres_batched")
    X, flag, relres, iters, cycles
end

# Per-edge data for DZ on the device (FFT_solver_QS_S_type.jl:106-115):
# R, Cd m x 4 (R real, Cd complex), Zs0 (m), Lp (m), surface flags (m).
function peec_edge_terms_set(ctx, R::Matrix{Float64}, Cd::Matrix{ComplexF64}, Zs0::Vector{ComplexF64},
                             Lp::Vector{Float64}, surface::AbstractVector{Bool}, esc_R, esc_freq)
    m = length(Lp)
    Rt = Matrix(transpose(R)); Ct = Matrix(transpose(Cd))   # 4 per edge, edge after edge
    sf = Int8.(surface)
    rc = GC.@preserve Rt Ct Zs0 Lp sf ccall((:peec_edge_terms_set, LIBPEEC), Cint,
        (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{ComplexF64}, Ptr{ComplexF64}, Ptr{Float64}, Ptr{Int8}, Float64, Float64),
        ctx, m, Rt, Ct, Zs0, Lp, sf, esc_R, esc_freq)
    peec_check(rc, "peec_edge_terms_set")
end

function peec_frequency_set_slot_edges(ctx, slot, w, Yle::SparseMatrixCSC{ComplexF64,Int64})
    Y = CscC128(Yle)
    rc = GC.@preserve Yle ccall((:peec_frequency_set_slot_edges, LIBPEEC), Cint,
        (Ptr{Cvoid}, Cint, Float64, Ref{CscC128}), ctx, slot, w, Y)
    peec_check(rc, "peec_frequency_set_slot_edges")
end

peec_lu_factor_slots(ctx, slot0, count) =
    peec_check(ccall((:peec_lu_factor_slots, LIBPEEC), Cint, (Ptr{Cvoid}, Cint, Cint), ctx, slot0, count),
               "peec_lu_factor_slots")

# One batch of the sweep on the device: port_nodes is P x 2 (1-based node
# indices, as the reference's port_nodes); returns V[s, c1, c2], the node
# voltage difference of port c2 in the solution of (slot s, port c1).
function peec_sweep_batch(ctx, port_nodes::AbstractMatrix, Is, nslots; x0 = nothing, tol = 1e-9, restart = 100,
                          max_cycles = 20, mixed = true, inner_tol = 0.0)
    P = size(port_nodes, 1)
    nrhs = nslots * P
    pn = Int64[port_nodes[c, k] - 1 for k in 1:2, c in 1:P]   # (n1, n2) per port, 0-based
    vp = zeros(ComplexF64, P, nrhs)                           # vp[c2, r]
    flag = zeros(Cint, nrhs); relres = zeros(Float64, nrhs); iters = zeros(Cint, nrhs); cycles = zeros(Cint, nrhs)
    X0 = x0 === nothing ? C_NULL : Matrix{ComplexF64}(x0)
    rc = GC.@preserve pn vp flag relres iters cycles X0 ccall((:peec_sweep_batch, LIBPEEC), Cint,
        (Ptr{Cvoid}, Cint, Ptr{Int64}, Float64, Ptr{ComplexF64}, Float64, Cint, Cint, Cint, Float64,
         Ptr{ComplexF64}, Ptr{ComplexF64}, Ptr{Cint}, Ptr{Float64}, Ptr{Cint}, Ptr{Cint}),
        ctx, P, pn, Is, X0, tol, restart, max_cycles, mixed, inner_tol, vp, C_NULL, flag, relres, iters, cycles)
    peec_check(rc, "peec_sweep_batch")
    V = permutedims(reshape(vp, P, P, nslots), (3, 2, 1))     # V[s, c1, c2]
    V, flag, relres, iters, cycles
end

peec_fft_tune(ctx, nrhs; fp64 = true, fp32 = true) =
    peec_check(ccall((:peec_fft_tune, LIBPEEC), Cint, (Ptr{Cvoid}, Cint, Cint, Ptr{Float64}), ctx, nrhs,
                     Cint(fp64) | (Cint(fp32) << 1), C_NULL), "peec_fft_tune")

# ---------------------------------------------------------------------------
# ChainRulesCore rule for J(sigma) = sum_f |Z_pp(f)|^2.
#
#   solve_ports(sigma, f)  -> (S::Matrix{ComplexF64} P x P, V::Matrix{ComplexF64} N x P)
#       the forward sweep at one frequency (batched or plain path above),
#   adjoint_solve(sigma, f, g) -> lambda::Vector{ComplexF64} (N)
#       K^T lambda = g through peec_gmres_batched(...; adjoint = true) on
#       conj(g), conjugated back,
#   dDZ(sigma, f) -> SparseMatrixCSC{ComplexF64} m x nvox,  d DZ_e / d sigma_v
#       (half-voxel resistances R = l/2 / (sigma W T), compute_diagonals.jl:10-16,
#       and the skin term where it wins the switch, FFT_solver_QS_S_type.jl:107-114).
#
# With A = (I - S)^-1: dZ_pp = -4 lambda_I' (dDZ .* (V_I * A[p, :])), where
# K^T lambda = sum_j A[j, p] g_j and g_j selects the node-voltage difference
# of port j (S_ij = 2 g_j' V_i / R0 - delta_ij, dV_i = -K^-1 dK V_i, dK =
# diag(dDZ) on the edge block; the Python mirror is
# solver_amd/sweep.py: impedance_objective_gradient).
using ChainRulesCore
using LinearAlgebra

struct PortObjective{F1,F2,F3}
    freqs::Vector{Float64}
    port::Int
    m::Int
    port_vector::Function      # weights -> node excitation vector g (length N, zero off the node block)
    solve_ports::F1
    adjoint_solve::F2
    dDZ::F3
    R0::Float64
end

function (J::PortObjective)(sigma)
    s = 0.0
    for f in J.freqs
        S, _ = J.solve_ports(sigma, f)
        Z = J.R0 * ((I - S) \ (I + S))
        s += abs2(Z[J.port, J.port])
    end
    s
end

function ChainRulesCore.rrule(J::PortObjective, sigma)
    val = 0.0
    grad = zeros(length(sigma))
    for f in J.freqs
        S, V = J.solve_ports(sigma, f)
        P = size(S, 1)
        A = inv(Matrix{ComplexF64}(I, P, P) - S)
        Z = J.R0 * A * (I + S)
        z = Z[J.port, J.port]
        lam = J.adjoint_solve(sigma, f, J.port_vector(A[:, J.port]))
        U = V[1:J.m, :] * A[J.port, :]
        dZ = -4 .* (transpose(J.dDZ(sigma, f)) * (lam[1:J.m] .* U))
        val += abs2(z)
        grad .+= 2 .* real.(conj(z) .* dZ)
    end
    pullback(dJ) = (NoTangent(), dJ .* grad)
    return val, pullback
end
