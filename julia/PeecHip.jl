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
# Device state behind the reference's FFT-solver functions.  The spectra, the
# operator, the frequency terms and the LU factors live in the context; the
# Julia objects the reference passes around (FFTCP, FFTCLp, lu, the FFTW
# plans) become handles or are only used to tell whether the device copy is
# current.

# What compute_FFT_mutual_coupling_mats returns in place of the spectra:
# FFTCP and FFTCLp are both this handle.
struct PeecSpectra
    ctx::Ptr{Cvoid}
    id::Any
end

mutable struct PeecDeviceState
    operator::UInt64      # objectid of the incidence_selection uploaded
    freq_key::Any         # (w, objectid(DZ), objectid(Yle)) of the frequency terms uploaded
    factors::UInt64       # objectid of the LU object uploaded
end
const PEEC_STATE = Dict{Ptr{Cvoid},PeecDeviceState}()
const PEEC_CURRENT = Ref{Ptr{Cvoid}}(C_NULL)   # context of the running sweep (precond_3_3_Kt_fft! has no id)
peec_state(ctx) = get!(() -> PeecDeviceState(0, nothing, 0), PEEC_STATE, ctx)

function peec_kernels_build(ctx, origin, Nx, Ny, Nz, sx, sy, sz, esc_Lp, esc_P)
    o = Float64[origin[1], origin[2], origin[3]]
    rc = GC.@preserve o ccall((:peec_kernels_build, LIBPEEC), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Cint, Cint, Cint, Float64, Float64, Float64, Float64, Float64),
        ctx, o, Nx, Ny, Nz, sx, sy, sz, esc_Lp, esc_P)
    peec_check(rc, "peec_kernels_build")
end

# A, Gamma, mat_map_Lp[:, 1], exp_P (3 x 3, column-major) and diag(P)
function peec_operator_set(ctx, incidence_selection, expansions, diag_P)
    st = peec_state(ctx)
    st.operator == objectid(incidence_selection) && return
    A = SparseMatrixCSC{Float64,Int64}(incidence_selection["A"])
    G = SparseMatrixCSC{Float64,Int64}(incidence_selection["Gamma"])
    lp = [SparseMatrixCSC{Float64,Int64}(expansions["mat_map_Lp"][d, 1]) for d in 1:3]
    pm = [SparseMatrixCSC{Float64,Int64}(expansions["exp_P"][i, j]) for i in 1:3, j in 1:3]   # vec: i fastest
    dP = Vector{Float64}(vec(diag_P))
    lpc = [CscF64(x) for x in lp]
    pmc = [CscF64(x) for x in vec(pm)]
    rc = GC.@preserve A G lp pm dP lpc pmc ccall((:peec_operator_set, LIBPEEC), Cint,
        (Ptr{Cvoid}, Int64, Int64, Int64, Ref{CscF64}, Ref{CscF64}, Ptr{CscF64}, Ptr{CscF64}, Ptr{Float64}),
        ctx, size(lp[1], 2), size(lp[2], 2), size(lp[3], 2), CscF64(A), CscF64(G), lpc, pmc, dP)
    peec_check(rc, "peec_operator_set")
    st.operator = objectid(incidence_selection)
    st.freq_key = nothing
    st.factors = 0
end

function peec_frequency_set(ctx, w, DZ::AbstractVecOrMat, Yle::SparseMatrixCSC{ComplexF64,Int64})
    st = peec_state(ctx)
    key = (w, objectid(DZ), objectid(Yle))
    st.freq_key == key && return
    dz = Vector{ComplexF64}(vec(DZ))
    rc = GC.@preserve dz Yle ccall((:peec_frequency_set, LIBPEEC), Cint,
        (Ptr{Cvoid}, Float64, Ptr{ComplexF64}, Ref{CscC128}), ctx, w, dz, CscC128(Yle))
    peec_check(rc, "peec_frequency_set")
    st.freq_key = key
end

# UMFPACK's F.L * F.U == (F.Rs .* SS)[F.p, F.q] is the library's convention
function peec_lu_set_factors(ctx, F)
    st = peec_state(ctx)
    st.factors == objectid(F) && return
    L = SparseMatrixCSC{ComplexF64,Int64}(F.L); U = SparseMatrixCSC{ComplexF64,Int64}(F.U)
    p = Int64.(F.p) .- 1; q = Int64.(F.q) .- 1; Rs = Vector{Float64}(F.Rs)
    rc = GC.@preserve L U p q Rs ccall((:peec_lu_set_factors, LIBPEEC), Cint,
        (Ptr{Cvoid}, Int64, Ref{CscC128}, Ref{CscC128}, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}),
        ctx, size(L, 1), CscC128(L), CscC128(U), p, q, Rs)
    peec_check(rc, "peec_lu_set_factors")
    st.factors = objectid(F)
end

# compute_FFT_mutual_coupling_mats.jl:1-33.  The nine kernel rows and their
# spectra are computed on the device from the grid description: the voxel
# origin follows from the first Lp centre (build_centers_Lp_with_air.jl,
# build_center_air_Voxels.jl: centres at origin + s .* (i - 1/2), Lp centres
# half a cell further along their axis).
function compute_FFT_mutual_coupling_mats(circulant_centers, escalings, Nx, Ny, Nz, QS_Rcc_FW, id, chan)
    QS_Rcc_FW in (1, 2) || error("libpeec_hip builds the quasi-static (1) and retarded Rcc (2) kernels")
    ctx = peec_ctx(id); peec_sync_stop!(id)
    sx, sy, sz = circulant_centers["sx"], circulant_centers["sy"], circulant_centers["sz"]
    half = (sx / 2, sy / 2, sz / 2)
    axis = findfirst(d -> size(circulant_centers[d], 1) > 0, ("Lpx", "Lpy", "Lpz"))
    axis === nothing && error("compute_FFT_mutual_coupling_mats: a 1 x 1 x 1 grid has no edges")
    c = circulant_centers[("Lpx", "Lpy", "Lpz")[axis]][1, :]
    origin = [c[d] - half[d] - (d == axis ? half[d] : 0.0) for d in 1:3]
    if QS_Rcc_FW == 1
        peec_kernels_build(ctx, origin, Nx, Ny, Nz, sx, sy, sz, escalings["Lp"], escalings["P"])
    else
        # compute_rows_Rcc_P + compute_rows_Rcc_Lp: static rows and distances;
        # the spectra follow per frequency (peec_rcc_spectra)
        o = Float64[origin[1], origin[2], origin[3]]
        rc = GC.@preserve o ccall((:peec_kernels_build_rcc, LIBPEEC), Cint,
            (Ptr{Cvoid}, Ptr{Float64}, Cint, Cint, Cint, Float64, Float64, Float64, Float64, Float64),
            ctx, o, Nx, Ny, Nz, sx, sy, sz, escalings["Lp"], escalings["P"])
        peec_check(rc, "peec_kernels_build_rcc")
    end
    PEEC_CURRENT[] = ctx
    h = PeecSpectra(ctx, id)
    return h, h
end

# DZ of one frequency (FFT_solver_QS_S_type.jl:106-114): the self impedance
# of every edge plus its skin-effect impedance, of which only the larger
# real part is kept on surface edges, plus j w Lp.
function peec_edge_impedance(diagonals, escalings, Zs_info, wk)
    Zself = vec(compute_Z_self(diagonals["R"], diagonals["Cd"], wk))
    Zs = vec(escalings["R"] * (Zs_info["Zs"] * sqrt(wk / escalings["freq"])))
    se = vec(Zs_info["surface_edges"])
    dre = real.(Zs[se]) .- real.(Zself[se])
    keep_skin = se[dre .> 0]
    keep_self = se[dre .< 0]
    Zself[keep_skin] .= im .* imag.(Zself[keep_skin])
    Zs[keep_self] .= im .* imag.(Zs[keep_self])
    return (Zself .+ real.(Zs)) .+ im * wk .* vec(diagonals["Lp"])
end

# FFT_solver_QS_S_type.jl:1-234.  Same arguments and the same Dict result
# ("S", "Z", "Y", "f").  Per frequency the host still assembles SS and
# factors it with UMFPACK (the reference's pivoting and scaling); the
# frequency terms and factors then go to the device, where every port is
# preconditioned and solved by gmres_custom's algorithm, warm-started from
# the previous frequency's solution as the reference does (Vrest).
function FFT_solver_QS_S_type(freq, escalings, incidence_selection, FFTCP, FFTCLp, diagonals, ports,
                              ports_scatter_value, lumped_elements, expansions, GMRES_settings, Zs_info, QS_Rcc_FW,
                              id, chan, commentsEnabled::Bool)
    QS_Rcc_FW in (1, 2) || error("libpeec_hip runs the quasi-static (1) and retarded Rcc (2) solvers")
    FFTCP isa PeecSpectra || error("FFT_solver_QS_S_type: FFTCP must come from compute_FFT_mutual_coupling_mats")
    ctx = FFTCP.ctx; PEEC_CURRENT[] = ctx
    A = incidence_selection["A"]; Gamma = incidence_selection["Gamma"]
    m, n = size(A); ns = size(Gamma, 2)
    N = m + ns + n
    peec_operator_set(ctx, incidence_selection, expansions, diagonals["P"])
    invP = sparse(1:ns, 1:ns, 1 ./ vec(diagonals["P"]), ns, ns)
    fs = freq .* escalings["freq"]
    P = size(ports["port_nodes"], 1)
    S = zeros(ComplexF64, P, P, length(fs))
    Vrest = zeros(ComplexF64, N, P)
    maxit = GMRES_settings["Inner_Iter"]
    for k in eachindex(fs)
        peec_sync_stop!(id)
        is_stop_requested(id) && return nothing
        wk = 2pi * fs[k]
        if QS_Rcc_FW == 2   # compute_Circulant_Lp_Rcc / compute_Circulant_P_sup_Rcc (:61-64)
            peec_check(ccall((:peec_rcc_spectra, LIBPEEC), Cint, (Ptr{Cvoid}, Float64), ctx, fs[k] / escalings["freq"]),
                       "peec_rcc_spectra")
        end
        Yle = build_Yle_FFT(lumped_elements, [], ports, escalings, n, wk / escalings["freq"], ports_scatter_value)
        DZ = peec_edge_impedance(diagonals, escalings, Zs_info, wk)
        peec_frequency_set(ctx, wk, DZ, Yle)
        invZ = sparse(1:m, 1:m, 1 ./ DZ, m, m)
        F = lu(Yle + transpose(A) * (invZ * A) + im * wk * (Gamma * invP) * transpose(Gamma))
        peec_lu_set_factors(ctx, F)
        for c1 in 1:P
            n1, n2 = Int(ports["port_nodes"][c1, 1]), Int(ports["port_nodes"][c1, 2])
            is = zeros(Float64, n); is[n1] = escalings["Is"]; is[n2] = -escalings["Is"]
            tn = Vector{ComplexF64}(undef, N)
            peec_check(GC.@preserve is tn ccall((:peec_precond_rhs, LIBPEEC), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Ptr{ComplexF64}), ctx, is, tn), "peec_precond_rhs")
            V, flag, relres, it, resvec = gmres_custom(tn, false, GMRES_settings["tol"][k], maxit, Vrest[:, c1], wk,
                incidence_selection, FFTCP, FFTCLp, DZ, Yle, expansions, invZ, invP, F, nothing, nothing, nothing,
                nothing, nothing, id, chan, c1)
            flag == 99 && return nothing
            commentsEnabled && println("Flag $flag - Iteration = $k - port $c1, iterations $(it)")
            Vrest[:, c1] = V
            for c2 in c1:P
                n3, n4 = Int(ports["port_nodes"][c2, 1]), Int(ports["port_nodes"][c2, 2])
                dv = 2 * (V[m+ns+n3] - V[m+ns+n4])
                S[c1, c2, k] = c1 == c2 ? (dv - ports_scatter_value) / ports_scatter_value : dv / ports_scatter_value
                S[c2, c1, k] = S[c1, c2, k]
            end
        end
    end
    return Dict("S" => S, "Z" => s2z_fft(S, ports_scatter_value), "Y" => s2y_fft(S, ports_scatter_value),
                "f" => fs ./ escalings["freq"])
end

# compute_Matrix_vector_fft.jl:1-81: y = M^-1 K x.  The spectra, plans and
# factors the reference passes are handles here; the device copies of the
# frequency terms (w, DZ, Yle) and of the factors (lu) are refreshed when
# the caller passes different objects.
function ComputeMatrixVectorFTT(x, w::Float64, incidence_selection::Dict, FFTCP, FFTCLp, DZ, Yle, expansions, invZ,
                                invP, lu, PLIVector, PVector, PLI2Vector, P2Vector, chi2Vector)
    FFTCP isa PeecSpectra || error("ComputeMatrixVectorFTT: FFTCP must come from compute_FFT_mutual_coupling_mats")
    ctx = FFTCP.ctx; PEEC_CURRENT[] = ctx
    peec_state(ctx).operator == objectid(incidence_selection) ||
        error("ComputeMatrixVectorFTT: the operator of this geometry is not on the device (FFT_solver_QS_S_type sets it)")
    peec_frequency_set(ctx, w, DZ, Yle)
    peec_lu_set_factors(ctx, lu)
    xx = Vector{ComplexF64}(vec(x)); y = similar(xx)
    peec_check(GC.@preserve xx y ccall((:peec_matvec, LIBPEEC), Cint,
        (Ptr{Cvoid}, Ptr{ComplexF64}, Ptr{ComplexF64}), ctx, xx, y), "peec_matvec")
    return y
end

# FFT_solver_QS_S_type.jl:237-262: Y = M^-1 [0; 0; X3] with the factors F
# (uploaded if F is not the object on the device).  The context is the one
# of the running sweep.
function precond_3_3_Kt_fft!(F, invZ, invP, A, Gamma, n1, n2, X3, Y, resProd)
    ctx = PEEC_CURRENT[]
    ctx == C_NULL && error("precond_3_3_Kt_fft!: no libpeec_hip context (compute_FFT_mutual_coupling_mats creates it)")
    peec_lu_set_factors(ctx, F)
    is = Vector{Float64}(vec(X3)); tn = Vector{ComplexF64}(undef, length(Y))
    peec_check(GC.@preserve is tn ccall((:peec_precond_rhs, LIBPEEC), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{ComplexF64}), ctx, is, tn), "peec_precond_rhs")
    Y[:] .= tn
    return Y
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
