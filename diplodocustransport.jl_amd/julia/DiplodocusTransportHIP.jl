# DiplodocusTransportHIP.jl -- HIP (AMD MI355X) backend for DiplodocusTransport.jl's explicit steps.
#
# `include` this file inside `module DiplodocusTransport`, after src/Backends.jl and
# src/StructsAndDictionaries.jl.  It adds:
#   * HIPBackend <: BackendType                      (beside CPUBackend/CUDABackend, src/Backends.jl:3-4)
#   * DtxHandle, the owner of one libdtx.so simulation (device state and operators)
#   * dtx_attach!(method, PhaseSpace, Initial, Injection, BinM, EmiM, FluxM, backend)
#       the constructor branch `elseif Backend isa HIPBackend` of ForwardEulerStruct /
#       ForwardSymplecticEulerStruct (src/SteppingStructs.jl:57-75, 175-193)
#   * dtx_step!(method, t_start, t_stop, dt, Verbose) -> (dt, save)
#       the whole functor body (src/SteppingMethods.jl:12-107, 120-243) as one ccall
#   * DtxState <: AbstractVector{Float64}, the value of `method.f` on this backend, so that
#       Solve's `@. method.f = method.f_init` and `copyto!(f, method.f)` (src/Solver.jl:22,40)
#       work unchanged.
#
# Every numerical operation runs in libdtx.so (C ABI: include/dtx.h).  Arrays are passed in the
# reference's own layouts: M_Bin (N^2 x N, column-major), SparseMatrixCSC{Float64,Int64}
# (1-based), state vectors in GlobalIndices_To_StateIndex order.

const libdtx = get(ENV, "DTX_LIBRARY", normpath(joinpath(@__DIR__, "..", "libdtx.so")))

const DTX_MAX_SPECIES = 16
const DTX_PRECISIONS = Dict(:FP64 => Int32(0), :FP32 => Int32(1), :BF16x3 => Int32(2))

"""
    HIPBackend(; device=0, precision=:FP64)

MI355X backend.  `precision` selects the arithmetic of the binary gain contraction (`:FP64`,
`:FP32` MFMA, or `:BF16x3` split-bf16 MFMA); f, the Euler update, losses and moments stay Float64.
"""
struct HIPBackend <: BackendType
    device::Int32
    precision::Int32
end
HIPBackend(; device::Integer=0, precision::Symbol=:FP64) =
    HIPBackend(Int32(device), get(DTX_PRECISIONS, precision) do
        error("HIPBackend precision must be :FP64, :FP32 or :BF16x3")
    end)

function Base.zeros(::HIPBackend, T::Type{<:AbstractFloat}, dims::Int...)
    return zeros(T, dims...)   # host scratch only: device buffers belong to the DtxHandle
end

# Mirrors of the C structs (same field order and types => same layout).
struct DtxConfig
    num_species::Int32
    px_num::NTuple{DTX_MAX_SPECIES,Int32}
    py_num::NTuple{DTX_MAX_SPECIES,Int32}
    pz_num::NTuple{DTX_MAX_SPECIES,Int32}
    x_num::Int32
    y_num::Int32
    z_num::Int32
    precision::Int32
    method::Int32
    device::Int32
    dt0::Float64
end

struct DtxStepStats
    finite::Int32
    cr_evaluated::Int32
    save::Int32
    step_index::Int32
    cr::Float64
    cr_bin::Float64
    cr_emi::Float64
    cr_flux::Float64
    cr_momentum::Float64
    cr_space::Float64
    dt::Float64
    t_start::Float64
    sum_f::Float64
end

mutable struct DtxHandle
    ptr::Ptr{Cvoid}
    n::Int
    num_species::Int
    n_space::Int
end

function dtx_error(ptr::Ptr{Cvoid})
    msg = ccall((:dtx_last_error, libdtx), Cstring, (Ptr{Cvoid},), ptr)
    return unsafe_string(msg)
end

dtx_check(h::DtxHandle, rc::Cint) = rc == 0 || error("libdtx: " * dtx_error(h.ptr))

function dtx_close!(h::DtxHandle)
    if h.ptr != C_NULL
        ccall((:dtx_destroy, libdtx), Cvoid, (Ptr{Cvoid},), h.ptr)
        h.ptr = C_NULL
    end
    return nothing
end

"""
    DtxState(handle)

`method.f` on the HIP backend: the state lives in HBM.  Reads copy the whole state to the host
(dtx_get_f); `copyto!(state, v)` and broadcast assignment upload it (dtx_set_f).
"""
struct DtxState <: AbstractVector{Float64}
    h::DtxHandle
end
Base.size(s::DtxState) = (s.h.n,)
Base.IndexStyle(::Type{DtxState}) = IndexLinear()
function Base.copyto!(dest::AbstractVector{Float64}, src::DtxState)
    length(dest) == src.h.n || throw(DimensionMismatch("state length"))
    buf = dest isa Vector{Float64} ? dest : Vector{Float64}(undef, src.h.n)
    dtx_check(src.h, ccall((:dtx_get_f, libdtx), Cint, (Ptr{Cvoid}, Ptr{Float64}), src.h.ptr, buf))
    buf === dest || copyto!(dest, buf)
    return dest
end
Base.getindex(s::DtxState, i::Int) = copyto!(Vector{Float64}(undef, s.h.n), s)[i]
Base.collect(s::DtxState) = copyto!(Vector{Float64}(undef, s.h.n), s)
function Base.copyto!(dest::DtxState, src::AbstractVector)
    v = convert(Vector{Float64}, src)
    length(v) == dest.h.n || throw(DimensionMismatch("state length"))
    dtx_check(dest.h, ccall((:dtx_set_f, libdtx), Cint, (Ptr{Cvoid}, Ptr{Float64}), dest.h.ptr, v))
    return dest
end
# `@. method.f = method.f_init` (src/Solver.jl:22) lowers to materialize!(dest, bc)
Base.copyto!(dest::DtxState, bc::Base.Broadcast.Broadcasted) =
    copyto!(dest, Base.materialize(bc))

"""
    dtx_energy_weights(PhaseSpace) -> Vector{Float64}

Per-state T^00 weight of one cell, written as StressEnergyTensor writes it
(src/DistributionMoments.jl:145 massive, :162 massless) and expanded over (py, pz) in the state
order of src/GlobalToStateIndices.jl:7-25.
"""
function dtx_energy_weights(PhaseSpace)
    Momentum = PhaseSpace.Momentum
    Grids = PhaseSpace.Grids
    w = Float64[]
    for s in eachindex(PhaseSpace.name_list)
        pr = Grids.pxr_list[s]
        m = Grids.mass_list[s]
        px = Momentum.px_num_list[s]
        wp = Vector{Float64}(undef, px)
        for p in 1:px
            wp[p] = m != 0e0 ?
                (pr[p]sqrt(m^2 + pr[p]^2) - pr[p+1]sqrt(m^2 + pr[p+1]^2) +
                 m^2*(asinh(pr[p]/m) - asinh(pr[p+1]/m))) / (2(pr[p] - pr[p+1])) :
                (pr[p] + pr[p+1]) / 2
        end
        append!(w, repeat(wp, Momentum.py_num_list[s] * Momentum.pz_num_list[s]))
    end
    return w
end

"""
    dtx_attach!(method, PhaseSpace, Initial, Injection, BinM, EmiM, FluxM, backend; symplectic)

Creates the libdtx simulation for `method` and uploads its operators.  Called from the
`elseif Backend isa HIPBackend` branch of the stepping-struct constructors; afterwards
`method.f` is a DtxState and `method.HIP` holds the handle.

With `binary_arrays` (one `(name_locs, GainMatrix3, GainMatrix4, LossMatrix1, LossMatrix2,
GainScale, LossScale)` per interaction, as LoadMatrices_Binary forms them) the binary operator
is assembled on the handle from the collision arrays (`binary_mode = (mode=Ani(), sparse=false)`
as Bin_Mode / Bin_sparse) and `BinM.M_Bin` is not read: BuildBinaryMatrices' N^2 x N array need
not be built.
"""
function dtx_attach!(method, PhaseSpace, Initial::Vector{Float64}, Injection::Vector{Float64},
                     BinM, EmiM, FluxM, backend::HIPBackend; symplectic::Bool=false,
                     flux_synth=nothing, binary_arrays=nothing,
                     binary_mode=(mode=nothing, sparse=false))
    Momentum = PhaseSpace.Momentum
    Space = PhaseSpace.Space
    S = length(PhaseSpace.name_list)
    S <= DTX_MAX_SPECIES || error("HIPBackend supports at most $DTX_MAX_SPECIES species")
    pad(v) = ntuple(i -> i <= S ? Int32(v[i]) : Int32(0), DTX_MAX_SPECIES)
    cfg = DtxConfig(Int32(S), pad(Momentum.px_num_list), pad(Momentum.py_num_list),
                    pad(Momentum.pz_num_list), Int32(Space.x_num), Int32(Space.y_num),
                    Int32(Space.z_num), backend.precision, Int32(symplectic ? 1 : 0),
                    backend.device, Float64(PhaseSpace.Time.dt0))
    out = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:dtx_create, libdtx), Cint, (Ref{DtxConfig}, Ref{Ptr{Cvoid}}), cfg, out)
    rc == 0 || error("libdtx: " * dtx_error(C_NULL))
    n = length(Initial)
    h = DtxHandle(out[], n, S, Space.x_num * Space.y_num * Space.z_num)
    finalizer(dtx_close!, h)

    N = sum(Momentum.px_num_list .* Momentum.py_num_list .* Momentum.pz_num_list)
    set_csc(which, A) = begin
        F = SparseMatrixCSC{Float64,Int64}(A)
        dtx_check(h, ccall((:dtx_set_flux_csc, libdtx), Cint,
                           (Ptr{Cvoid}, Int32, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}),
                           h.ptr, Int32(which), Int64(n), F.colptr, F.rowval, F.nzval))
    end
    if flux_synth !== nothing
        # the synthetic Cartesian stencils, 1 ./ Ap and Vol built on the device instead of the
        # FluxM CSC upload; dtx_set_vectors then takes C_NULL for Ap and Vol
        dtx_synth_flux_handle!(h, flux_synth...)
    elseif symplectic
        set_csc(1, FluxM.X_Flux)     # DTX_FLUX_X
        set_csc(2, FluxM.P_Flux)     # DTX_FLUX_P
    else
        set_csc(0, FluxM.X_Flux + FluxM.P_Flux)   # F_Flux (src/SteppingStructs.jl:60)
    end
    if !isempty(BinM.Binary_list)
        if binary_arrays !== nothing
            # the collision arrays of every interaction, in Binary_list order: tuples
            # (name_locs, GainMatrix3, GainMatrix4, LossMatrix1, LossMatrix2, GainScale,
            # LossScale) as LoadMatrices_Binary forms them, assembled on the handle
            for (locs, G3, G4, L1, L2, gs, ls) in binary_arrays
                dtx_add_binary_interaction!(h, PhaseSpace, locs, G3, G4, L1, L2;
                                            mode=get(binary_mode, :mode, nothing),
                                            sparse=get(binary_mode, :sparse, false),
                                            GainScale=gs, LossScale=ls)
            end
        elseif BinM.M_Bin isa AbstractSparseMatrix
            size(BinM.M_Bin) == (N^2, N) || error("M_Bin must be (N^2 x N) with N = $N")
            # Bin_sparse: the CSC as Julia holds it (src/SteppingStructs.jl:178-179)
            B = SparseMatrixCSC{Float64,Int64}(BinM.M_Bin)
            dtx_check(h, ccall((:dtx_set_binary_csc, libdtx), Cint,
                               (Ptr{Cvoid}, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}),
                               h.ptr, Int64(N), B.colptr, B.rowval, B.nzval))
        else
            size(BinM.M_Bin) == (N^2, N) || error("M_Bin must be (N^2 x N) with N = $N")
            M = Matrix{Float64}(BinM.M_Bin)
            dtx_check(h, ccall((:dtx_set_binary_dense, libdtx), Cint,
                               (Ptr{Cvoid}, Int64, Ptr{Float64}), h.ptr, Int64(N), M))
        end
        dom = BinM.Domain
        if dom === nothing
            dtx_check(h, ccall((:dtx_set_binary_domain, libdtx), Cint,
                               (Ptr{Cvoid}, Int64, Ptr{Int64}), h.ptr, Int64(-1), C_NULL))
        else
            d = Vector{Int64}(dom)
            dtx_check(h, ccall((:dtx_set_binary_domain, libdtx), Cint,
                               (Ptr{Cvoid}, Int64, Ptr{Int64}), h.ptr, Int64(length(d)), d))
        end
    end
    if !isempty(EmiM.Emission_list)
        if EmiM.M_Emi isa AbstractSparseMatrix
            # Emi_sparse (src/SteppingStructs.jl:183-184): no dense n x n copy
            E = SparseMatrixCSC{Float64,Int64}(EmiM.M_Emi)
            dtx_check(h, ccall((:dtx_set_emission_csc, libdtx), Cint,
                               (Ptr{Cvoid}, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}),
                               h.ptr, Int64(n), E.colptr, E.rowval, E.nzval))
        else
            E = Matrix{Float64}(EmiM.M_Emi)
            dtx_check(h, ccall((:dtx_set_emission_dense, libdtx), Cint,
                               (Ptr{Cvoid}, Int64, Ptr{Float64}), h.ptr, Int64(n), E))
        end
    end
    ap, vol = flux_synth === nothing ?
        (Vector{Float64}(FluxM.Ap_Flux), Vector{Float64}(FluxM.Vol)) :
        (Ptr{Float64}(C_NULL), Ptr{Float64}(C_NULL))
    dtx_check(h, ccall((:dtx_set_vectors, libdtx), Cint,
                       (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                       h.ptr, ap, vol, Initial, Injection))
    dtx_check(h, ccall((:dtx_set_moment_weights, libdtx), Cint, (Ptr{Cvoid}, Ptr{Float64}),
                       h.ptr, dtx_energy_weights(PhaseSpace)))
    method.HIP = h
    method.f = DtxState(h)
    return method
end

"""
    dtx_step!(method, t_start, t_stop, dt, Verbose) -> (dt, save)

The functor on the HIP backend: the dt clip, dt_scale and the whole update run in libdtx; the
CFL / non-finite diagnostics are printed with the reference's wording when the device reports
them.  With the `save` result true the call also waits for the step to finish.
"""
function dtx_step!(method, t_start, t_stop, dt, Verbose::Int64)
    h = method.HIP
    dt_out = Ref{Float64}(0.0)
    save = Ref{Int32}(0)
    dtx_check(h, ccall((:dtx_step, libdtx), Cint,
                       (Ptr{Cvoid}, Float64, Float64, Float64, Int32, Ref{Float64}, Ref{Int32},
                        Ptr{DtxStepStats}),
                       h.ptr, Float64(t_start), Float64(t_stop), Float64(dt), Int32(Verbose),
                       dt_out, save, C_NULL))
    if save[] != 0
        dtx_check(h, ccall((:dtx_synchronize, libdtx), Cint, (Ptr{Cvoid},), h.ptr))
    end
    return dt_out[], save[] != 0
end

"""
    dtx_moments(method) -> (number, energy)

Number N^0 and energy T^00 of the current state per (cell, species), each a
(num_species x n_space) matrix, from the FP64 sums the step kernel keeps.
"""
function dtx_moments(method)
    h = method.HIP
    num = Matrix{Float64}(undef, h.num_species, h.n_space)
    en = Matrix{Float64}(undef, h.num_species, h.n_space)
    dtx_check(h, ccall((:dtx_moments, libdtx), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}),
                       h.ptr, num, en))
    return num, en
end

# dtx_flux_synth of include/dtx.h (field order and types as in C)
struct DtxFluxSynth
    xr::Ptr{Float64}
    yr::Ptr{Float64}
    zr::Ptr{Float64}
    pr::Ptr{Float64}
    ur::Ptr{Float64}
    hr::Ptr{Float64}
    mass::NTuple{DTX_MAX_SPECIES,Float64}
    charge::NTuple{DTX_MAX_SPECIES,Float64}
    sync_K::Float64
    axes::Int32
    periodic::Int32
end

"""
    dtx_synth_flux!(method, xr, yr, zr, pr_list, ur_list, hr_list, mass, charge;
                    sync_K=0.0, axes=(:x,), periodic=true)

Builds the synthetic Cartesian flux operators (upwind advection along the given axes,
synchrotron cooling along p when sync_K > 0), 1 ./ Ap and Vol on the device, in place of the
X_Flux / P_Flux CSC upload (src/Fluxes/BuildFluxMatrices.jl:10-184).  `pr_list[s]`,
`ur_list[s]`, `hr_list[s]` are the bounds of species s; `mass` and `charge` one value per species.
On a method already attached it replaces the uploaded flux operators, 1 ./ Ap and Vol (FluxM is
then no longer what the method steps); to build them on the device from the start, pass the
same arguments to the constructor as `dtx_attach!(...; flux_synth=(xr, yr, zr, pr_list,
ur_list, hr_list, mass, charge))` (keywords as a trailing NamedTuple are not accepted there:
sync_K, axes and periodic go in as positional `sync_K`, `axes`, `periodic` after `charge`).
"""
function dtx_synth_flux!(method, xr, yr, zr, pr_list, ur_list, hr_list, mass, charge;
                         sync_K=0.0, axes=(:x,), periodic=true)
    dtx_synth_flux_handle!(method.HIP, xr, yr, zr, pr_list, ur_list, hr_list, mass, charge,
                           sync_K, axes, periodic)
end

function dtx_synth_flux_handle!(h::DtxHandle, xr, yr, zr, pr_list, ur_list, hr_list, mass,
                                charge, sync_K=0.0, axes=(:x,), periodic=true)
    pad(v) = ntuple(i -> i <= length(v) ? Float64(v[i]) : 0.0, DTX_MAX_SPECIES)
    bufs = map(v -> Vector{Float64}(v),
               (xr, yr, zr, reduce(vcat, pr_list), reduce(vcat, ur_list), reduce(vcat, hr_list)))
    GC.@preserve bufs begin
        spec = DtxFluxSynth(map(pointer, bufs)..., pad(mass), pad(charge), Float64(sync_K),
                            Int32((:x in axes ? 1 : 0) | (:y in axes ? 2 : 0)),
                            Int32(periodic ? 1 : 0))
        dtx_check(h, ccall((:dtx_synth_flux, libdtx), Cint, (Ptr{Cvoid}, Ref{DtxFluxSynth}),
                           h.ptr, spec))
    end
    return nothing
end

# dtx_binary_interaction of include/dtx.h (field order and types as in C)
struct DtxBinaryInteraction
    species::NTuple{4,Int32}
    mode::Int32
    sparse::Int32
    gain_scale::Float64
    loss_scale::Float64
    gain3::Ptr{Float64}
    gain4::Ptr{Float64}
    loss1::Ptr{Float64}
    loss2::Ptr{Float64}
    dpy::Ptr{Float64}
    dpz::Ptr{Float64}
end

const DTX_BIN_MODES = Dict(:Ani => Int32(0), :Axi => Int32(1), :Iso => Int32(2))

"""
    dtx_add_binary_interaction!(h, PhaseSpace, name_locs, GainMatrix3, GainMatrix4,
                                LossMatrix1, LossMatrix2; mode=Ani(), sparse=false,
                                GainScale=1.0, LossScale=1.0)

The `Fill_M_Bin!` call of `LoadMatrices_Binary` (src/Collisions/LoadMatrices.jl:136) on the
handle: adds one interaction's gain and loss terms, with the Ani / Axi / Iso averaging, straight
to the packed operator (no N^2 x N `M_Bin`).  `name_locs` are 1-based as in the reference;
`GainScale` / `LossScale` already include `Bin_Norm` (LoadMatrices.jl:129-133).  `h` is a
`DtxHandle` or a method whose `HIP` field holds one.
"""
function dtx_add_binary_interaction!(h::DtxHandle, PhaseSpace, name_locs, GainMatrix3,
                                     GainMatrix4, LossMatrix1, LossMatrix2; mode=nothing,
                                     sparse::Bool=false, GainScale=1.0, LossScale=1.0)
    m = mode === nothing ? :Ani : nameof(typeof(mode))
    haskey(DTX_BIN_MODES, m) || error("Bin_Mode must be Ani(), Axi() or Iso()")
    Grids = PhaseSpace.Grids
    bufs = (Array{Float64}(GainMatrix3), Array{Float64}(GainMatrix4),
            Array{Float64}(LossMatrix1), Array{Float64}(LossMatrix2),
            Vector{Float64}(reduce(vcat, Grids.dpy_list)),
            Vector{Float64}(reduce(vcat, Grids.dpz_list)))
    GC.@preserve bufs begin
        x = DtxBinaryInteraction(ntuple(i -> Int32(name_locs[i] - 1), 4), DTX_BIN_MODES[m],
                                 Int32(sparse ? 1 : 0), Float64(GainScale), Float64(LossScale),
                                 map(pointer, bufs)...)
        dtx_check(h, ccall((:dtx_add_binary_interaction, libdtx), Cint,
                           (Ptr{Cvoid}, Ref{DtxBinaryInteraction}), h.ptr, x))
    end
    return nothing
end

dtx_add_binary_interaction!(method, args...; kw...) =
    dtx_add_binary_interaction!(method.HIP, args...; kw...)
