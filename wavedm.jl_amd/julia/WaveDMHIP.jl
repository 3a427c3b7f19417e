# WaveDMHIP.jl -- Julia binding of libwavedm_hip.so (include/wavedm.h).
#
# `ccall` wrappers over the C ABI (one `HIPPlan` = one device-resident box),
# plus, in WaveDMHIPBackend.jl (included at the end), methods of WaveDM.jl's
# own solver-step functions at the reference signatures for a
# `HIPDeviceConfig`: apply_kick_step! (src/solver/KDK.jl:37-77),
# apply_drift_step! (:88-101), compute_gravitational_potential
# (src/solver/gravity.jl:7-35), parallel_poisson (src/parallel_poisson.jl:36-48)
# and compute_averaged_fields (src/io/save.jl:79-95).  The plan-level wrappers
# here have names of their own (kick!, drift!, poisson, potential, step!) so
# they never shadow WaveDM's.  Arrays cross as `pointer(A)` (column-major, x
# fastest, interleaved complex), exactly the layout the library expects.
#
# This file is not executed in this repository (the image has no Julia); the
# Python harness wavedm.jl_amd/kdk.py issues the same C-ABI call sequence and
# is what the tests run.
module WaveDMHIP

using ChainRulesCore

export HIPPlan, ssfm, set_psi!, get_psi, set_border!, set_static_potential!, set_hernquist!,
       set_quadrupole!, set_kappa!, set_tidal_model!, poisson, potential, kick!, drift!, step!, reduce_diagnostics,
       split_energy, cfl_limits, step_properties, accelerations, density_accumulate!, density_average,
       density_reset!, set_mass_ledger!, set_ic_velocity!, baryon_mesh_potential

const libwdm = get(ENV, "WAVEDM_HIP_LIB", joinpath(@__DIR__, "..", "lib", "libwavedm_hip.so"))

const WDM_F64 = Cint(0)
const WDM_F32 = Cint(1)
const WDM_C128 = Cint(0)
const WDM_C64 = Cint(1)
const WDM_GRID_CONSISTENT = Cuint(1)
const WDM_SPLIT_REFERENCE = Cint(0)
const WDM_SPLIT_STRANG = Cint(1)
const WDM_V_HERNQUIST = Cint(2)
const WDM_V_QUADRUPOLE = Cint(4)
const WDM_V_KAPPA = Cint(8)
const WDM_RED_COUNT = 6
const WDM_DIAG_COUNT = 22
const WDM_ADJ_DT_SCALE = Cuint(1)
const WDM_ADJ_NPARAM = 12

struct WaveDMHIPError <: Exception
    code::Int
    msg::String
end
Base.showerror(io::IO, e::WaveDMHIPError) = print(io, "libwavedm_hip error ", e.code, ": ", e.msg)

function check(rc::Cint, h::Ptr{Cvoid})
    if rc != 0
        msg = unsafe_string(ccall((:wdm_last_error, libwdm), Cstring, (Ptr{Cvoid},), h))
        throw(WaveDMHIPError(Int(rc), msg))
    end
    return nothing
end

"""
    HIPPlan(Nx, Ny, Nz, Xmax, Ymax, Zmax; precision=:f64, consistent_grid=false,
            hbar_over_m=1.0, four_pi_G=4π)

Device state of one simulation box (ψ, spectra, Poisson buffers, k tables).
The reference grid `x = LinRange(-Xmax, Xmax, Nx)` (Δ = 2Xmax/(Nx-1)) is the
default; `consistent_grid=true` uses Δ = 2Xmax/Nx.  The reference's code units
are `hbar_over_m = 1`, `four_pi_G = 4π`.
"""
mutable struct HIPPlan
    h::Ptr{Cvoid}
    dims::NTuple{3,Int}
    precision::Symbol
    tape::Int  # counts taped forward runs (ssfm); a pullback checks it is still its own
    function HIPPlan(Nx::Integer, Ny::Integer, Nz::Integer, Xmax::Real, Ymax::Real, Zmax::Real;
                     precision::Symbol=:f64, consistent_grid::Bool=false,
                     hbar_over_m::Real=1.0, four_pi_G::Real=4π)
        r = Ref{Ptr{Cvoid}}(C_NULL)
        prec = precision === :f32 ? WDM_F32 : WDM_F64
        flags = consistent_grid ? WDM_GRID_CONSISTENT : Cuint(0)
        rc = ccall((:wdm_plan_create, libwdm), Cint,
                   (Ref{Ptr{Cvoid}}, Cint, Cint, Cint, Cdouble, Cdouble, Cdouble, Cint, Cuint),
                   r, Nx, Ny, Nz, Xmax, Ymax, Zmax, prec, flags)
        check(rc, C_NULL)
        p = new(r[], (Int(Nx), Int(Ny), Int(Nz)), precision, 0)
        finalizer(destroy!, p)
        check(ccall((:wdm_set_coeffs, libwdm), Cint, (Ptr{Cvoid}, Cdouble, Cdouble), p.h, hbar_over_m,
                    four_pi_G), p.h)
        return p
    end
end

function destroy!(p::HIPPlan)
    if p.h != C_NULL
        ccall((:wdm_plan_destroy, libwdm), Cint, (Ptr{Cvoid},), p.h)
        p.h = C_NULL
    end
    return nothing
end

_dtype(::Type{ComplexF64}) = WDM_C128
_dtype(::Type{ComplexF32}) = WDM_C64

function set_psi!(p::HIPPlan, ψ::Array{T,3}) where {T<:Union{ComplexF64,ComplexF32}}
    size(ψ) == p.dims || throw(DimensionMismatch("ψ is $(size(ψ)), plan is $(p.dims)"))
    check(ccall((:wdm_set_psi, libwdm), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint), p.h, ψ, _dtype(T)), p.h)
    return p
end

function get_psi(p::HIPPlan, ::Type{T}=ComplexF64) where {T<:Union{ComplexF64,ComplexF32}}
    ψ = Array{T,3}(undef, p.dims)
    check(ccall((:wdm_get_psi, libwdm), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint), p.h, ψ, _dtype(T)), p.h)
    return ψ
end

"Absorbing border along x (the reference's rank-1 `border`, src/simulation_main.jl:415)."
function set_border!(p::HIPPlan, border::Union{Nothing,AbstractVector{<:Real}})
    b = border === nothing ? C_NULL : Vector{Float64}(border)
    check(ccall((:wdm_set_border, libwdm), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Cint), p.h, b, 1), p.h)
    return p
end

"Static external potential on the grid (replaces `V.(xxx, yyy, zzz, ψ)` when V does not depend on t or ψ)."
function set_static_potential!(p::HIPPlan, V::Union{Nothing,Array{<:Real,3}})
    v = V === nothing ? C_NULL : Array{Float64,3}(V)
    check(ccall((:wdm_set_static_potential, libwdm), Cint, (Ptr{Cvoid}, Ptr{Cdouble}), p.h, v), p.h)
    return p
end

function _model!(p::HIPPlan, kind::Cint, params::Vector{Float64})
    check(ccall((:wdm_set_potential_model, libwdm), Cint, (Ptr{Cvoid}, Cint, Ptr{Cdouble}, Cint),
                p.h, kind, params, length(params)), p.h)
    return p
end
set_hernquist!(p::HIPPlan; M, a, center=(0.0, 0.0, 0.0)) =
    _model!(p, WDM_V_HERNQUIST, Float64[M, a, center...])
set_quadrupole!(p::HIPPlan; T0::AbstractMatrix, axis=(0.0, 0.0, 1.0), period=1.0, center=(0.0, 0.0, 0.0)) =
    _model!(p, WDM_V_QUADRUPOLE, Float64[T0[1, 1], T0[2, 2], T0[3, 3], T0[1, 2], T0[1, 3], T0[2, 3],
                                         axis..., period, center...])
set_kappa!(p::HIPPlan, κ::Real) = _model!(p, WDM_V_KAPPA, Float64[κ])

"""
    set_tidal_model!(p, MW_Phi, lo, hi, traj_t, traj_xyz)

The MW tidal potential in every kick (src/physics/tidal_fields.jl:64-146, applied: the
reference's `Φ_all += pot_tidal` never reaches its kick).  `MW_Phi` on the uniform grid
from `lo` to `hi`; `traj_t` the kick times of the `traj_xyz` rows (an n×3 matrix, the
df_traj positions over uL): a kick at t uses the last row with traj_t ≤ t.
`MW_Phi = nothing` removes it.
"""
function set_tidal_model!(p::HIPPlan, MW_Phi::Union{Nothing,Array{<:Real,3}}, lo=(0.0, 0.0, 0.0),
                          hi=(1.0, 1.0, 1.0), traj_t::AbstractVector{<:Real}=Float64[],
                          traj_xyz::AbstractMatrix{<:Real}=zeros(0, 3))
    if MW_Phi === nothing
        check(ccall((:wdm_set_tidal_model, libwdm), Cint,
                    (Ptr{Cvoid}, Ptr{Cdouble}, Cint, Cint, Cint, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble},
                     Ptr{Cdouble}, Cint), p.h, C_NULL, 0, 0, 0, C_NULL, C_NULL, C_NULL, C_NULL, 0), p.h)
        return p
    end
    tab = Array{Float64,3}(MW_Phi)
    rows = permutedims(Matrix{Float64}(traj_xyz))  # x, y, z of each row contiguous
    check(ccall((:wdm_set_tidal_model, libwdm), Cint,
                (Ptr{Cvoid}, Ptr{Cdouble}, Cint, Cint, Cint, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble},
                 Ptr{Cdouble}, Cint),
                p.h, tab, size(tab)..., Float64[lo...], Float64[hi...], Vector{Float64}(traj_t), rows,
                length(traj_t)), p.h)
    return p
end

"""
    poisson(p::HIPPlan, ρ::Array{Float64,3}; scale=1.0) -> Φ

Periodic ∇²Φ = scale·(ρ - ⟨ρ⟩) on the plan's lattice (k = 2πm/(2Xmax), k=0 → 0),
the `Periodic()` arithmetic of src/parallel_poisson.jl:54-57.  Returns the real part (the
reference's complex output has a round-off imaginary part).  `scale = 0` takes the plan's
4πG.
"""
function poisson(p::HIPPlan, ρ::Array{Float64,3}; scale::Real=1.0)
    size(ρ) == p.dims || throw(DimensionMismatch("ρ is $(size(ρ)), plan is $(p.dims)"))
    Φ = Array{Float64,3}(undef, p.dims)
    check(ccall((:wdm_poisson, libwdm), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble),
                p.h, ρ, Φ, scale), p.h)
    return Φ
end

"""
    baryon_mesh_potential(p::HIPPlan, ρ_baryon; four_pi_G=nothing) -> Φ_b

Φ_b = 4πG ∇⁻²(ρ_baryon - ⟨ρ_baryon⟩) for `baryon_mode = :mesh`
(src/simulation_main.jl:244) on the plan's lattice and precision, real part.
`four_pi_G = nothing` takes the plan's coefficient; an explicit 0 gives Φ_b = 0.  Install
it with `set_static_potential!(p, Φ_b)` so every kick adds it (src/solver/gravity.jl:24-28).
"""
function baryon_mesh_potential(p::HIPPlan, ρ_baryon::Array{Float64,3}; four_pi_G::Union{Nothing,Real}=nothing)
    four_pi_G === nothing && return poisson(p, ρ_baryon; scale=0.0)  # 0: the plan's coefficient
    iszero(four_pi_G) && return zeros(Float64, p.dims)
    return poisson(p, ρ_baryon; scale=four_pi_G)
end

"""
    potential(p::HIPPlan; Φ_b=nothing) -> (Φ_all, Φ_WaveDM)

Φ_WaveDM = 4πG ∇⁻²(|ψ|² - ⟨|ψ|²⟩) of the plan's ψ (src/solver/gravity.jl:13-23);
Φ_all adds the mesh baryon potential `Φ_b` when one is given (gravity.jl:24-28).
"""
function potential(p::HIPPlan; Φ_b::Union{Nothing,Array{Float64,3}}=nothing)
    Φ = Array{Float64,3}(undef, p.dims)
    check(ccall((:wdm_get_phi, libwdm), Cint, (Ptr{Cvoid}, Ptr{Cdouble}), p.h, Φ), p.h)
    return (Φ_b === nothing ? Φ : Φ .+ Φ_b), Φ
end

"ψ ← FFT[exp(-i dt_kick (Φ[ψ] + V(t))) ψ] on the device (the kick of src/solver/KDK.jl:37-77); the plan keeps the spectrum."
kick!(p::HIPPlan, dt_kick::Real, t::Real=0.0) =
    (check(ccall((:wdm_kick, libwdm), Cint, (Ptr{Cvoid}, Cdouble, Cdouble), p.h, dt_kick, t), p.h); p)

"ψ ← border .* IFFT[exp(-i k² dt/2) ψ_k] (the drift of src/solver/KDK.jl:88-101)."
drift!(p::HIPPlan, dt::Real) = (check(ccall((:wdm_drift, libwdm), Cint, (Ptr{Cvoid}, Cdouble), p.h, dt), p.h); p)

"""
    step!(p::HIPPlan, nsteps, dt; dt_kick=dt/2, t0=0.0, split=:reference)

`nsteps` iterations of the loop body of src/simulation_main.jl:536-551
(`split=:reference`: one kick of `dt_kick` then one drift of `dt`), or Strang
KDK (`split=:strang`), fused on the device.
"""
function step!(p::HIPPlan, nsteps::Integer, dt::Real; dt_kick::Real=dt / 2, t0::Real=0.0,
               split::Symbol=:reference)
    mode = split === :strang ? WDM_SPLIT_STRANG : WDM_SPLIT_REFERENCE
    check(ccall((:wdm_step, libwdm), Cint, (Ptr{Cvoid}, Cint, Cdouble, Cdouble, Cdouble, Cint),
                p.h, nsteps, dt, dt_kick, t0, mode), p.h)
    return p
end

"(mass, kinetic, w_self, v_ext, psi_max, phi_max) of the plan's ψ, FP64 (read-only device passes)."
function reduce_diagnostics(p::HIPPlan)
    out = zeros(Float64, WDM_RED_COUNT)
    check(ccall((:wdm_reduce, libwdm), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Cint), p.h, out, WDM_RED_COUNT), p.h)
    return (mass=out[1], kinetic=out[2], w_self=out[3], v_ext=out[4], psi_max=out[5], phi_max=out[6])
end

"""
    split_energy(r, dt, dt_kick; split=:reference)

The energy the loop's split conserves, from `reduce_diagnostics`: one kick of `dt_kick`
per drift of `dt` (src/simulation_main.jl:538-547) advances H = K + (dt_kick/dt)(W + V_ext);
the Strang split kicks for a whole dt per step, H = K + W + V_ext.
"""
split_energy(r, dt::Real, dt_kick::Real; split::Symbol=:reference) =
    r.kinetic + (split === :strang ? 1.0 : dt_kick / dt) * (r.w_self + r.v_ext)

"""
    step_properties(p::HIPPlan) -> NamedTuple

The loop's per-step diagnostics of the plan's ψ, evaluated on the device
(src/simulation_main.jl:550-558,641-656; src/visualization/runtime.jl:210-238),
under the Prop.csv column names of src/io/save.jl:46-67 (lengths in code
units: multiply by uL as the loop does).  `peak` is the 1-based
CartesianIndex of findmax(ρ).
"""
function step_properties(p::HIPPlan)
    o = zeros(Float64, WDM_DIAG_COUNT)
    check(ccall((:wdm_diagnostics, libwdm), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Cint), p.h, o, WDM_DIAG_COUNT), p.h)
    return (TotalMass=o[1], rho_max=o[2], peak=CartesianIndex(Int(o[3]) + 1, Int(o[4]) + 1, Int(o[5]) + 1),
            R=o[6], R1=o[7], R2=o[8], R3=o[9], R4=o[10], R5=o[11], R6=o[12], R7=o[13], R8=o[14], R9=o[15],
            MomentumX=o[16], MomentumY=o[17], MomentumZ=o[18], PE_abs=o[19], KE=o[20], QE=o[21], Virial=o[22])
end

"ax_all, ay_all, az_all = grad_central(-Δ..., Φ_all) of the plan's ψ (src/simulation_main.jl:652)."
function accelerations(p::HIPPlan)
    ax, ay, az = (Array{Float64}(undef, p.dims) for _ in 1:3)
    check(ccall((:wdm_acceleration, libwdm), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}),
                p.h, ax, ay, az), p.h)
    return ax, ay, az
end

"ψ = sqrt.(ρ) .* exp.(im*θ) with θ = solve_vector_equation(vx, vy, vz, Δ...) (src/physics/schrodinger.jl:8-30), on the device."
function set_ic_velocity!(p::HIPPlan, ρ::Array{Float64,3}, vx::Array{Float64,3}, vy::Array{Float64,3},
                          vz::Array{Float64,3}; remove_bulk::Bool=true)
    check(ccall((:wdm_ic_velocity, libwdm), Cint,
                (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cint),
                p.h, ρ, vx, vy, vz, remove_bulk), p.h)
    return p
end

"Opt-in mass projection for FP32 plans (no counterpart in the reference)."
set_mass_ledger!(p::HIPPlan, on::Bool) =
    (check(ccall((:wdm_set_mass_ledger, libwdm), Cint, (Ptr{Cvoid}, Cint), p.h, on), p.h); p)

"buffer_ψ2 += ρ on the device (src/simulation_main.jl:659-660)."
density_accumulate!(p::HIPPlan) = (check(ccall((:wdm_density_accumulate, libwdm), Cint, (Ptr{Cvoid},), p.h), p.h); p)

"(averaged_ψ2, average_N) of the accumulated steps (src/io/save.jl:79-81)."
function density_average(p::HIPPlan)
    n = Ref{Clonglong}(0)
    check(ccall((:wdm_density_average, libwdm), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Ref{Clonglong}), p.h, C_NULL, n), p.h)
    n[] == 0 && return (nothing, 0)
    avg = Array{Float64}(undef, p.dims)
    check(ccall((:wdm_density_average, libwdm), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Ref{Clonglong}), p.h, avg, n), p.h)
    return avg, Int(n[])
end

density_reset!(p::HIPPlan) = (check(ccall((:wdm_density_reset, libwdm), Cint, (Ptr{Cvoid},), p.h), p.h); p)

"(dt_phase, dt_grav, dt_SI) of src/physics/poisson.jl:35-50 on the plan's ψ."
function cfl_limits(p::HIPPlan; κ::Real=0.0)
    out = zeros(Float64, 3)
    check(ccall((:wdm_cfl, libwdm), Cint, (Ptr{Cvoid}, Cdouble, Ptr{Cdouble}), p.h, κ, out), p.h)
    return (out[1], out[2], out[3])
end

# --------------------------------------------------------------------------
# Differentiable forward run (the gradient form): the reference imports Zygote
# (src/WaveDM.jl:42) but defines no rule for the step; this rrule runs the
# device adjoint (wdm_adjoint_forward / wdm_adjoint_backward).

"""
    ssfm(p::HIPPlan, ψ0, θ, nsteps, dt; dt_kick=dt/2, ckpt_every=20) -> ψN

`nsteps` reference-split steps from `ψ0` under the Hernquist host
θ = (M, a, cx, cy, cz) plus the plan's self-gravity, with device checkpoints
every `ckpt_every` steps.  Differentiable in `ψ0` and `θ` through the rrule
below (Zygote's complex convention: the cotangent of ψ is ∂L/∂Re ψ + i ∂L/∂Im ψ).
"""
function ssfm(p::HIPPlan, ψ0::Array{ComplexF64,3}, θ::NTuple{5,Float64}, nsteps::Integer, dt::Real;
              dt_kick::Real=dt / 2, ckpt_every::Integer=20)
    check(ccall((:wdm_clear_potential_models, libwdm), Cint, (Ptr{Cvoid},), p.h), p.h)
    set_hernquist!(p; M=θ[1], a=θ[2], center=θ[3:5])
    set_psi!(p, ψ0)
    check(ccall((:wdm_adjoint_forward, libwdm), Cint, (Ptr{Cvoid}, Cint, Cdouble, Cdouble, Cdouble, Cint),
                p.h, nsteps, dt, dt_kick, 0.0, ckpt_every), p.h)
    p.tape += 1
    return get_psi(p)
end

function ChainRulesCore.rrule(::typeof(ssfm), p::HIPPlan, ψ0::Array{ComplexF64,3}, θ::NTuple{5,Float64},
                              nsteps::Integer, dt::Real; dt_kick::Real=dt / 2, ckpt_every::Integer=20)
    ψN = ssfm(p, ψ0, θ, nsteps, dt; dt_kick, ckpt_every)
    token = p.tape
    function ssfm_pullback(ΔψN)
        # the plan holds one tape: a later ssfm on the same plan replaced it
        p.tape == token || error("ssfm pullback: the plan was re-taped by a later ssfm call; " *
                                 "use one HIPPlan per differentiated call")
        g = Array{ComplexF64,3}(unthunk(ΔψN))
        g0 = similar(g)
        gp = zeros(Float64, WDM_ADJ_NPARAM)
        # the library also refuses (WDM_ERR_STATE) if the plan's models or ψ changed since
        check(ccall((:wdm_adjoint_backward, libwdm), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Cvoid}, Ptr{Cdouble}, Cint, Cuint),
                    p.h, g, WDM_C128, g0, gp, WDM_ADJ_NPARAM, Cuint(0)), p.h)
        return NoTangent(), NoTangent(), g0, Tangent{typeof(θ)}(gp[1:5]...), NoTangent(), NoTangent()
    end
    return ψN, ssfm_pullback
end

# methods of WaveDM's solver-step functions for a HIPDeviceConfig
include("WaveDMHIPBackend.jl")

end # module
