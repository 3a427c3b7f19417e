# WaveDMHIPBackend.jl -- the HIP path behind WaveDM.jl's own solver-step
# functions, at the reference signatures.
#
# WaveDM's loop (src/simulation_main.jl:536-551) calls
#
#   Φ_all, Φ_WaveDM, spec = apply_kick_step!(device_ψ, ψ, V, rho, rho_max, rho_max_id, grid,
#                                            config_gravity, config_tidal, config_device, dt_kick, i, t)
#   device_ψ = apply_drift_step!(spec, linear_phase, border, gpu, DeviceArray)
#   ψ = collect(device_ψ) |> DA
#
# and its Poisson callers (src/simulation_main.jl:244,304, src/ic/generation.jl:54-86,266,
# src/io/save.jl:83-85, src/solver/gravity.jl:23-26) call
# parallel_poisson(Δ, N, ρ, boundary, config_device).  This file adds methods of those
# WaveDM functions for a `HIPDeviceConfig` (and the `HIPField` handle the kick returns),
# so a run switches to libwavedm_hip.so by building its `config_device` as
# `HIPDeviceConfig()` instead of `DeviceConfig(; gpu, distributed)` (src/simulation_main.jl:197);
# the loop lines themselves are unchanged.  Nothing here is type piracy: every method
# has a HIP type in its signature.
#
# Reference interfaces mirrored (same argument meaning and return values):
#   apply_kick_step!                 src/solver/KDK.jl:37-77
#   apply_drift_step!                src/solver/KDK.jl:88-101
#   compute_gravitational_potential  src/solver/gravity.jl:7-35
#   parallel_poisson                 src/parallel_poisson.jl:36-48
#   compute_averaged_fields          src/io/save.jl:79-95
#
# Not executed in this repository (the image has no Julia); wavedm.jl_amd/kdk.py issues
# the same C-ABI call sequence and is what the tests run.
module WaveDMHIPBackend

using WaveDM
using WaveDM: SimulationGrid, GravityConfig, TidalFieldConfig
using ..WaveDMHIP
using ..WaveDMHIP: HIPPlan, libwdm, check, set_psi!, get_psi, set_static_potential!, set_border!

export HIPDeviceConfig, HIPField, HIPPotential

"""
    HIPDeviceConfig(; precision=:f64, hbar_over_m=1.0, four_pi_G=4π)

The `config_device` of a run on one MI355X.  Read like WaveDM's `DeviceConfig`
(`kind`, `gpu`, `DeviceArray`, `DA`, `distribute_array`, `nthreads`, `nworkers`):
`DeviceArray` and `DA` are `identity` (host arrays cross the boundary; the state lives in
the simulation plan, created at the first kick from the grid).  The reference's code
units are the defaults: ħ/m = 1 and the literal 4π of src/solver/gravity.jl:23.
"""
mutable struct HIPDeviceConfig
    precision::Symbol
    hbar_over_m::Float64
    four_pi_G::Float64
    plan::Union{Nothing,HIPPlan}               # the simulation state
    grid::Union{Nothing,SimulationGrid}        # its grid (set at the first kick)
    poisson_plans::Dict{Any,HIPPlan}           # parallel_poisson lattices
    border_id::UInt                            # objectid of the installed absorber
    drift_id::UInt                             # objectid of the decoded linear_phase
    drift_dt::Float64
end
HIPDeviceConfig(; precision::Symbol=:f64, hbar_over_m::Real=1.0, four_pi_G::Real=4π) =
    HIPDeviceConfig(precision, Float64(hbar_over_m), Float64(four_pi_G), nothing, nothing, Dict{Any,HIPPlan}(),
                    UInt(0), UInt(0), NaN)

function Base.getproperty(c::HIPDeviceConfig, s::Symbol)
    s === :kind && return :hip
    s === :gpu && return true
    (s === :DeviceArray || s === :DA || s === :distribute_array || s === :device_array ||
     s === :local_array) && return identity
    s === :nthreads && return Threads.nthreads()
    s === :nworkers && return 1
    return getfield(c, s)
end

"""
Handle on the simulation plan's state: what `apply_kick_step!` returns as `spec` (k-space)
and `apply_drift_step!` as `device_ψ` (real space).  `collect` copies ψ to the host.
"""
struct HIPField
    plan::HIPPlan
    config::HIPDeviceConfig
end
Base.collect(f::HIPField) = get_psi(f.plan)
Base.size(f::HIPField) = f.plan.dims

"""
    HIPPotential(; hernquist=nothing, quadrupole=nothing, κ=nothing, static=nothing)

A `V` for the kick that the device evaluates itself (no host `V.(xxx, yyy, zzz, ψ)` and no
per-kick upload): Hernquist `(M, a, center)`, a rotating quadrupole
`(T0, axis, period, center)` evaluated at the kick time, κ|ψ|², a static array.  Any
other `V(x, y, z, ψ)` is evaluated on the host each kick, as src/solver/KDK.jl:64 does.
"""
Base.@kwdef struct HIPPotential <: Function
    hernquist::Union{Nothing,NamedTuple} = nothing
    quadrupole::Union{Nothing,NamedTuple} = nothing
    κ::Union{Nothing,Float64} = nothing
    static::Union{Nothing,Array{Float64,3}} = nothing
end
(v::HIPPotential)(x, y, z, ψ) = error("HIPPotential is evaluated on the device")

function _sim_plan!(c::HIPDeviceConfig, grid::SimulationGrid)
    if c.plan === nothing
        c.plan = HIPPlan(grid.Nx, grid.Ny, grid.Nz, grid.Xmax, grid.Ymax, grid.Zmax;
                         precision=c.precision, hbar_over_m=c.hbar_over_m, four_pi_G=c.four_pi_G)
        c.grid = grid
    end
    return c.plan
end

function _install_potential!(p::HIPPlan, V, ψ, grid::SimulationGrid, Φ_b)
    check(ccall((:wdm_clear_potential_models, libwdm), Cint, (Ptr{Cvoid},), p.h), p.h)
    static = Φ_b === nothing ? nothing : Array{Float64,3}(real.(collect(Φ_b)))
    if V isa HIPPotential
        V.hernquist === nothing || WaveDMHIP.set_hernquist!(p; V.hernquist...)
        V.quadrupole === nothing || WaveDMHIP.set_quadrupole!(p; V.quadrupole...)
        V.κ === nothing || WaveDMHIP.set_kappa!(p, V.κ)
        V.static === nothing || (static = static === nothing ? V.static : static .+ V.static)
    else
        Vh = real.(V.(grid.xxx, grid.yyy, grid.zzz, ψ))
        if any(!iszero, Vh)
            static = static === nothing ? Array{Float64,3}(Vh) : static .+ Vh
        end
    end
    set_static_potential!(p, static)
    return nothing
end

_phi(p::HIPPlan) = (Φ = Array{Float64,3}(undef, p.dims);
                    check(ccall((:wdm_get_phi, libwdm), Cint, (Ptr{Cvoid}, Ptr{Cdouble}), p.h, Φ), p.h); Φ)

# ---------------------------------------------------------------- parallel_poisson
"""
    parallel_poisson(Δ, N, ρ, boundary, config::HIPDeviceConfig; gpu_force=false) -> Φ

∇²Φ = ρ - ⟨ρ⟩ with period `N .* Δ` per axis on the device (the `Periodic()` branch of
src/parallel_poisson.jl:36-68; one plan kept per lattice).  Returns a host complex array,
as the reference's solver does (imaginary part zero).
"""
function WaveDM.parallel_poisson(Δ, N, ρ, boundary, config::HIPDeviceConfig; gpu_force::Bool=false)
    boundary isa WaveDM.Periodic || throw(ArgumentError("the HIP Poisson solver is periodic only"))
    ρh = Array{Float64,3}(real.(collect(ρ)))
    half = ntuple(a -> 0.5 * N[a] * Δ[a], 3)
    key = (size(ρh), half)
    p = get!(config.poisson_plans, key) do
        HIPPlan(size(ρh)..., half...; precision=config.precision)
    end
    Φ = WaveDMHIP.poisson(p, ρh; scale=1.0)
    return complex.(Φ)
end

# ---------------------------------------------------------------- gravity
"""
    compute_gravitational_potential(device_ψ, config_gravity, grid, config_device::HIPDeviceConfig)

(Φ_all, Φ_WaveDM) of src/solver/gravity.jl:7-35 (periodic): Φ_WaveDM = 4πG∇⁻²(|ψ|² - ⟨|ψ|²⟩)
on the device, Φ_all = Φ_WaveDM + Φ_b unless `baryon_mode == :ignored`.  Host FP64 arrays.
"""
function WaveDM.compute_gravitational_potential(device_ψ, config_gravity::GravityConfig, grid::SimulationGrid,
                                               config_device::HIPDeviceConfig)
    config_gravity.boundary isa WaveDM.Periodic ||
        throw(ArgumentError("the HIP path has periodic gravity only (tree gravity stays on the host)"))
    Φ = if device_ψ isa HIPField
        _phi(device_ψ.plan)
    else
        ψh = collect(device_ψ)
        ρ = abs.(ψh) .^ 2
        real.(WaveDM.parallel_poisson(grid.Δ, [grid.Nx - 1, grid.Ny - 1, grid.Nz - 1], ρ, WaveDM.Periodic(),
                                      config_device)) .* config_device.four_pi_G
    end
    config_gravity.baryon_mode == :ignored && return Φ, Φ
    return Φ .+ real.(collect(config_gravity.Φ_b)), Φ
end

# ---------------------------------------------------------------- kick
"""
    apply_kick_step!(device_ψ, ψ, V, rho, rho_max, rho_max_id, grid, config_gravity, config_tidal,
                     config_device::HIPDeviceConfig, dt, i, t) -> (Φ_all, Φ_WaveDM, spec)

src/solver/KDK.jl:37-77 on the device.  `device_ψ` is a host array (first call: it is
uploaded into the simulation plan) or the `HIPField` the previous drift returned.  The
kick adds `V` (a `HIPPotential`, evaluated in-kernel at the kick time `t[i-1]`, or any
`V(x, y, z, ψ)` evaluated on the host ψ and uploaded) and Φ_b (baryons on the mesh) to the
self-gravity of ψ, rotates by exp(-i dt (Φ_all + V)) and transforms; `spec` is the plan in
k-space.  Φ_all and Φ_WaveDM are the host potentials of the state before the kick, as in
the reference.  The MW tidal branch (src/solver/KDK.jl:59-61) is not taken: in the
reference it does not change the potential the kick applies.
"""
function WaveDM.apply_kick_step!(device_ψ, ψ, V::F, rho, rho_max, rho_max_id::CartesianIndex{3},
                                 grid::SimulationGrid, config_gravity::GravityConfig,
                                 config_tidal::TidalFieldConfig, config_device::HIPDeviceConfig,
                                 dt::Real, i::Int, t::Vector{<:Real}) where {F<:Function}
    p = _sim_plan!(config_device, grid)
    device_ψ isa HIPField || set_psi!(p, Array{ComplexF64,3}(collect(device_ψ)))
    Φ_b = config_gravity.baryon_mode == :ignored ? nothing : config_gravity.Φ_b
    _install_potential!(p, V, ψ, grid, Φ_b)
    Φ_all, Φ_WaveDM = WaveDM.compute_gravitational_potential(HIPField(p, config_device), config_gravity, grid,
                                                             config_device)
    check(ccall((:wdm_kick, libwdm), Cint, (Ptr{Cvoid}, Cdouble, Cdouble), p.h, dt, t[i-1]), p.h)
    return Φ_all, Φ_WaveDM, HIPField(p, config_device)
end

# ---------------------------------------------------------------- drift
# dt of a linear_phase = fftshift(exp(i Lap dt/2)) (src/solver/fft_operators.jl:8-15): the
# first mode along an axis of length > 1 has Lap = -(2π/(2X))^2 and a phase far below π,
# so dt = -2 arg(linear_phase[mode]) / (2π/(2X))^2 to FP64 accuracy.
function _decode_dt(lp::AbstractArray, grid::SimulationGrid)
    for (a, n, X) in ((1, grid.Nx, grid.Xmax), (2, grid.Ny, grid.Ymax), (3, grid.Nz, grid.Zmax))
        n > 1 || continue
        idx = ntuple(b -> b == a ? 2 : 1, 3)
        k = 2π / (2X)
        return -2 * angle(lp[idx...]) / k^2
    end
    return 0.0
end

"""
    apply_drift_step!(spec::HIPField, linear_phase, border, gpu, DeviceArray) -> device_ψ

src/solver/KDK.jl:88-101 on the device: the k-space factor exp(-i k² dt/2) is evaluated
in-kernel from the dt that `linear_phase` encodes (decoded once per array), and the
absorber `border` (the reference's length-Nx vector) is applied along x in the inverse x
pass (uploaded once per array; all ones means none).
"""
function WaveDM.apply_drift_step!(spec::HIPField, linear_phase, border, gpu::Bool, DeviceArray)
    p, c = spec.plan, spec.config
    if c.drift_id != objectid(linear_phase)
        c.drift_dt = _decode_dt(linear_phase, c.grid)
        c.drift_id = objectid(linear_phase)
    end
    if c.border_id != objectid(border)
        b = collect(border)
        set_border!(p, all(isone, b) ? nothing : vec(Float64.(b)))
        c.border_id = objectid(border)
    end
    check(ccall((:wdm_drift, libwdm), Cint, (Ptr{Cvoid}, Cdouble), p.h, c.drift_dt), p.h)
    return HIPField(p, c)
end

# ---------------------------------------------------------------- averages
"""
    compute_averaged_fields(average, buffer_ψ2, average_N, baryon_mode, a_all, Φ_b, Δ, Nx, Ny, Nz,
                            config_device::HIPDeviceConfig)

src/io/save.jl:79-95 with the averaged potential solved on the device.
"""
function WaveDM.compute_averaged_fields(average, buffer_ψ2, average_N, baryon_mode, a_all, Φ_b, Δ, Nx, Ny, Nz,
                                        config_device::HIPDeviceConfig)
    average || return [], a_all .* 0
    averaged_ψ2 = buffer_ψ2 / average_N
    Φ = config_device.four_pi_G .* real.(WaveDM.parallel_poisson(Δ, [Nx - 1, Ny - 1, Nz - 1], averaged_ψ2,
                                                                 WaveDM.Periodic(), config_device))
    baryon_mode == :ignored || (Φ = Φ .+ real.(collect(Φ_b)))
    ax, ay, az = WaveDM.grad_central(-Δ..., Φ)
    return averaged_ψ2, sqrt.(ax[:, :, div(end, 2)] .^ 2 .+ ay[:, :, div(end, 2)] .^ 2 .+ az[:, :, div(end, 2)] .^ 2)
end

end # module
