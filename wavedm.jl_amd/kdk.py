"""Host-side mirror of WaveDM.jl's solver-step API over the HIP C ABI.

Same names, argument meaning and error behaviour as the reference functions
(Julia `!` dropped); the arithmetic runs in libwavedm_hip.so.  Geometry and
time-step helpers are plain host logic, as in the reference.

  setup_grid                       src/physics/poisson.jl:8-15
  setup_coordinates                src/physics/poisson.jl:22-28
  compute_timestep                 src/physics/poisson.jl:35-50
  setup_fft_operators              src/solver/fft_operators.jl:8-15
  setup_absorption_boundary        src/solver/fft_operators.jl:22-27
  parallel_poisson                 src/parallel_poisson.jl:36-57
  baryon_mesh_potential            src/simulation_main.jl:244 (baryon_mode = :mesh)
  compute_gravitational_potential  src/solver/gravity.jl:7-35
  apply_kick_step                  src/solver/KDK.jl:37-77
  apply_drift_step                 src/solver/KDK.jl:88-101
  run_loop                         src/simulation_main.jl:536-551 (the fast path is Plan.step)
  split_energy                     the energy the loop's split conserves (Q1 of the loop)
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .plan import Plan


# ---------------------------------------------------------------- geometry
def _linrange(a, b, n):
    if n == 1:
        return np.array([float(a)])
    t = np.arange(n, dtype=np.float64) / float(n - 1)
    return (1.0 - t) * float(a) + t * float(b)


def setup_grid(Xmax, Ymax, Zmax, Nx, Ny, Nz, consistent=False):
    """x = LinRange(-Xmax, Xmax, Nx) etc.; Delta = x[2]-x[1]; cell volume prod(Delta)."""
    if consistent:
        axes = [-X + np.arange(N) * (2.0 * X / N) for X, N in ((Xmax, Nx), (Ymax, Ny), (Zmax, Nz))]
        d = np.array([2.0 * Xmax / Nx, 2.0 * Ymax / Ny, 2.0 * Zmax / Nz])
    else:
        axes = [_linrange(-X, X, N) for X, N in ((Xmax, Nx), (Ymax, Ny), (Zmax, Nz))]
        d = np.array([a[1] - a[0] if len(a) > 1 else 2.0 * X
                      for a, X in zip(axes, (Xmax, Ymax, Zmax))])
    return axes[0], axes[1], axes[2], d, float(np.prod(d))


def setup_coordinates(x, y, z):
    """3-D coordinate arrays (xxx[:,0,0]==x, ...) and r; only needed for host V callbacks."""
    xxx = np.broadcast_to(np.asarray(x)[:, None, None], (len(x), len(y), len(z))).copy(order="F")
    yyy = np.broadcast_to(np.asarray(y)[None, :, None], (len(x), len(y), len(z))).copy(order="F")
    zzz = np.broadcast_to(np.asarray(z)[None, None, :], (len(x), len(y), len(z))).copy(order="F")
    return xxx, yyy, zzz, np.sqrt(xxx ** 2 + yyy ** 2 + zzz ** 2)


def compute_timestep(dx, Phi_all, kappa, psi, Tmax, Nt, autoset_timestep, autoset_timestep_ratio,
                     hbar_over_m=1.0):
    """Reference CFL: dt_phase = 2 dx^2/pi, dt_grav = 2 pi/max|Phi|,
    dt_SI = 2 pi/(|kappa| max|psi|) (hbar/m made explicit).  ``Phi_all`` and
    ``psi`` may be host arrays, or ``psi`` a Plan (FP64 device reductions)."""
    if isinstance(psi, Plan):
        lims = psi.cfl(kappa)
    else:
        mh = 1.0 / hbar_over_m
        pmax = float(np.max(np.abs(Phi_all)))
        smax = float(np.max(np.abs(psi)))
        lims = (2.0 * dx * dx / math.pi / hbar_over_m,
                2.0 * math.pi / (mh * pmax) if pmax > 0 else math.inf,
                2.0 * math.pi / abs(kappa) / mh / smax if kappa != 0 and smax > 0 else math.inf)
    if autoset_timestep:
        Nt = int(math.ceil(Tmax / (autoset_timestep_ratio * min(lims))))
    t = _linrange(0.0, Tmax, Nt)
    dt = float(np.float32(Tmax / Nt))
    return t, dt, Nt


@dataclass
class DriftOperator:
    """Stands for `linear_phase = fftshift(exp(i Lap dt/2))`: the device
    evaluates the factor per mode from dt, nothing N-sized is built or uploaded."""
    Xmax: float
    Ymax: float
    Zmax: float
    Nx: int
    Ny: int
    Nz: int
    dt: float


def setup_fft_operators(Xmax, Ymax, Zmax, Nx, Ny, Nz, dt):
    return DriftOperator(Xmax, Ymax, Zmax, Nx, Ny, Nz, float(dt))


def setup_absorption_boundary(Xmax, Ymax, Zmax, x, y, z, absorb_coeff, dt):
    """1-D mask broadcast along x, as the reference builds it from 1-D x, y, z."""
    wx, wy, wz = Xmax / 50.0, Ymax / 50.0, Zmax / 50.0
    x, y, z = np.asarray(x), np.asarray(y), np.asarray(z)
    s = (6.0 - np.tanh((x + Xmax) / wx) + np.tanh((x - Xmax) / wx)
         - np.tanh((y + Ymax) / wy) + np.tanh((y - Ymax) / wy)
         - np.tanh((z + Zmax) / wz) + np.tanh((z - Zmax) / wz))
    return np.exp(-absorb_coeff * s * dt)


# ---------------------------------------------------------------- configs
@dataclass
class DeviceConfig:
    """`DeviceConfig` with the one backend this package has: the HIP device."""
    precision: str = "f64"
    kind: str = "hip"
    gpu: bool = True


@dataclass
class GravityConfig:
    """Periodic self-gravity (src/core/configs.jl:161-170) with baryon_mode
    :ignored, or :mesh with a fixed baryon potential ``Phi_b`` added to
    Phi_WaveDM in every kick (src/solver/gravity.jl:24-28).  ``Phi_b`` is the
    real field from ``baryon_mesh_potential`` (src/simulation_main.jl:244); a
    complex one (the reference's Poisson output) contributes its real part."""
    boundary: str = "periodic"
    baryon_mode: str = "ignored"
    four_pi_G: float = 4.0 * math.pi
    Phi_b: np.ndarray | None = None

    def __post_init__(self):
        mode = str(self.baryon_mode).lstrip(":")
        if mode not in ("ignored", "mesh"):
            raise ValueError(f"Unsupported baryon_mode {self.baryon_mode}: the HIP path takes :ignored or :mesh")
        self.baryon_mode = mode
        if mode == "mesh":
            if self.Phi_b is None:
                raise ValueError("baryon_mode :mesh needs the baryon potential Phi_b")
            phi_b = np.asarray(self.Phi_b)
            if phi_b.ndim != 3:
                raise ValueError("Phi_b must be a 3-D field on the simulation grid")
            self.Phi_b = np.asarray(phi_b.real, dtype=np.float64)

    def baryon_potential(self):
        """Phi_b when baryons are on the mesh, else None."""
        return self.Phi_b if self.baryon_mode == "mesh" else None


@dataclass
class SimulationGrid:
    Xmax: float
    Ymax: float
    Zmax: float
    Nx: int
    Ny: int
    Nz: int
    consistent: bool = False
    x: np.ndarray = field(init=False)
    y: np.ndarray = field(init=False)
    z: np.ndarray = field(init=False)
    delta: np.ndarray = field(init=False)
    unit_cell_volumn: float = field(init=False)

    def __post_init__(self):
        self.x, self.y, self.z, self.delta, self.unit_cell_volumn = setup_grid(
            self.Xmax, self.Ymax, self.Zmax, self.Nx, self.Ny, self.Nz, self.consistent)


# ---------------------------------------------------------------- potentials
@dataclass
class Hernquist:
    M: float
    a: float
    center: tuple = (0.0, 0.0, 0.0)


@dataclass
class RotatingQuadrupole:
    T0: np.ndarray
    axis: tuple
    period: float
    center: tuple = (0.0, 0.0, 0.0)


@dataclass
class Kappa:
    kappa: float


def _install_potential(plan: Plan, V, t, phi_b=None):
    """Device models are installed once; a host callable is evaluated each
    kick on the host psi (the reference's V.(xxx,yyy,zzz,psi), KDK.jl:64).
    The static field the kick adds is the sum of the array and callable
    items and the baryon potential ``phi_b``.  V = None leaves the plan's
    device models as they are."""
    static = phi_b
    if V is not None:
        items = V if isinstance(V, (list, tuple)) else [V]
        plan.clear_potential_models()
        for v in items:
            if v is None:
                continue
            if isinstance(v, Hernquist):
                plan.set_hernquist(v.M, v.a, v.center)
                continue
            if isinstance(v, RotatingQuadrupole):
                plan.set_quadrupole(v.T0, v.axis, v.period, v.center)
                continue
            if isinstance(v, Kappa):
                plan.set_kappa(v.kappa)
                continue
            if callable(v):
                x, y, z, _, _ = setup_grid(*plan.xmax, *plan.shape, consistent=plan.consistent_grid)
                xxx, yyy, zzz, _ = setup_coordinates(x, y, z)
                field_v = np.asarray(v(xxx, yyy, zzz, plan.get_psi())).real
            else:
                field_v = np.asarray(v).real
            field_v = np.asarray(field_v, dtype=np.float64)
            static = field_v if static is None else static + field_v
    plan.set_static_potential(static)


# ---------------------------------------------------------------- solver API
# parallel_poisson keeps one plan per (shape, box, precision), as the
# reference's callers reuse one FFT plan per grid; clear_poisson_plans frees them.
_poisson_plans: dict = {}


def clear_poisson_plans():
    """Destroy the plans parallel_poisson keeps (their device buffers are freed)."""
    while _poisson_plans:
        _, plan = _poisson_plans.popitem()
        plan.close()


def parallel_poisson(delta, N, rho, boundary="periodic", backend=None, gpu_force=False):
    """Solve lap(Phi) = rho - <rho> periodically with period N_i * Delta_i per
    axis (src/parallel_poisson.jl:36-57 -> PhysicalFFT.fft_poisson).  Returns a
    complex array like the reference (imaginary part zero)."""
    if str(boundary).lower().strip(":") not in ("periodic", "periodic()"):
        raise NotImplementedError("only periodic boundaries are on the HIP path")
    rho = np.asarray(rho, dtype=np.float64)
    shape = rho.shape
    xmax = tuple(0.5 * float(N[i]) * float(delta[i]) for i in range(3))
    prec = getattr(backend, "precision", "f64") if backend is not None else "f64"
    key = (shape, xmax, prec)
    plan = _poisson_plans.get(key)
    if plan is None:
        plan = Plan(shape, xmax, prec)
        _poisson_plans[key] = plan
    phi = plan.poisson(rho, scale=1.0)
    return phi.astype(np.complex128)


def baryon_mesh_potential(plan: Plan, rho_baryon, four_pi_G=None):
    """Phi_b = 4 pi G * parallel_poisson(Delta, [Nx-1, Ny-1, Nz-1], rho_baryon)
    for baryon_mode = :mesh (src/simulation_main.jl:244, src/ic/generation.jl:54),
    solved on ``plan``'s device lattice (period 2 Xmax per axis, as the
    reference's N Delta) in the plan's precision; real part, host FP64.
    ``four_pi_G`` None takes the plan's coefficient (the reference's literal 4 pi
    in code units); an explicit 0 gives Phi_b = 0 (no solve).  The plan's psi
    is left as it was."""
    rho = np.asarray(rho_baryon, dtype=np.float64)
    if rho.shape != plan.shape:
        raise ValueError(f"rho_baryon must have shape {plan.shape}, got {rho.shape}")
    scale = plan.four_pi_G if four_pi_G is None else float(four_pi_G)
    if scale == 0.0:
        # the C ABI reads scale 0 as "the plan's coefficient": a zero
        # coefficient is answered here
        return np.zeros(plan.shape, dtype=np.float64)
    return plan.poisson(rho, scale=scale)


def compute_gravitational_potential(device_psi: Plan, config_gravity=None, grid=None, config_device=None):
    """rho = |psi|^2; rho -= mean(rho); Phi = 4 pi G lap^-1 rho (gravity.jl:13-23).
    Returns host (Phi_all, Phi_WaveDM): the same array with baryons ignored,
    Phi_all = Phi_WaveDM + Phi_b with baryons on the mesh (gravity.jl:24-28)."""
    phi = device_psi.get_phi()
    phi_b = config_gravity.baryon_potential() if config_gravity is not None else None
    if phi_b is not None:
        return phi + phi_b, phi
    return phi, phi


def apply_kick_step(device_psi: Plan, psi=None, V=None, rho=None, rho_max=None, rho_max_id=None, grid=None,
                    config_gravity=None, config_tidal=None, config_device=None, dt=0.0, i=0, t=None,
                    want_potential=True):
    """psi <- FFT[exp(-i dt (Phi + V)) psi].  Returns (Phi_all, Phi_WaveDM, spec)
    where spec is the same device state now holding the spectrum (the
    reference consumes device_psi and returns a new spectrum array)."""
    tk = 0.0
    if t is not None:
        tk = float(t[i - 1]) if hasattr(t, "__len__") else float(t)
    phi_b = config_gravity.baryon_potential() if config_gravity is not None else None
    if V is not None or phi_b is not None:
        _install_potential(device_psi, V, tk, phi_b)
    phi_all = phi_wdm = None
    if want_potential:
        phi_all, phi_wdm = compute_gravitational_potential(device_psi, config_gravity, grid, config_device)
    device_psi.kick(dt, tk)
    return phi_all, phi_wdm, device_psi


def apply_drift_step(spec: Plan, linear_phase: DriftOperator, border=None, gpu=True, DeviceArray=None):
    """spec .*= linear_phase; psi = ifft(spec); psi = border .* psi."""
    spec.set_border(border if border is not None and not np.all(np.asarray(border) == 1.0) else None)
    spec.drift(linear_phase.dt)
    return spec


def split_energy(red, dt, dt_kick, split="reference"):
    """The energy the loop's split conserves, from a ``Plan.reduce()`` dict.

    One reference iteration is K(dt_kick) D(dt) (src/simulation_main.jl:536-547,
    one kick of dt_kick = dt/2 when KDK_flag): to first order in dt it advances
    the flow of H = K + (dt_kick/dt) (W + V_ext), i.e. Schroedinger-Poisson with
    both the self-gravity and the external potential scaled by dt_kick/dt.
    That H (kinetic + scaled self-gravity W = 1/2 sum rho Phi dV + scaled
    sum rho V dV) is what a static potential keeps constant up to the
    splitting error; K + W + V_ext is not (it moves at O(1) when
    dt_kick != dt).  The Strang split K(dt/2) D(dt) K(dt/2) kicks for a whole
    dt per step, so there H = K + W + V_ext."""
    s = 1.0 if split == "strang" else float(dt_kick) / float(dt)
    return red["kinetic"] + s * (red["w_self"] + red["v_ext"])


def run_loop(plan: Plan, Nt, dt, *, KDK_flag=True, t0=0.0, snapshot_every=0, on_snapshot=None,
             split="reference", config_gravity=None):
    """The simulation loop (Nt-1 iterations of K(dt_kick) D(dt), Q1/Q2) on the fused
    device path, copying psi to the host only at snapshot cadence.  With baryons
    on the mesh, Phi_b is uploaded once as the plan's static potential (it
    replaces a static field set before); device potential models stay."""
    phi_b = config_gravity.baryon_potential() if config_gravity is not None else None
    if phi_b is not None:
        plan.set_static_potential(phi_b)
    dt_kick = 0.5 * dt if KDK_flag else dt
    nsteps = Nt - 1
    done = 0
    chunk = snapshot_every if snapshot_every > 0 else nsteps
    while done < nsteps:
        n = min(chunk, nsteps - done)
        plan.step(n, dt, dt_kick, t0 + done * dt, split)
        done += n
        if on_snapshot is not None:
            on_snapshot(done, plan)
    return plan
