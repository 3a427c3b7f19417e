"""FP64 CPU oracle for the WaveDM.jl split-step Fourier (SSFM) kick/drift step.

TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library imports this module;
only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
use it, always to check results (or as the timed CPU restatement), never as the
thing measured on the GPU.

This is a restatement, op for op, of the reference's CPU path (FFTW replaced by
numpy's pocketfft, or scipy.fft with worker threads for the timed baseline).
Arrays use the reference's index semantics: shape ``(Nx, Ny, Nz)``, Fortran
order, so memory is x-fastest exactly like a Julia ``Array{ComplexF64,3}``.

Functions and the reference lines they restate:

* ``linrange``                       Julia ``LinRange`` point values (``lerpi``)
* ``setup_grid``                     src/physics/poisson.jl:8-15
* ``setup_coordinates``              src/physics/poisson.jl:22-28
* ``compute_timestep``               src/physics/poisson.jl:35-50
* ``setup_fft_operators``            src/solver/fft_operators.jl:8-15
* ``setup_absorption_boundary``      src/solver/fft_operators.jl:22-27
* ``fft_poisson`` / ``parallel_poisson``
                                     src/parallel_poisson.jl:36-57 -> PhysicalFFT.fft_poisson
                                     (external, absent: convention restated, see below)
* ``compute_gravitational_potential`` src/solver/gravity.jl:7-35 (periodic, :ignored baryons)
* ``apply_kick_step``                src/solver/KDK.jl:37-77
* ``apply_drift_step``               src/solver/KDK.jl:88-101
* ``run_ssfm``                       loop body src/simulation_main.jl:536-551
* ``setup_initial_conditions``       src/physics/schrodinger.jl:37-46

Physical constants: the reference works in code units where hbar/m = 1 and the
Poisson source is 4*pi*rho (gravity.jl:23).  Every function here takes
``hbar_over_m`` and ``four_pi_G`` so the same restatement serves the
kpc-Msun-Gyr configurations (reference code units are hbar_over_m=1,
four_pi_G=4*pi).

Pinning status (checked in tests/test_oracle.py):
* FFT convention -> brute-force DFT matrices (closed form).
* drift factor -> discrete plane-wave eigenmodes acquire exp(-i hbar/m k^2 dt/2).
* Poisson -> single Fourier modes, k=0, linearity, reference test_parallel sanity.
* kick -> |psi| preserved exactly, phase = -(m/hbar) dt (Phi+V).
* whole step -> mass conservation, exact reversibility, lattice symmetry,
  3-D harmonic-oscillator ground state (consistent-grid mode).
* PhysicalFFT.fft_poisson's exact k-period convention is parity unpinned: it is
  not in the reference tree; the restatement uses period N_arg*Delta, with
  N_arg = Nx-1 as passed by gravity.jl:23 (so period = 2*Xmax).
"""
from __future__ import annotations

import math
import os

import numpy as np

__all__ = [
    "linrange", "setup_grid", "setup_coordinates", "compute_timestep",
    "setup_fft_operators", "setup_absorption_boundary", "fft_poisson",
    "parallel_poisson", "compute_gravitational_potential", "apply_kick_step",
    "apply_drift_step", "run_ssfm", "setup_initial_conditions", "fft_order_int",
    "Backend", "mass", "energies",
]


class Backend:
    """FFT provider.  ``numpy`` (default, single thread) or ``scipy`` with
    ``workers`` threads (used only for the timed CPU baseline)."""

    def __init__(self, kind: str = "numpy", workers: int | None = None):
        self.kind = kind
        self.workers = workers or os.cpu_count() or 1
        if kind == "scipy":
            import scipy.fft as sfft  # noqa: F401
            self._sfft = sfft

    def fftn(self, a):
        if self.kind == "scipy":
            return self._sfft.fftn(a, workers=self.workers)
        return np.fft.fftn(a)

    def ifftn(self, a):
        if self.kind == "scipy":
            return self._sfft.ifftn(a, workers=self.workers)
        return np.fft.ifftn(a)


_NUMPY = Backend("numpy")


def linrange(start: float, stop: float, n: int) -> np.ndarray:
    """Julia ``LinRange(start, stop, n)`` point values.

    Julia evaluates element j (0-based) as ``(1-t)*start + t*stop`` with
    ``t = j/(n-1)`` (Base ``lerpi``); restated here so the grid matches the
    reference to the last ulp as far as IEEE arithmetic allows."""
    if n == 1:
        return np.array([float(start)])
    j = np.arange(n, dtype=np.float64)
    t = j / float(n - 1)
    return (1.0 - t) * float(start) + t * float(stop)


def setup_grid(Xmax, Ymax, Zmax, Nx, Ny, Nz, consistent=False):
    """src/physics/poisson.jl:8-15.  ``consistent=True`` is the build's optional
    grid where Delta = 2*Xmax/N (period equals N*Delta); the reference uses
    LinRange(-Xmax, Xmax, N), i.e. Delta = 2*Xmax/(N-1) with both endpoints."""
    if consistent:
        x = -Xmax + np.arange(Nx) * (2.0 * Xmax / Nx)
        y = -Ymax + np.arange(Ny) * (2.0 * Ymax / Ny)
        z = -Zmax + np.arange(Nz) * (2.0 * Zmax / Nz)
        d = np.array([2.0 * Xmax / Nx, 2.0 * Ymax / Ny, 2.0 * Zmax / Nz])
    else:
        x = linrange(-Xmax, Xmax, Nx)
        y = linrange(-Ymax, Ymax, Ny)
        z = linrange(-Zmax, Zmax, Nz)
        d = np.array([x[1] - x[0] if Nx > 1 else 2.0 * Xmax,
                      y[1] - y[0] if Ny > 1 else 2.0 * Ymax,
                      z[1] - z[0] if Nz > 1 else 2.0 * Zmax])
    return x, y, z, d, float(np.prod(d))


def setup_coordinates(x, y, z):
    """src/physics/poisson.jl:22-28 (xxx[:,1,1]==x, yyy[1,:,1]==y, zzz[1,1,:]==z)."""
    Nx, Ny, Nz = len(x), len(y), len(z)
    one = np.ones((Nx, Ny, Nz), order="F")
    xxx = one * x[:, None, None]
    yyy = one * y[None, :, None]
    zzz = one * z[None, None, :]
    r = np.sqrt(xxx ** 2 + yyy ** 2 + zzz ** 2)
    return (np.asfortranarray(xxx), np.asfortranarray(yyy),
            np.asfortranarray(zzz), np.asfortranarray(r))


def compute_timestep(dx, phi_all, kappa, psi, Tmax, Nt, autoset, ratio,
                     hbar_over_m=1.0):
    """src/physics/poisson.jl:35-50, with hbar/m made explicit (=1 in code units).

    dt_phase = 2 dx^2/(pi hbar/m); dt_grav = 2 pi/((m/hbar) max|Phi|);
    dt_SI = 2 pi/((m/hbar)|kappa| max|psi|) (the reference uses max|psi|, not
    |psi|^2).  dt is rounded through Float32 as the reference does."""
    m_over_hbar = 1.0 / hbar_over_m
    dt_phase = 2.0 * dx * dx / math.pi / hbar_over_m
    pmax = float(np.max(np.abs(phi_all)))
    dt_grav = 2.0 * math.pi / (m_over_hbar * pmax) if pmax > 0 else math.inf
    psimax = float(np.max(np.abs(psi)))
    if kappa == 0 or psimax == 0:
        dt_si = math.inf
    else:
        dt_si = 2.0 * math.pi / abs(kappa) / m_over_hbar / psimax
    if autoset:
        lim = min(dt_phase, dt_grav, dt_si)
        Nt = int(math.ceil(Tmax / (ratio * lim)))
    t = linrange(0.0, Tmax, Nt)
    dt = float(np.float32(Tmax / Nt))
    return t, dt, Nt, (dt_phase, dt_grav, dt_si)


def fft_order_int(n: int) -> np.ndarray:
    """Signed integer frequency index in FFT order: 0,1,..,ceil(n/2)-1,-floor(n/2),..,-1."""
    return np.fft.fftfreq(n, d=1.0 / n).round().astype(np.int64)


def _shifted_freq_1d(n: int, Xmax: float) -> np.ndarray:
    """fft_operators.jl:9-11 then the fftshift of :13, for one axis.

    f_j = LinRange(-n/(4X), n/(4X) - 1/(2X), n)[j]; fftshift(A)[m] =
    A[(m - n div 2) mod n] (Julia circshift semantics)."""
    if n == 1:
        # build extension for embedded 1-D/2-D runs: a length-1 axis carries k=0
        # only (the reference is 3-D; its formula would give f = -1/(4X) here)
        return np.zeros(1)
    f = linrange(-n / 4.0 / Xmax, n / 4.0 / Xmax - 1.0 / 2.0 / Xmax, n)
    idx = (np.arange(n) - n // 2) % n
    return f[idx]


def setup_fft_operators(Xmax, Ymax, Zmax, Nx, Ny, Nz, dt, hbar_over_m=1.0):
    """src/solver/fft_operators.jl:8-15: fftshift(exp(i * Lap * dt/2)) with
    Lap = (2 pi i kx)^2 + (2 pi i ky)^2 + (2 pi i kz)^2; hbar/m multiplies Lap."""
    fx = _shifted_freq_1d(Nx, Xmax)
    fy = _shifted_freq_1d(Ny, Ymax)
    fz = _shifted_freq_1d(Nz, Zmax)
    lap = ((2j * np.pi * fx[:, None, None]) ** 2 + (2j * np.pi * fy[None, :, None]) ** 2
           + (2j * np.pi * fz[None, None, :]) ** 2)
    return np.asfortranarray(np.exp(1j * hbar_over_m * lap * dt / 2.0))


def setup_absorption_boundary(Xmax, Ymax, Zmax, x, y, z, absorb_coeff, dt):
    """src/solver/fft_operators.jl:22-27 called with 1-D x,y,z as the simulation loop does
    (simulation_main.jl:415): the result is a length-Nx vector whose element i
    mixes x[i], y[i], z[i] and is broadcast along dim 1 only."""
    wx, wy, wz = Xmax / 50.0, Ymax / 50.0, Zmax / 50.0
    x, y, z = np.asarray(x), np.asarray(y), np.asarray(z)
    s = (6.0 - np.tanh((x + Xmax) / wx) + np.tanh((x - Xmax) / wx)
         - np.tanh((y + Ymax) / wy) + np.tanh((y - Ymax) / wy)
         - np.tanh((z + Zmax) / wz) + np.tanh((z - Zmax) / wz))
    return np.exp(-absorb_coeff * s * dt)


def fft_poisson(delta, N, rho, be: Backend = _NUMPY):
    """Restatement of PhysicalFFT.fft_poisson(Delta, N, rho, Periodic(), CPU())
    as called by _poisson_local (src/parallel_poisson.jl:54-57).

    Solves lap(u) = rho periodically: u_k = -rho_k/|k|^2 with k=0 -> 0 and
    k_i = 2 pi m_i / (N_i Delta_i), m_i the FFT-order integer.  Returns the
    complex inverse FFT (the reference output is complex, test_parallel.jl:90).
    Convention parity unpinned (PhysicalFFT is not in the reference tree)."""
    rho = np.asarray(rho)
    nx, ny, nz = rho.shape
    kx = 2.0 * np.pi * fft_order_int(nx) / (N[0] * delta[0])
    ky = 2.0 * np.pi * fft_order_int(ny) / (N[1] * delta[1])
    kz = 2.0 * np.pi * fft_order_int(nz) / (N[2] * delta[2])
    k2 = kx[:, None, None] ** 2 + ky[None, :, None] ** 2 + kz[None, None, :] ** 2
    with np.errstate(divide="ignore", invalid="ignore"):
        g = np.where(k2 == 0.0, 0.0, -1.0 / k2)
    return np.asfortranarray(be.ifftn(be.fftn(rho) * g))


def parallel_poisson(delta, N, rho, be: Backend = _NUMPY):
    """src/parallel_poisson.jl:36-57 (serial/threads branch)."""
    return fft_poisson(delta, N, rho, be)


def compute_gravitational_potential(psi, delta, four_pi_G=4.0 * np.pi,
                                    be: Backend = _NUMPY, poisson_N=None):
    """src/solver/gravity.jl:13-14,22-23 (periodic, baryon_mode=:ignored).

    rho = |psi|^2; rho .-= mean(rho); Phi = 4 pi * parallel_poisson(Delta,
    [Nx-1,Ny-1,Nz-1], rho); four_pi_G replaces the literal 4 pi."""
    rho = np.abs(psi) ** 2
    rho = rho - rho.mean()
    nx, ny, nz = psi.shape
    if poisson_N is None:
        poisson_N = (max(nx - 1, 1), max(ny - 1, 1), max(nz - 1, 1))
    return four_pi_G * parallel_poisson(delta, poisson_N, rho, be)


def apply_kick_step(psi, phi_all, V_field, dt_kick, hbar_over_m=1.0,
                    be: Backend = _NUMPY):
    """src/solver/KDK.jl:63-73: potential = V + Phi; nonlinear =
    exp(-i dt (m/hbar) potential) .* psi; spec = fft(nonlinear)."""
    pot = phi_all if V_field is None else phi_all + V_field
    nonlinear = np.exp(-1j * dt_kick * pot / hbar_over_m) * psi
    return be.fftn(nonlinear)


def apply_drift_step(spec, linear_phase, border, be: Backend = _NUMPY):
    """src/solver/KDK.jl:89-96: spec .*= linear_phase; newpsi = ifft(spec);
    psi = border .* newpsi (border is a length-Nx vector broadcast along x)."""
    spec = spec * linear_phase
    newpsi = be.ifftn(spec)
    if border is None:
        return np.asfortranarray(newpsi)
    b = np.asarray(border)
    if b.ndim == 1:
        return np.asfortranarray(b[:, None, None] * newpsi)
    return np.asfortranarray(b * newpsi)


def setup_initial_conditions(IC, xxx, yyy, zzz):
    """src/physics/schrodinger.jl:37-46."""
    if callable(IC):
        psi = np.asarray(IC(xxx, yyy, zzz), dtype=np.complex128)
    else:
        psi = np.asarray(IC, dtype=np.complex128)
    psi = np.asfortranarray(psi)
    sqrt_rho = np.abs(psi)
    return psi, sqrt_rho, sqrt_rho ** 2


SPLIT_REFERENCE = 0
SPLIT_STRANG = 1


def run_ssfm(psi0, Xmax, nsteps, dt, dt_kick=None, *, hbar_over_m=1.0,
             four_pi_G=4.0 * np.pi, V=None, border=None, t0=0.0,
             self_gravity=True, consistent_grid=False, split=SPLIT_REFERENCE,
             be: Backend = _NUMPY, poisson_N=None, callback=None):
    """Loop body of src/simulation_main.jl:536-551.

    Reference split (Q1): each iteration kicks once with dt_kick (0.5*dt when
    KDK_flag) and drifts with the full dt.  ``split=SPLIT_STRANG`` runs the
    documented K(dt/2) D(dt) K(dt/2) with Phi recomputed after the drift.

    ``V`` is None, an array (static potential), or a callable
    ``V(xxx, yyy, zzz, psi, t)`` evaluated on the host every kick like the
    reference's V.(xxx,yyy,zzz,psi) (KDK.jl:64); ``t`` is the kick time
    t0 + step*dt.  ``Xmax`` is a 3-tuple (Xmax, Ymax, Zmax)."""
    psi = np.asfortranarray(np.asarray(psi0, dtype=np.complex128))
    nx, ny, nz = psi.shape
    X, Y, Z = Xmax
    x, y, z, delta, _ = setup_grid(X, Y, Z, nx, ny, nz, consistent=consistent_grid)
    if poisson_N is None:
        if consistent_grid:
            poisson_N = (nx, ny, nz)
        else:
            poisson_N = (max(nx - 1, 1), max(ny - 1, 1), max(nz - 1, 1))
    coords = None
    if callable(V):
        coords = setup_coordinates(x, y, z)
    lp = setup_fft_operators(X, Y, Z, nx, ny, nz, dt, hbar_over_m)
    if dt_kick is None:
        dt_kick = 0.5 * dt

    def potential(ps, t):
        if self_gravity:
            phi = compute_gravitational_potential(ps, delta, four_pi_G, be, poisson_N)
        else:
            phi = np.zeros(ps.shape)
        if V is None:
            v = None
        elif callable(V):
            v = V(coords[0], coords[1], coords[2], ps, t)
        else:
            v = V
        return phi, v

    def kick_real(ps, dtk, t):
        phi, v = potential(ps, t)
        pot = phi if v is None else phi + v
        return np.exp(-1j * dtk * pot / hbar_over_m) * ps

    if split == SPLIT_REFERENCE:
        for s in range(nsteps):
            t = t0 + s * dt
            phi, v = potential(psi, t)
            spec = apply_kick_step(psi, phi, v, dt_kick, hbar_over_m, be)
            psi = apply_drift_step(spec, lp, border, be)
            if callback is not None:
                callback(s, psi)
        return psi
    # Strang: K(dt/2) D(dt) K(dt/2) per step, Phi recomputed after each drift.
    for s in range(nsteps):
        t = t0 + s * dt
        psi = kick_real(psi, 0.5 * dt, t)
        spec = be.fftn(psi)
        psi = apply_drift_step(spec, lp, border, be)
        psi = kick_real(psi, 0.5 * dt, t + dt)
        if callback is not None:
            callback(s, psi)
    return np.asfortranarray(psi)


def mass(psi, dV):
    """Sum |psi|^2 dV (math.fsum: exactly rounded)."""
    return math.fsum((np.abs(np.asarray(psi)) ** 2).ravel().tolist()) * dV


def energies(psi, Xmax, hbar_over_m=1.0, four_pi_G=4.0 * np.pi,
             consistent_grid=False, V_field=None, poisson_N=None):
    """Spectral energies on the lattice used by the step:
    K = (hbar/m)^2/2 * sum_k k^2 |psi_k|^2 * dV/N, W = 1/2 sum rho Phi dV,
    Vext = sum rho V dV (all real, FP64)."""
    psi = np.asarray(psi)
    nx, ny, nz = psi.shape
    X, Y, Z = Xmax
    _, _, _, delta, dV = setup_grid(X, Y, Z, nx, ny, nz, consistent=consistent_grid)
    L = (2 * X, 2 * Y, 2 * Z)
    kx = 2 * np.pi * fft_order_int(nx) / L[0]
    ky = 2 * np.pi * fft_order_int(ny) / L[1]
    kz = 2 * np.pi * fft_order_int(nz) / L[2]
    k2 = kx[:, None, None] ** 2 + ky[None, :, None] ** 2 + kz[None, None, :] ** 2
    ph = np.fft.fftn(psi)
    N = psi.size
    K = 0.5 * hbar_over_m ** 2 * float(np.sum(k2 * np.abs(ph) ** 2)) * dV / N
    if poisson_N is None:
        poisson_N = (nx, ny, nz) if consistent_grid else (max(nx - 1, 1), max(ny - 1, 1), max(nz - 1, 1))
    phi = compute_gravitational_potential(psi, delta, four_pi_G, poisson_N=poisson_N).real
    rho = np.abs(psi) ** 2
    W = 0.5 * float(np.sum(rho * phi)) * dV
    Vx = 0.0 if V_field is None else float(np.sum(rho * V_field)) * dV
    return K, W, Vx
