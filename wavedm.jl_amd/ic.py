"""Seeded initial conditions for the benchmark configurations.

psi0 = sqrt(rho) exp(i theta) as in the reference's setup
(src/ic/generation.jl:185-187, src/physics/schrodinger.jl:37-46), with the
configuration-defined profiles: a soliton rho_c / (1 + 0.091 (r/r_c)^2)^8
(the standard fuzzy-dark-matter core profile), Gaussian random phase fields and
soliton superpositions.  Randomness comes from SplitMix64 (Vigna's public
algorithm) so any host language can regenerate identical draws.
"""
from __future__ import annotations

import math

import numpy as np

_M64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & _M64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _M64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        """Uniform in (0, 1]."""
        return ((self.next_u64() >> 11) + 1) * (1.0 / 9007199254740992.0)

    def normal(self) -> float:
        """Box-Muller (cosine branch)."""
        u1, u2 = self.uniform(), self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


def grid_coords(shape, xmax, consistent=False):
    """Per-axis coordinates of the plan grid (LinRange(-X, X, N) by default)."""
    from .kdk import setup_grid
    x, y, z, _, _ = setup_grid(*xmax, *shape, consistent=consistent)
    return x, y, z


def soliton_density(r, rho_c, r_c):
    return rho_c / (1.0 + 0.091 * (r / r_c) ** 2) ** 8


def random_phase_field(shape, kmax_modes: float, rms: float, seed: int) -> np.ndarray:
    """Real Gaussian random field on the grid's discrete Fourier modes with
    |n| <= kmax_modes (integer lattice vectors), Hermitian, no k=0 term,
    rescaled to the requested rms."""
    rng = SplitMix64(seed)
    nx, ny, nz = shape
    K = int(math.floor(kmax_modes))
    coef = np.zeros(shape, dtype=np.complex128)
    for a in range(-K, K + 1):
        for b in range(-K, K + 1):
            for c in range(-K, K + 1):
                if a * a + b * b + c * c > kmax_modes * kmax_modes:
                    continue
                # one draw per +/- pair: keep the lexicographically positive vector
                if (a, b, c) <= (0, 0, 0) and (a, b, c) != (0, 0, 0):
                    continue
                if (a, b, c) == (0, 0, 0):
                    continue
                re, im = rng.normal(), rng.normal()
                coef[a % nx, b % ny, c % nz] += re + 1j * im
                coef[(-a) % nx, (-b) % ny, (-c) % nz] += re - 1j * im
    theta = np.fft.ifftn(coef).real
    s = math.sqrt(float(np.mean(theta ** 2)))
    return np.asfortranarray(theta * (rms / s if s > 0 else 0.0))


def soliton_random_phase(shape, xmax, rho_c=1e8, r_c=1.0, sigma=0.3, kcut=8, seed=2024,
                         consistent=False) -> np.ndarray:
    """Centred soliton times exp(i theta), theta a Gaussian random phase field
    (rms sigma rad, |n| <= kcut modes)."""
    x, y, z = grid_coords(shape, xmax, consistent)
    r = np.sqrt(x[:, None, None] ** 2 + y[None, :, None] ** 2 + z[None, None, :] ** 2)
    amp = np.sqrt(soliton_density(r, rho_c, r_c))
    theta = random_phase_field(shape, kcut, sigma, seed)
    return np.asfortranarray(amp * np.exp(1j * theta))


def soliton_set(n, seed, r_pos, rc_mu=math.log(0.8), rc_sigma=0.3, rho_c0=1e8, v_sigma_kms=20.0):
    """Seeded parameters of n solitons: positions uniform in a ball of radius
    r_pos, log-normal core radii, rho_c = rho_c0 (r_c/kpc)^-4, uniform phases,
    Gaussian velocity components (km/s -> kpc/Gyr).  Returns a (n, 9) array:
    x, y, z, r_c, rho_c, phase, vx, vy, vz."""
    rng = SplitMix64(seed)
    out = np.zeros((n, 9))
    for i in range(n):
        while True:
            p = [2.0 * rng.uniform() - 1.0 for _ in range(3)]
            if p[0] ** 2 + p[1] ** 2 + p[2] ** 2 <= 1.0:
                break
        rc = math.exp(rc_mu + rc_sigma * rng.normal())
        out[i, 0:3] = [r_pos * v for v in p]
        out[i, 3] = rc
        out[i, 4] = rho_c0 * rc ** -4
        out[i, 5] = 2.0 * math.pi * rng.uniform()
        out[i, 6:9] = [v_sigma_kms * 1.0227121650537077 * rng.normal() for _ in range(3)]
    return out


def device_rows(params, xmax, hbar_over_m):
    """Rows for Plan.set_ic_solitons: velocity kicks rounded to the box's
    Fourier lattice (the same rounding soliton_superposition applies)."""
    L = [2.0 * v for v in xmax]
    out = np.zeros((len(params), 9))
    for i, row in enumerate(params):
        cx_, cy_, cz_, rc, rhoc, ph, vx, vy, vz = row
        k = [2.0 * math.pi / Li * round(v / hbar_over_m * Li / (2.0 * math.pi)) for v, Li in zip((vx, vy, vz), L)]
        out[i] = [cx_, cy_, cz_, rc, rhoc, ph, *k]
    return out


def soliton_superposition(shape, xmax, params, hbar_over_m, consistent=False) -> np.ndarray:
    """sum_i sqrt(rho_i(|x - x_i|)) exp(i (phase_i + k_i . x)), with minimum-image
    distances in the periodic box and each velocity kick k_i = v_i/(hbar/m)
    rounded to the box's Fourier lattice so psi0 stays periodic."""
    x, y, z = grid_coords(shape, xmax, consistent)
    L = [2.0 * v for v in xmax]
    psi = np.zeros(shape, dtype=np.complex128)
    for row in params:
        cx_, cy_, cz_, rc, rhoc, ph, vx, vy, vz = row
        d = []
        for coord, c0, Li in ((x, cx_, L[0]), (y, cy_, L[1]), (z, cz_, L[2])):
            dd = coord - c0
            dd = dd - Li * np.round(dd / Li)
            d.append(dd)
        r = np.sqrt(d[0][:, None, None] ** 2 + d[1][None, :, None] ** 2 + d[2][None, None, :] ** 2)
        k = [2.0 * math.pi / Li * round(v / hbar_over_m * Li / (2.0 * math.pi)) for v, Li in zip((vx, vy, vz), L)]
        phase = ph + k[0] * x[:, None, None] + k[1] * y[None, :, None] + k[2] * z[None, None, :]
        psi += np.sqrt(soliton_density(r, rhoc, rc)) * np.exp(1j * phase)
    return np.asfortranarray(psi)


# ---------------------------------------------------------------- acceleration -> velocity
def compute_acceleration_field(plan, rho_halo, want=("phi", "ax", "ay", "az", "a_cdm")):
    """compute_acceleration_field + generate_velocity_field's magnitude on the
    device (src/ic/generation.jl:246-291, periodic gravity): Phi = 4 pi G
    lap^-1 (rho_halo - mean) on the plan's lattice, a = -grad Phi (central
    differences), a_CDM = |a| on the interior, 0 on the faces and at the
    1-based centre cell.  Returns a dict of the requested host arrays."""
    import ctypes as C
    from . import _capi as A
    rho = np.ascontiguousarray(np.asarray(rho_halo, dtype=np.float64).ravel(order="F"))
    if rho.size != plan.ncell:
        raise ValueError(f"rho_halo must have shape {plan.shape}")
    names = ("phi", "ax", "ay", "az", "a_cdm")
    outs = {k: np.empty(plan.ncell, dtype=np.float64) for k in names if k in want}
    ptr = lambda a: a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None  # noqa: E731
    rc = plan.lib.wdm_ic_acceleration(plan._h, ptr(rho), *[ptr(outs.get(k)) for k in names])
    A.check(plan.lib, plan._h, rc)
    return {k: v.reshape(plan.shape, order="F") for k, v in outs.items()}


def zero_center_velocity(vx, vy, vz, bulk_center_size):
    """adjust_velocity_field's bulk_center_size step (src/ic/generation.jl:316-326):
    zero the velocity on the 1-based ranges div(N,2)-div(s,2)+1 :
    div(N,2)-div(s,2)+s of every axis.  Host arrays, returned as copies."""
    out = [np.array(v, dtype=np.float64, copy=True) for v in (vx, vy, vz)]
    s = int(bulk_center_size)
    if s > 0:
        sl = tuple(slice(n // 2 - s // 2, n // 2 - s // 2 + s) for n in out[0].shape)
        for v in out:
            v[sl] = 0.0
    return tuple(out)


def velocity_ic(plan, rho_halo, velocity, bulk_center_size=0, bulk_perturb=False):
    """The reference's IC chain from a halo density (src/ic/generation.jl:179-187,
    303-337) on the device: accelerations and a_CDM, the velocity field
    ``velocity(x, y, z, a_CDM) -> (vx, vy, vz)`` (the AstroIC velocity formulas
    the reference calls are not in this image: the caller supplies them),
    centre zeroing, removal of the mass-weighted mean velocity, the line
    integrals of solve_vector_equation and psi = sqrt(rho) exp(i theta), left
    in ``plan``.  ``bulk_perturb`` (the reference's zoom/shrink filters, also
    absent) is not available."""
    if bulk_perturb:
        raise NotImplementedError("bulk_perturb needs the zoom/shrink filters, which are not in this image")
    acc = compute_acceleration_field(plan, rho_halo, want=("a_cdm",))
    x, y, z = grid_coords(plan.shape, plan.xmax, plan.consistent_grid)
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    vx, vy, vz = velocity(X, Y, Z, acc["a_cdm"])
    vx, vy, vz = zero_center_velocity(vx, vy, vz, bulk_center_size)
    plan.set_ic_velocity(rho_halo, vx, vy, vz, remove_bulk=True)
    return plan
