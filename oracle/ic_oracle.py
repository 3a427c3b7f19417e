"""FP64 restatement of the reference's velocity -> phase initial condition --
test infrastructure (only tests/ import it).

src/physics/schrodinger.jl:8-30 solve_vector_equation: theta is the sum of
three line integrals, each along its own axis from the first cell of that
axis, theta[:, j, k] += cumul_integrate(x, vx[:, j, k]) with x = (0:N-1) dx,
then the same along y (vy) and z (vz).  src/ic/generation.jl:330-336 removes
the mass-weighted mean velocity first; src/ic/generation.jl:185-187 assembles
psi = sqrt(rho) exp(i theta).

The acceleration -> velocity step before it (src/ic/generation.jl:246-337):
compute_acceleration_field (periodic branch :264-270): Phi = 4 pi *
parallel_poisson(Delta, [N-1...], rho_halo) and a = grad_central(-Delta, Phi);
generate_velocity_field's a_CDM (:285-291): |a| on the interior cells
[2:end-1]^3, 0 on the faces, then a_CDM[div(end,2), div(end,2), div(end,2)]
*= 0 (1-based); adjust_velocity_field's bulk_center_size zeroing (:316-326)
on the 1-based ranges div(N,2) - div(s,2) + 1 : div(N,2) - div(s,2) + s.  The
velocity formulas of that step (AstroIC.freefall_velocity_acc,
rotational_velocity_acc) and the bulk_perturb filters (zoom, shrink) come
from packages absent from the image: parity unpinned, taken as host inputs.
grad_central comes from PhysicalFDM (absent): central differences with
periodic neighbours on the faces here; a_CDM uses interior cells only, so it
does not depend on that face rule.

cumul_integrate comes from PhysicalParticles (absent from the image); it is
restated as the cumulative trapezoid rule of NumericalIntegration.jl's
published algorithm, out[0] = 0, out[i] = 1/2 sum_{l<=i} (x[l] - x[l-1])
(y[l] + y[l-1]) (parity unpinned at the rounding level; pinned by the closed
forms in tests/test_ic_oracle.py).  In units where hbar/m != 1 the phase is
theta / (hbar/m); the reference's code units have hbar/m = 1.
"""
from __future__ import annotations

import numpy as np


def cumul_integrate(x, y):
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    out = np.zeros_like(y)
    acc = 0.0
    for i in range(1, len(y)):
        acc += (x[i] - x[i - 1]) * (y[i] + y[i - 1])
        out[i] = 0.5 * acc
    return out


def cumul_integrate_axis(v, d, axis):
    """cumul_integrate along one axis of a 3-D array (vectorised restatement)."""
    n = v.shape[axis]
    x = np.arange(n, dtype=np.float64) * d
    dx = np.diff(x)
    shape = [1, 1, 1]
    shape[axis] = n - 1
    vs = np.moveaxis(v, axis, 0)
    terms = (dx.reshape(-1, 1, 1)) * (vs[1:] + vs[:-1])
    acc = np.concatenate([np.zeros((1,) + vs.shape[1:]), np.cumsum(terms, axis=0)], axis=0)
    return np.moveaxis(0.5 * acc, 0, axis)


def solve_vector_equation(vx, vy, vz, dx, dy, dz):
    theta = cumul_integrate_axis(vx, dx, 0)
    theta = theta + cumul_integrate_axis(vy, dy, 1)
    theta = theta + cumul_integrate_axis(vz, dz, 2)
    return theta


def remove_bulk_velocity(rho, vx, vy, vz):
    s = rho.sum()
    return vx - (vx * rho).sum() / s, vy - (vy * rho).sum() / s, vz - (vz * rho).sum() / s


def ic_from_velocity(rho, vx, vy, vz, d, hbar_over_m=1.0, remove_bulk=True):
    if remove_bulk:
        vx, vy, vz = remove_bulk_velocity(rho, vx, vy, vz)
    theta = solve_vector_equation(vx, vy, vz, *d)
    return np.sqrt(rho) * np.exp(1j * theta / hbar_over_m)


def compute_acceleration_field(rho_halo, delta, four_pi_G=4.0 * np.pi):
    """(Phi, ax, ay, az) of src/ic/generation.jl:264-272 (periodic)."""
    from oracle import ssfm_oracle as O
    from oracle import diag_oracle as D
    rho = np.asarray(rho_halo, dtype=np.float64)
    nx, ny, nz = rho.shape
    poisson_N = (max(nx - 1, 1), max(ny - 1, 1), max(nz - 1, 1))
    phi = four_pi_G * O.parallel_poisson(delta, poisson_N, rho - rho.mean()).real
    acc = tuple(-D.central_diff(phi, a, delta[a]) for a in range(3))
    return (phi,) + acc


def a_cdm(ax, ay, az):
    """generate_velocity_field's acceleration magnitude (src/ic/generation.jl:289-291)."""
    out = np.zeros(np.shape(ax))
    sl = (slice(1, -1),) * 3
    out[sl] = np.sqrt(ax[sl] ** 2 + ay[sl] ** 2 + az[sl] ** 2)
    c = tuple(n // 2 - 1 for n in out.shape)  # Julia div(end, 2), 1-based
    out[c] *= 0
    return out


def zero_center_velocity(vx, vy, vz, bulk_center_size):
    """adjust_velocity_field's bulk_center_size zeroing (src/ic/generation.jl:316-326)."""
    out = [np.array(v, dtype=np.float64, copy=True) for v in (vx, vy, vz)]
    if bulk_center_size > 0:
        rng = []
        for n in out[0].shape:
            lo = n // 2 - bulk_center_size // 2 + 1          # 1-based first
            hi = n // 2 - bulk_center_size // 2 + bulk_center_size  # 1-based last
            rng.append(slice(lo - 1, hi))
        for v in out:
            v[tuple(rng)] = 0.0
    return tuple(out)
