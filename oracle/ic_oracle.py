"""FP64 restatement of the reference's velocity -> phase initial condition --
test infrastructure (only tests/ import it).

src/physics/schrodinger.jl:8-30 solve_vector_equation: theta is the sum of
three line integrals, each along its own axis from the first cell of that
axis, theta[:, j, k] += cumul_integrate(x, vx[:, j, k]) with x = (0:N-1) dx,
then the same along y (vy) and z (vz).  src/ic/generation.jl:330-336 removes
the mass-weighted mean velocity first; src/ic/generation.jl:185-187 assembles
psi = sqrt(rho) exp(i theta).

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
