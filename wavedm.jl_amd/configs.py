"""The benchmark workloads (kpc, Msun, Gyr; m_a = 1e-22 eV).

Each builder returns a dict with the grid, box half-width, coefficients, the
initial condition (device soliton rows or a host array) and the external
potential, so tests, the benchmark and Julia callers build identical runs.

* c0: 64^3, L = 20 kpc, centred soliton (r_c = 1 kpc, rho_c = 1e8) times a
  0.3 rad Gaussian random phase field (|n| <= 8); V = 0; dt = 1e-3 Gyr
  (Float32-rounded); 100 steps; FP64.
* c1: 256^3, L = 40 kpc, 64 solitons within r < 10 kpc + static Hernquist host
  (M = 1e11, a = 20 kpc) 50 kpc off the box centre; 500 steps; autoset dt.
* c2: 384^3, L = 20 kpc, 128 solitons within r < 5 kpc + tidal quadrupole
  V = 1/2 x^T T(t) x rotating with period 1 Gyr, eigenvalues +-(0.5-2)e3
  Gyr^-2; 1000 steps; autoset dt.
* c3: batched ensemble of 48 boxes: 16 each of 64^3 (L = 10 kpc), 128^3
  (L = 20 kpc) and 160^3 (L = 30 kpc), each a seeded 8-32 soliton
  superposition (positions within L/4 of the centre) with self-gravity and no
  external V; FP32; 200 steps; autoset dt per box.
* c4: 128^3, L = 20 kpc, 16 solitons + the c1 Hernquist host; 200 steps.
"""
from __future__ import annotations

import math

import numpy as np

from . import ic as IC
from . import units as U

AUTOSET_RATIO = 0.9  # the reference's autoset_timestep_ratio default (src/simulation_main.jl:53)


def _quadrupole(seed):
    rng = IC.SplitMix64(seed)
    lam = [(0.5 + 1.5 * rng.uniform()) * 1e3 * (1 if rng.uniform() < 0.5 else -1) for _ in range(3)]
    axis = np.array([rng.normal() for _ in range(3)])
    axis /= np.linalg.norm(axis)
    # fixed seeded orientation of the principal axes: rotate diag(lam) by a
    # seeded angle about a second seeded axis
    ax2 = np.array([rng.normal() for _ in range(3)])
    ax2 /= np.linalg.norm(ax2)
    ang = 2 * math.pi * rng.uniform()
    K = np.array([[0, -ax2[2], ax2[1]], [ax2[2], 0, -ax2[0]], [-ax2[1], ax2[0], 0]])
    R = np.eye(3) + math.sin(ang) * K + (1 - math.cos(ang)) * K @ K
    T0 = R @ np.diag(lam) @ R.T
    return T0, axis


def build(name: str, seed: int = 2026):
    hm = U.hbar_over_m(1.0)
    base = dict(hbar_over_m=hm, four_pi_G=U.FOUR_PI_G)
    if name == "c0":
        shape, X = (64, 64, 64), (10.0, 10.0, 10.0)
        return dict(base, name=name, shape=shape, xmax=X, precision="f64", steps=100,
                    dt=float(np.float32(1e-3)), psi0=IC.soliton_random_phase(shape, X, seed=seed),
                    sol_rows=None, potential=None)
    if name in ("c1", "c4"):
        shape, X, nsol, rpos, steps = ((256,) * 3, (20.0,) * 3, 64, 10.0, 500) if name == "c1" else \
            ((128,) * 3, (10.0,) * 3, 16, 5.0, 200)
        params = IC.soliton_set(nsol, seed, rpos)
        return dict(base, name=name, shape=shape, xmax=X, precision="f32", steps=steps, dt=None,
                    psi0=None, sol_rows=IC.device_rows(params, X, hm),
                    potential=("hernquist", dict(M=1e11, a=20.0, center=(50.0, 0.0, 0.0))))
    if name == "c3":
        members = []
        k = 0
        for n, L in ((64, 10.0), (128, 20.0), (160, 30.0)):
            for _ in range(16):
                rng = IC.SplitMix64(seed + 1000 + k)
                nsol = 8 + int(rng.uniform() * 25)  # 8..32
                X = (L / 2,) * 3
                params = IC.soliton_set(nsol, seed + 1000 + k, L / 4)
                members.append(dict(base, name=f"c3[{k}]", shape=(n,) * 3, xmax=X, precision="f32", steps=200,
                                    dt=None, psi0=None, sol_rows=IC.device_rows(params, X, hm), potential=None))
                k += 1
        return dict(base, name=name, members=members, steps=200, precision="f32")
    if name == "c2":
        shape, X = (384,) * 3, (10.0,) * 3
        params = IC.soliton_set(128, seed, 5.0)
        T0, axis = _quadrupole(seed + 1)
        return dict(base, name=name, shape=shape, xmax=X, precision="f32", steps=1000, dt=None,
                    psi0=None, sol_rows=IC.device_rows(params, X, hm),
                    potential=("quadrupole", dict(T0=T0, axis=tuple(axis), period=1.0, center=(0.0, 0.0, 0.0))))
    raise KeyError(name)


def install(plan, cfg):
    """Set psi0 and the external potential of a configuration on a plan."""
    if cfg["sol_rows"] is not None:
        plan.set_ic_solitons(cfg["sol_rows"])
    else:
        plan.set_psi(cfg["psi0"])
    plan.clear_potential_models()
    pot = cfg["potential"]
    if pot is not None:
        kind, kw = pot
        if kind == "hernquist":
            plan.set_hernquist(kw["M"], kw["a"], kw["center"])
        elif kind == "quadrupole":
            plan.set_quadrupole(kw["T0"], kw["axis"], kw["period"], kw["center"])


def autoset_dt(plan, nsteps, ratio=AUTOSET_RATIO):
    """compute_timestep with autoset (src/physics/poisson.jl:35-50): the step
    limit from the device FP64 CFL reductions of the initial state, Tmax
    chosen so that Nt = nsteps, dt rounded through Float32."""
    lims = plan.cfl(0.0)
    lim = min(lims)
    Tmax = nsteps * ratio * lim
    return float(np.float32(Tmax / nsteps)), lims
