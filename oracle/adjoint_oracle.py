"""FP64 CPU adjoint of the reference-split SSFM step -- test infrastructure.

Only tests/ (and bench.py's CPU leg) may import this module; it is the
reference the device adjoint (wdm_adjoint_forward / wdm_adjoint_backward) is
compared against, never part of the product path.

The forward map restates ssfm_oracle.run_ssfm with split=SPLIT_REFERENCE
(src/simulation_main.jl:536-547, src/solver/KDK.jl:37-101,
src/solver/gravity.jl:13-23), with the kick potential
    Phi[psi] + V_H(x; M, a, c) + V_Q(x, t; T0) + kappa |psi|^2
where V_H = -G M / (|x - c| + a) is the Hernquist host, V_Q = 1/2 u^T T0 u
with u = R(t)^T (x - c_Q) the rotating tidal quadrupole (R(t) the rotation
about the axis by 2 pi t / period; kick n at t0 + n dt), and sigma scales dt
and dt_kick.  The reference defines no adjoint (it imports Zygote,
src/WaveDM.jl:42, but has no rule for the step); this one is pinned by central
finite differences of the forward map in tests/test_adjoint_oracle.py.

Gradients are packed g = dL/dRe(psi) + i dL/dIm(psi).
"""
from __future__ import annotations

import numpy as np

from . import ssfm_oracle as O

PARAM_NAMES = ("M", "a", "cx", "cy", "cz", "sigma", "Txx", "Tyy", "Tzz", "Txy", "Txz", "Tyz")


def rotation(axis, angle):
    """Rotation by ``angle`` about ``axis`` (Rodrigues)."""
    n = np.asarray(axis, dtype=np.float64)
    n = n / np.linalg.norm(n)
    K = np.array([[0.0, -n[2], n[1]], [n[2], 0.0, -n[0]], [-n[1], n[0], 0.0]])
    return np.eye(3) + np.sin(angle) * K + (1.0 - np.cos(angle)) * (K @ K)


class Problem:
    """Grid, coefficients and potential of one adjoint run."""

    def __init__(self, shape, Xmax, dt, dt_kick=None, *, hbar_over_m=1.0, four_pi_G=4.0 * np.pi,
                 hernquist=None, kappa=0.0, border=None, self_gravity=True, quadrupole=None, t0=0.0):
        self.shape = tuple(shape)
        self.Xmax = tuple(Xmax)
        nx, ny, nz = self.shape
        x, y, z, self.delta, self.dV = O.setup_grid(*self.Xmax, nx, ny, nz)
        self.xxx, self.yyy, self.zzz = O.setup_coordinates(x, y, z)[:3]
        self.poisson_N = (max(nx - 1, 1), max(ny - 1, 1), max(nz - 1, 1))
        self.dt = float(dt)
        self.dt_kick = 0.5 * self.dt if dt_kick is None else float(dt_kick)
        self.hbar_over_m = float(hbar_over_m)
        self.four_pi_G = float(four_pi_G)
        self.G = self.four_pi_G / (4.0 * np.pi)
        self.hernquist = None if hernquist is None else tuple(float(v) for v in hernquist)  # M, a, cx, cy, cz
        # quadrupole: (T0 as xx, yy, zz, xy, xz, yz; axis; period; centre)
        self.quadrupole = None
        if quadrupole is not None:
            T0, axis, period, cen = quadrupole
            self.quadrupole = (tuple(float(v) for v in T0), tuple(float(v) for v in axis), float(period),
                               tuple(float(v) for v in cen))
        self.t0 = float(t0)
        self.kappa = float(kappa)
        self.border = None if border is None else np.asarray(border, dtype=np.float64)
        self.self_gravity = self_gravity
        # drift exponent w_k: the reference operator is exp(-i sigma w_k)
        # (src/solver/fft_operators.jl:8-15 with dt -> sigma dt)
        self.w = self._drift_exponent()

    def _drift_exponent(self):
        nx, ny, nz = self.shape
        fx = O._shifted_freq_1d(nx, self.Xmax[0])
        fy = O._shifted_freq_1d(ny, self.Xmax[1])
        fz = O._shifted_freq_1d(nz, self.Xmax[2])
        k2 = ((2 * np.pi * fx)[:, None, None] ** 2 + (2 * np.pi * fy)[None, :, None] ** 2
              + (2 * np.pi * fz)[None, None, :] ** 2)
        return self.hbar_over_m * k2 * self.dt / 2.0

    # -- pieces of the forward map ---------------------------------------------
    def phi(self, psi):
        if not self.self_gravity:
            return np.zeros(self.shape)
        return np.real(O.compute_gravitational_potential(psi, self.delta, self.four_pi_G,
                                                         poisson_N=self.poisson_N))

    def op_G(self, s):
        """The linear Poisson operator of the forward: four_pi_G lap^-1 (s - <s>)."""
        return self.four_pi_G * np.real(O.parallel_poisson(self.delta, self.poisson_N, s - s.mean()))

    def v_host(self, hq=None, t=0.0, T0=None):
        hq = self.hernquist if hq is None else hq
        v = np.zeros(self.shape)
        if hq is not None:
            M, a, cx, cy, cz = hq
            r = np.sqrt((self.xxx - cx) ** 2 + (self.yyy - cy) ** 2 + (self.zzz - cz) ** 2)
            v = v - self.G * M / (r + a)
        if self.quadrupole is not None:
            # V_Q = 1/2 u^T T0 u is linear in the six components of T0
            T = self.quadrupole[0] if T0 is None else T0
            for Tj, c in zip(T, self.dv_quad(t)):
                v = v + Tj * c
        return v

    def quad_u(self, t):
        """u = R(t)^T (x - c_Q) on the grid."""
        _, axis, period, cen = self.quadrupole
        ang = 2.0 * np.pi * t / period if period != 0.0 else 0.0
        R = rotation(axis, ang)
        d = (self.xxx - cen[0], self.yyy - cen[1], self.zzz - cen[2])
        return [R[0, a] * d[0] + R[1, a] * d[1] + R[2, a] * d[2] for a in range(3)]

    def dv_quad(self, t):
        """dV_Q/d(Txx, Tyy, Tzz, Txy, Txz, Tyz) at time t (V_Q is linear in T0)."""
        u = self.quad_u(t)
        return [0.5 * u[0] ** 2, 0.5 * u[1] ** 2, 0.5 * u[2] ** 2, u[0] * u[1], u[0] * u[2], u[1] * u[2]]

    def dv_host(self):
        """dV_H/d(M, a, cx, cy, cz) on the grid."""
        M, a, cx, cy, cz = self.hernquist
        d = (self.xxx - cx, self.yyy - cy, self.zzz - cz)
        r = np.sqrt(d[0] ** 2 + d[1] ** 2 + d[2] ** 2)
        ra = r + a
        dvdr = self.G * M / ra ** 2
        with np.errstate(divide="ignore", invalid="ignore"):
            dc = [np.where(r > 0, -dvdr * di / r, 0.0) for di in d]
        return [-self.G / ra, dvdr, *dc]

    def drift(self, psi, sigma=1.0):
        out = np.fft.ifftn(np.exp(-1j * sigma * self.w) * np.fft.fftn(psi))
        if self.border is not None:
            out = self.border[:, None, None] * out
        return out

    def kick_time(self, n, sigma=1.0):
        return self.t0 + n * sigma * self.dt

    def step(self, psi, sigma=1.0, hq=None, t=0.0, T0=None):
        tau = sigma * self.dt_kick / self.hbar_over_m
        pot = self.phi(psi) + self.v_host(hq, t, T0) + self.kappa * np.abs(psi) ** 2
        psip = np.exp(-1j * tau * pot) * psi
        return self.drift(psip, sigma)

    def forward(self, psi0, nsteps, sigma=1.0, hq=None, T0=None):
        psi = np.asarray(psi0, dtype=np.complex128)
        for n in range(nsteps):
            psi = self.step(psi, sigma, hq, self.kick_time(n, sigma), T0)
        return psi

    # -- adjoint ------------------------------------------------------------------
    def backward(self, psi0, nsteps, gN):
        """(g0, dL/d[M, a, cx, cy, cz, sigma, Txx, Tyy, Tzz, Txy, Txz, Tyz]) for
        the loss gradient gN at psi_N, sigma = 1 and the problem's potential
        parameters (zero for an absent model)."""
        states = [np.asarray(psi0, dtype=np.complex128)]
        for n in range(nsteps - 1):
            states.append(self.step(states[-1], t=self.kick_time(n)))
        tau = self.dt_kick / self.hbar_over_m
        dvh = self.dv_host() if self.hernquist is not None else None
        grads = np.zeros(12)
        g = np.asarray(gN, dtype=np.complex128)
        for n in range(nsteps - 1, -1, -1):
            psi = states[n]
            t = self.kick_time(n)
            vh = self.v_host(t=t)
            pot = self.phi(psi) + vh + self.kappa * np.abs(psi) ** 2
            E = np.exp(-1j * tau * pot)
            psip = E * psi
            bg = g if self.border is None else self.border[:, None, None] * g
            Fg = np.conj(np.exp(-1j * self.w)) * np.fft.fftn(bg)   # F g'
            gp = np.fft.ifftn(Fg)
            # drift part of dL/dsigma: (1/N) sum w Im(conj(F g') F psi')
            grads[5] += np.sum(self.w * np.imag(np.conj(Fg) * np.fft.fftn(psip))) / psi.size
            s = tau * np.imag(np.conj(gp) * psip)
            grads[5] += np.sum(s * pot)
            if dvh is not None:
                for j in range(5):
                    grads[j] += np.sum(s * dvh[j])
            if self.quadrupole is not None:
                for j, d in enumerate(self.dv_quad(t)):
                    grads[6 + j] += np.sum(s * d)
            u = (self.op_G(s) if self.self_gravity else 0.0) + self.kappa * s
            g = np.conj(E) * gp + 2.0 * u * psi
        return g, grads


def radial_bins(prob: Problem, nbins=40, center=(0.0, 0.0, 0.0)):
    """Bin index of every cell for nbins log-spaced shells r in [Delta, L/2)
    about ``center`` (-1 outside)."""
    r = np.sqrt((prob.xxx - center[0]) ** 2 + (prob.yyy - center[1]) ** 2 + (prob.zzz - center[2]) ** 2)
    rmin = min(prob.delta)
    rmax = min(prob.Xmax)
    edges = np.logspace(np.log10(rmin), np.log10(rmax), nbins + 1)
    idx = np.searchsorted(edges, r, side="right") - 1
    idx[(r < rmin) | (r >= rmax)] = -1
    return idx


def profile_loss(psi, idx, target, nbins=40):
    """L = sum_b ((rho_b - t_b)/t_b)^2, rho_b the mean |psi|^2 of shell b;
    returns (L, dL/dRe + i dL/dIm)."""
    rho = np.abs(psi) ** 2
    sel = idx >= 0
    cnt = np.bincount(idx[sel], minlength=nbins).astype(np.float64)
    sums = np.bincount(idx[sel], weights=rho[sel], minlength=nbins)
    ok = (cnt > 0) & (np.asarray(target) > 0)
    prof = np.where(cnt > 0, sums / np.maximum(cnt, 1), 0.0)
    t = np.where(ok, target, 1.0)
    res = np.where(ok, (prof - t) / t, 0.0)
    L = float(np.sum(res ** 2))
    dprof = np.where(ok, 2.0 * res / t / np.maximum(cnt, 1), 0.0)
    w = np.zeros(psi.shape)
    w[sel] = dprof[idx[sel]]
    return L, 2.0 * w * psi
