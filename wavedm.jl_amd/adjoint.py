"""Gradients through the SSFM step: the host side of wdm_adjoint_forward /
wdm_adjoint_backward (the adjoint runs on the device; this module only holds
the loss and the call sequence an ``rrule`` would issue).

The reference has no adjoint (it imports Zygote, src/WaveDM.jl:42, but defines
no rule for the step).  The loss here is the density-profile loss of the
gradient configuration: the L2 misfit of the log-spaced radial profile of
rho = |psi|^2 against a target profile.
"""
from __future__ import annotations

import numpy as np

PARAM_NAMES = ("M", "a", "cx", "cy", "cz", "sigma", "Txx", "Tyy", "Tzz", "Txy", "Txz", "Tyz")


def _coords(n, X):
    # LinRange(-X, X, n) as in setup_grid (src/physics/poisson.jl:8-15)
    if n == 1:
        return np.array([-X])
    t = np.arange(n) / (n - 1)
    return (1.0 - t) * (-X) + t * X


class ProfileLoss:
    """L = sum_b ((rho_b - t_b) / t_b)^2 over ``nbins`` log-spaced shells
    r in [Delta, min(Xmax)) about ``center``; rho_b is the mean |psi|^2 of the
    shell's cells.  Shells without cells or with t_b <= 0 are skipped."""

    def __init__(self, shape, xmax, nbins=40, center=(0.0, 0.0, 0.0)):
        self.shape = tuple(shape)
        self.nbins = int(nbins)
        xs = [_coords(n, X) for n, X in zip(self.shape, xmax)]
        d = [x[1] - x[0] if len(x) > 1 else 2.0 * X for x, X in zip(xs, xmax)]
        X, Y, Z = np.meshgrid(*xs, indexing="ij")
        r = np.sqrt((X - center[0]) ** 2 + (Y - center[1]) ** 2 + (Z - center[2]) ** 2)
        rmin, rmax = min(d), min(xmax)
        self.edges = np.logspace(np.log10(rmin), np.log10(rmax), self.nbins + 1)
        idx = np.searchsorted(self.edges, r, side="right") - 1
        idx[(r < rmin) | (r >= rmax)] = -1
        self.idx = idx
        self.sel = idx >= 0
        self.count = np.bincount(idx[self.sel], minlength=self.nbins).astype(np.float64)

    def profile(self, psi):
        rho = np.abs(psi) ** 2
        s = np.bincount(self.idx[self.sel], weights=rho[self.sel], minlength=self.nbins)
        return np.where(self.count > 0, s / np.maximum(self.count, 1.0), 0.0)

    def __call__(self, psi, target):
        """(L, dL/dRe(psi) + i dL/dIm(psi))."""
        target = np.asarray(target, dtype=np.float64)
        prof = self.profile(psi)
        ok = (self.count > 0) & (target > 0)
        t = np.where(ok, target, 1.0)
        res = np.where(ok, (prof - t) / t, 0.0)
        dprof = np.where(ok, 2.0 * res / (t * np.maximum(self.count, 1.0)), 0.0)
        w = np.zeros(self.shape)
        w[self.sel] = dprof[self.idx[self.sel]]
        return float(np.sum(res ** 2)), 2.0 * w * np.asarray(psi)


def value_and_grad(plan, psi0, nsteps, dt, loss, target, dt_kick=None, t0=0.0, ckpt_every=20, dt_scale=True):
    """Run ``nsteps`` reference-split steps from ``psi0`` on the device with
    checkpoints, evaluate ``loss(psi_N, target)`` and return
    (L, dL/dpsi0 packed as dL/dRe + i dL/dIm, the WDM_ADJ_NPARAM parameter
    gradients named by PARAM_NAMES: Hernquist M, a, centre; the dt scale
    sigma (NaN unless ``dt_scale``); the quadrupole's T0 components).
    This is the call sequence of the rrule's forward and pullback."""
    plan.set_psi(psi0)
    plan.adjoint_forward(nsteps, dt, dt_kick, t0, ckpt_every)
    cdt = np.complex64 if np.asarray(psi0).dtype == np.complex64 else np.complex128
    psiN = plan.get_psi(np.complex128)
    L, gN = loss(psiN, target)
    g0, gp = plan.adjoint_backward(gN.astype(cdt), dt_scale=dt_scale)
    return L, g0, gp
