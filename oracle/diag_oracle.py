"""FP64 restatement of the reference loop's per-step diagnostics -- test
infrastructure (only tests/ import it).

src/simulation_main.jl:550-558,641-656 and src/visualization/runtime.jl:203-232
compute, from the host copy of psi every step:
  rho_max, rho_max_id = findmax(|psi|^2)              (first maximum, column order)
  r_mass_center = |x - x[rho_max_id]|                   (plain Euclidean distance)
  radii = quantile(r_mass_center, weights(rho), 0.1:0.1:0.9)
  R = std(r[:, :, div(end, 2)], aweights(|psi[:, :, div(end, 2)]|^2)),  r = |x|
  total_halo_mass = sum(rho) * dV
and, for the virial plot, with grad_central finite differences:
  rho v = Im(conj(psi) grad psi); momentum = sum(rho v) dV
  V = -1/2 sum_int rho Phi dV;  K = 1/2 sum_int |rho v|^2 / rho dV;
  Q = 1/2 sum_int |grad sqrt(rho)|^2 dV;  virial = 2K + 2Q - V
(sum_int: cells [2:end-1] on every axis).  As in the reference, no hbar/m
factor enters rho v, KE or QE (its code units have hbar/m = 1), and a cell with
rho = 0 in the interior makes KE NaN (0/0).

Third-party pieces absent from the image (parity unpinned, restated from
their published definitions): StatsBase's weighted quantile (non-frequency
weights) and weighted std (analytic weights; with `corrected` not given
StatsBase takes the uncorrected estimate), and PhysicalFDM.grad_central
(central differences).  The boundary rule of grad_central is unknown: the
momentum sum, the one quantity that reaches the boundary cells, uses periodic
neighbours there (parity unpinned on the faces; the tests use fields that
vanish there).
"""
from __future__ import annotations

import numpy as np


def weighted_quantile(v, w, ps):
    """StatsBase quantile(v, weights(w), p) for non-frequency weights: sort by
    value, S_k = cumulative weight, h = p (W - w_1) + w_1, linear
    interpolation between the bracketing values."""
    v = np.asarray(v, dtype=np.float64).ravel()
    w = np.asarray(w, dtype=np.float64).ravel()
    nz = w != 0
    v, w = v[nz], w[nz]
    order = np.lexsort((w, v))  # sort the (v, w) pairs like sort!(collect(zip(v, w)))
    v, w = v[order], w[order]
    W = w.sum()
    w1 = w[0]
    S = np.cumsum(w)
    out = []
    for p in ps:
        h = p * (W - w1) + w1
        k = int(np.searchsorted(S, h, side="right"))  # first k (0-based) with S[k] > h
        if k >= len(v):
            out.append(v[-1])
            continue
        Sk = S[k]
        Skold = S[k - 1] if k > 0 else 0.0
        vk = v[k]
        vkold = v[k - 1] if k > 0 else 0.0
        out.append(vkold + (h - Skold) / (Sk - Skold) * (vk - vkold))
    return np.array(out)


def weighted_std(v, w):
    """StatsBase std(v, aweights(w)) without bias correction (its default):
    sqrt(sum w (v - mean)^2 / sum w)."""
    v = np.asarray(v, dtype=np.float64).ravel()
    w = np.asarray(w, dtype=np.float64).ravel()
    W = w.sum()
    mu = np.sum(w * v) / W
    return float(np.sqrt(np.sum(w * (v - mu) ** 2) / W))


def _cdiff(f, axis, d):
    return (np.roll(f, -1, axis=axis) - np.roll(f, 1, axis=axis)) / (2.0 * d)


def central_diff(f, axis, d):
    """grad_central along one axis: (f[i+1] - f[i-1]) / (2 d), periodic
    neighbours on the faces."""
    return _cdiff(f, axis, d)


def diagnostics(psi, phi, x, y, z, quantiles=np.arange(1, 10) / 10.0):
    """The reference's per-step diagnostics of psi and the real total potential
    phi (Phi_all) on the grid x, y, z (1-D coordinate vectors); phi may be None
    (then PE_abs = 0)."""
    psi = np.asarray(psi)
    rho = np.abs(psi) ** 2
    dx, dy, dz = x[1] - x[0], y[1] - y[0], z[1] - z[0]
    dV = dx * dy * dz
    out = {"mass": float(rho.sum() * dV)}
    flat = rho.ravel(order="F")
    imax = int(np.argmax(flat))  # first maximum in column order
    ix, iy, iz = np.unravel_index(imax, rho.shape, order="F")
    out.update(rho_max=float(flat[imax]), peak_index=imax, peak=(x[ix], y[iy], z[iz]))
    X, Y, Zc = np.meshgrid(x, y, z, indexing="ij")
    rc = np.sqrt((X - x[ix]) ** 2 + (Y - y[iy]) ** 2 + (Zc - z[iz]) ** 2)
    out["radii"] = weighted_quantile(rc, rho, quantiles)
    kz = z.size // 2 - 1 if z.size > 1 else 0  # Julia div(end, 2) (1-based)
    r = np.sqrt(X ** 2 + Y ** 2 + Zc ** 2)
    out["r_std_mid_plane"] = weighted_std(r[:, :, kz], rho[:, :, kz])
    grads = [_cdiff(psi, a, d) for a, d in ((0, dx), (1, dy), (2, dz))]
    rv = [np.imag(np.conj(psi) * g) for g in grads]
    out["momentum"] = tuple(float(c.sum() * dV) for c in rv)
    sl = (slice(1, -1),) * 3
    out["virial_potential"] = 0.0 if phi is None else float(-0.5 * np.sum(rho[sl] * phi[sl]) * dV)
    with np.errstate(divide="ignore", invalid="ignore"):
        kin = (rv[0] ** 2 + rv[1] ** 2 + rv[2] ** 2) / rho
    out["kinetic"] = float(0.5 * np.sum(kin[sl]) * dV)
    sr = np.abs(psi)
    q = sum(_cdiff(sr, a, d) ** 2 for a, d in ((0, dx), (1, dy), (2, dz)))
    out["quantum"] = float(0.5 * np.sum(q[sl]) * dV)
    out["virial"] = 2 * out["kinetic"] + 2 * out["quantum"] - out["virial_potential"]
    return out
