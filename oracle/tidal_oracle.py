"""FP64 restatement of the reference's MW tidal potential -- test
infrastructure (only tests/ import it).

src/physics/tidal_fields.jl:64-146, the MW_tidal_interpolate branch, as the
code intends it.  In the reference the result never reaches the kick: the
last line `Φ_all += pot_tidal` (:117) rebinds a local and the caller drops the
return, and cancel_field_gradient_at_center! rebinds `field -= ...` (:133)
before its in-place `.+=` (:143); the device model that follows this oracle is
therefore an opt-in deviation (the reference's effect is none).  Steps:

* pot = interpolate(MW_grid, MW_Phi, x + o) at every cell, o the trajectory
  offset of the kick (df_traj row id_t, over uL) (:83-84).  GridInterpolations
  (absent from the image) is restated as multilinear interpolation on the
  rectangular grid with coordinates clamped into its range (parity unpinned).
* pot .-= pot[rho_max_id] (:114), rho_max_id the density peak.
* cancel_field_gradient_at_center!(pot, rho_max_id, ...) (:126-146):
  g_a = (pot[c + e_a] - pot[c - e_a]) / 2 (index units), pot -= g_a (i_a -
  div(N_a, 2)) with i_a 1-based, then pot += sum(ramps * rho) / sum(rho).
  Neighbours past a face wrap around (the reference would index out of
  bounds there).
* The trajectory row: findfirstvalue (AstroSimBase, absent) is taken as the
  last trajectory sample at or before the kick time (parity unpinned); the
  device takes that table from the caller.
"""
from __future__ import annotations

import numpy as np


def interpolate(axes, table, pts):
    """Multilinear interpolation of ``table`` (shape len(ax) per axis) on the
    rectangular grid ``axes`` at points ``pts`` (..., 3), coordinates clamped
    into the grid."""
    pts = np.asarray(pts, dtype=np.float64)
    idx, w = [], []
    for a, ax in enumerate(axes):
        ax = np.asarray(ax, dtype=np.float64)
        c = np.clip(pts[..., a], ax[0], ax[-1])
        i = np.clip(np.searchsorted(ax, c, side="right") - 1, 0, len(ax) - 2)
        t = (c - ax[i]) / (ax[i + 1] - ax[i])
        idx.append(i)
        w.append(t)
    out = np.zeros(pts.shape[:-1])
    for bx in (0, 1):
        for by in (0, 1):
            for bz in (0, 1):
                wt = ((w[0] if bx else 1 - w[0]) * (w[1] if by else 1 - w[1]) * (w[2] if bz else 1 - w[2]))
                out += wt * table[idx[0] + bx, idx[1] + by, idx[2] + bz]
    return out


def tidal_potential(axes, table, x, y, z, offset, rho, center=None):
    """The tidal potential the kick would add (the reference's intent).
    ``center`` is the 0-based peak index (default: first argmax of rho in
    column order)."""
    X, Y, Zg = np.meshgrid(x, y, z, indexing="ij")
    pts = np.stack([X + offset[0], Y + offset[1], Zg + offset[2]], axis=-1)
    pot = interpolate(axes, table, pts)
    rho = np.asarray(rho, dtype=np.float64)
    if center is None:
        flat = int(np.argmax(rho.ravel(order="F")))
        center = np.unravel_index(flat, rho.shape, order="F")
    c = tuple(int(v) for v in center)
    pot = pot - pot[c]
    return cancel_field_gradient(pot, c, rho)


def cancel_field_gradient(field, c, rho):
    """src/physics/tidal_fields.jl:126-146 (its intent): the three central
    differences at c first, then each ramp subtracted and its rho-weighted sum
    accumulated, then the constant sum_PE / sum(rho) added."""
    n = field.shape
    field = np.array(field, dtype=np.float64, copy=True)
    grads = []
    for a in range(3):
        up, dn = list(c), list(c)
        up[a] = (c[a] + 1) % n[a]
        dn[a] = (c[a] - 1) % n[a]
        grads.append((field[tuple(up)] - field[tuple(dn)]) / 2.0)
    sum_pe = 0.0
    for a in range(3):
        shape = [1, 1, 1]
        shape[a] = n[a]
        ramp = ((np.arange(1, n[a] + 1) - n[a] // 2) * grads[a]).reshape(shape) * np.ones(n)
        field = field - ramp
        sum_pe += float(np.sum(ramp * rho))
    return field + sum_pe / float(np.sum(rho))


def trajectory_offset(times, xyz, t):
    """The trajectory sample of the kick at time t: the last sample with
    time <= t (the first one before the table starts)."""
    times = np.asarray(times, dtype=np.float64)
    k = int(np.searchsorted(times, t, side="right")) - 1
    return np.asarray(xyz, dtype=np.float64)[max(k, 0)]
