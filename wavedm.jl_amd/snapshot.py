"""Snapshots, restart and the property table of the reference loop.

The reference writes JLD2 files keyed "ψ", "Φ_all", "ψ_last_t",
"ψ2_averaged" (src/io/save.jl:27-34; per-step snapshots
src/simulation_main.jl:662-667) and reloads "ψ" as an initial condition
(src/simulation_main.jl:1033-1042); it keeps no time or step counter, so a run
cannot be continued exactly.  Here a snapshot is one uncompressed ``.npz``:
the same fields under ASCII keys (psi, phi_all, psi_last_t, psi2_averaged) in
the plan's precision and Julia's column order, plus a JSON header with the
grid, the coefficients, the time, the step counter and dt, so that
``restore`` followed by the remaining steps reproduces an uninterrupted run bit
for bit.  Potential models and the absorber are configuration, not state: the
caller installs them again, as the reference rebuilds them from its config.

Two pieces of device state beyond psi are saved so that this holds for every
plan option:
* the mass projection (wdm_set_mass_ledger): its on/off flag and the
  reference mass it holds the state to (restored with wdm_set_mass_reference;
  re-seeding it from the restored psi would move it by the last step's
  rounding).  The projection applies its pending scale when a step call
  returns, so "uninterrupted" means the same sequence of step calls: a run
  saved between two calls continues exactly as the second call would have.
* the time-averaged density: the FP64 device sum itself ("psi2_sum") and its
  count, next to the reference's "psi2_averaged" = sum / count, so the
  average continues where it stopped (the reference restarts it).

``PropTable`` is save_property_dataframe (src/io/save.jl:41-72): one row per
step of the wdm_diagnostics values under the reference's Prop.csv column
names and order, written as CSV.
"""
from __future__ import annotations

import csv
import json
import math

import numpy as np

FORMAT = "wavedm-hip-snapshot-1"


def save(path, plan, *, t, step, dt=None, phi_all=False, psi_last_t=None, average=False, meta=None):
    """Write the plan's state to ``path`` (``.npz``).  ``phi_all`` also stores
    the self-gravity potential of psi; ``average`` stores the device
    time-averaged density (when any was accumulated)."""
    dtype = np.complex64 if plan.precision == "f32" else np.complex128
    arrays = {"psi": plan.get_psi(dtype)}
    if phi_all:
        arrays["phi_all"] = plan.get_phi()
    if psi_last_t is not None:
        arrays["psi_last_t"] = np.asarray(psi_last_t)
    avg_n = 0
    if average:
        total, avg_n = plan.density_sum()
        if total is not None:
            arrays["psi2_averaged"] = total / float(avg_n)
            arrays["psi2_sum"] = total
    ledger_on, m_ref = plan.mass_ledger_state()
    header = {"format": FORMAT, "shape": list(plan.shape), "xmax": list(plan.xmax), "precision": plan.precision,
              "consistent_grid": bool(plan.consistent_grid), "hbar_over_m": plan.hbar_over_m,
              "four_pi_G": plan.four_pi_G, "self_gravity": bool(getattr(plan, "self_gravity", True)),
              "t": float(t), "step": int(step), "dt": None if dt is None else float(dt),
              "average_count": int(avg_n), "mass_ledger": ledger_on, "mass_reference": m_ref,
              "meta": meta or {}}
    arrays = {k: np.asfortranarray(v) for k, v in arrays.items()}
    with open(path, "wb") as f:
        np.savez(f, header=np.frombuffer(json.dumps(header).encode(), dtype=np.uint8), **arrays)


def load(path):
    """The header dict with the arrays added under their keys (Fortran order)."""
    with np.load(path, allow_pickle=False) as z:
        header = json.loads(bytes(z["header"]).decode())
        if header.get("format") != FORMAT:
            raise ValueError(f"{path}: not a {FORMAT} file")
        for k in z.files:
            if k != "header":
                header[k] = np.asfortranarray(z[k])
    return header


def _restore_state(plan, snap):
    plan.set_psi(snap["psi"])
    if snap.get("mass_ledger"):
        plan.set_mass_ledger(True)
        if snap.get("mass_reference") is not None:
            plan.set_mass_reference(snap["mass_reference"])
    elif snap.get("mass_ledger") is False:
        plan.set_mass_ledger(False)
    if "psi2_sum" in snap and snap.get("average_count", 0) > 0:
        plan.density_restore(snap["psi2_sum"], snap["average_count"])
    else:
        plan.density_reset()


def plan_from(snap, *, lib=None):
    """A new Plan with the snapshot's grid, precision and coefficients, holding
    its psi, mass projection state and density average (potential models and
    absorber are installed by the caller)."""
    from .plan import Plan
    p = Plan(tuple(snap["shape"]), tuple(snap["xmax"]), snap["precision"], consistent_grid=snap["consistent_grid"],
             hbar_over_m=snap["hbar_over_m"], four_pi_G=snap["four_pi_G"], self_gravity=snap["self_gravity"],
             lib=lib)
    _restore_state(p, snap)
    return p


def restore(plan, snap):
    """Load the snapshot's psi, mass projection state and density average into
    an existing plan of the same grid; returns (t, step).  A restored run
    continues with kick k at t + k*dt."""
    if tuple(snap["shape"]) != tuple(plan.shape) or snap["precision"] != plan.precision:
        raise ValueError("snapshot grid or precision differs from the plan")
    _restore_state(plan, snap)
    return snap["t"], snap["step"]


# Prop.csv columns (src/io/save.jl:46-67): always, then with plot_virial
PROP_COLUMNS = ("t", "R", "R1", "R2", "R3", "R4", "R5", "R6", "R7", "R8", "R9", "TotalMass")
VIRIAL_COLUMNS = ("PE_abs", "KE", "QE", "Virial", "MomentumX", "MomentumY", "MomentumZ")


class PropTable:
    """Rows of (t, diagnostics) -> Prop.csv.  ``scale`` maps a column name to
    the unit factor the reference applies when it records it (uL for the radii,
    uMomentum for the momenta; src/simulation_main.jl:557-558,642-650).

    TotalMass lags the other columns by one row, as in the reference: the loop
    pushes ``total_halo_mass`` (src/simulation_main.jl:557) before it
    recomputes it from the new state (:652), so row k holds the mass of the
    state of row k-1 and the first row the initial condition's mass
    (``initial_mass``, the reference's :349), with an infinite mass recorded
    as 0.  Without ``initial_mass`` the first row takes its own state's mass."""

    def __init__(self, virial=True, scale=None, initial_mass=None):
        self.columns = PROP_COLUMNS + (VIRIAL_COLUMNS if virial else ())
        self.scale = dict(scale or {})
        self.rows = []
        self._mass_prev = None if initial_mass is None else float(initial_mass)

    def append(self, t, diag):
        row = {"t": float(t)}
        for c in self.columns[1:]:
            if c != "TotalMass":
                row[c] = float(diag[c]) * self.scale.get(c, 1.0)
        m = float(diag["TotalMass"]) if self._mass_prev is None else self._mass_prev
        row["TotalMass"] = (0.0 if math.isinf(m) else m) * self.scale.get("TotalMass", 1.0)
        self._mass_prev = float(diag["TotalMass"])
        self.rows.append(row)

    def write(self, path):
        with open(path, "w", newline="") as f:
            w = csv.writer(f)
            w.writerow(self.columns)
            for r in self.rows:
                w.writerow([repr(r[c]) for c in self.columns])

    @staticmethod
    def read(path):
        with open(path, newline="") as f:
            rd = csv.reader(f)
            cols = next(rd)
            return cols, [[float(v) for v in row] for row in rd]
