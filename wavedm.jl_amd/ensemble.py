"""Batched ensemble of independent boxes (wdm_ensemble_*): many plans of
uneven grids stepped together (members of one shape in one launch per pass),
one HIP-graph launch per step, with per-shape autotuned pass tables.  The reference steps one box at a time
(src/simulation_main.jl:536-551); this is the host side of the batched form."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi as A


class Ensemble:
    """Members are Plan objects owned by the caller (they must stay open)."""

    def __init__(self, plans):
        self.plans = list(plans)
        if not self.plans:
            raise ValueError("an ensemble needs at least one plan")
        self.lib = self.plans[0].lib
        if any(p.lib is not self.lib for p in self.plans):
            raise ValueError("all members must use the same library")
        arr = (C.c_void_p * len(self.plans))(*[p._h for p in self.plans])
        h = C.c_void_p()
        A.check(self.lib, None, self.lib.wdm_ensemble_create(C.byref(h), arr, len(self.plans)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self.lib.wdm_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _vec(self, v, name):
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=np.float64), (len(self.plans),)))
        if not np.all(np.isfinite(a)):
            raise ValueError(f"{name} must be finite")
        return a

    def step(self, nsteps, dt, dt_kick=None, t0=0.0, graph=True):
        """``nsteps`` reference-split steps of every member; ``dt`` (and
        ``dt_kick``, default dt/2) scalar or one per member."""
        d = self._vec(dt, "dt")
        k = self._vec(0.5 * d if dt_kick is None else dt_kick, "dt_kick")
        A.check(self.lib, None, self.lib.wdm_ensemble_step(
            self._h, int(nsteps), d.ctypes.data_as(A._DP), k.ctypes.data_as(A._DP), float(t0), 1 if graph else 0))

    def layout(self):
        """(batches, members stepping alone) of the last step: members of one
        shape, precision, pass table and potential kind share one launch per
        pass."""
        nb, ns = C.c_int(0), C.c_int(0)
        A.check(self.lib, None, self.lib.wdm_ensemble_layout(self._h, C.byref(nb), C.byref(ns)))
        return nb.value, ns.value

    def autotune(self, trial_steps, dt):
        d = self._vec(dt, "dt")
        A.check(self.lib, None, self.lib.wdm_ensemble_autotune(self._h, int(trial_steps), d.ctypes.data_as(A._DP)))
