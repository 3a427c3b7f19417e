"""Python handle on a ``wdm_plan`` (the device-resident SSFM state).

Arrays cross the boundary in the reference's layout: numpy arrays of shape
(Nx, Ny, Nz) in Fortran order (x fastest), complex128 or complex64.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi as A


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Plan:
    """One grid, one precision, one HIP stream, all buffers device-resident.

    ``lib`` selects the native library (default: the HIP build)."""

    def __init__(self, shape, xmax, precision="f64", *, consistent_grid=False,
                 hbar_over_m=1.0, four_pi_G=4.0 * np.pi, self_gravity=True, lib=None):
        self.lib = lib if lib is not None else A.load()
        nx, ny, nz = (int(v) for v in shape)
        if np.isscalar(xmax):
            xmax = (float(xmax),) * 3
        self.shape = (nx, ny, nz)
        self.xmax = tuple(float(v) for v in xmax)
        self.precision = precision
        self.consistent_grid = consistent_grid
        prec = {"f64": A.WDM_F64, "f32": A.WDM_F32}[precision]
        flags = A.WDM_GRID_CONSISTENT if consistent_grid else 0
        h = C.c_void_p()
        rc = self.lib.wdm_plan_create(C.byref(h), nx, ny, nz, *self.xmax, prec, flags)
        A.check(self.lib, None, rc)
        self._h = h
        self.set_coeffs(hbar_over_m, four_pi_G)
        self.set_self_gravity(self_gravity)

    # -- lifetime -------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self.lib.wdm_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _chk(self, rc):
        A.check(self.lib, self._h, rc)

    # -- configuration --------------------------------------------------------
    @property
    def ncell(self):
        return self.shape[0] * self.shape[1] * self.shape[2]

    def set_coeffs(self, hbar_over_m, four_pi_G):
        self.hbar_over_m, self.four_pi_G = float(hbar_over_m), float(four_pi_G)
        self._chk(self.lib.wdm_set_coeffs(self._h, self.hbar_over_m, self.four_pi_G))

    def set_self_gravity(self, on: bool):
        self.self_gravity = bool(on)
        self._chk(self.lib.wdm_set_self_gravity(self._h, 1 if on else 0))

    def set_mass_ledger(self, on: bool):
        """Opt-in mass projection after every step (FP32 norm drift removal;
        needs self-gravity and no absorber)."""
        self._chk(self.lib.wdm_set_mass_ledger(self._h, 1 if on else 0))

    def mass_ledger_state(self):
        """(on, m_ref): whether the mass projection is on and the reference
        mass it holds (sum |psi|^2 without the cell volume; None while not
        seeded)."""
        on = C.c_int(0)
        m = C.c_double(0.0)
        self._chk(self.lib.wdm_get_mass_ledger(self._h, C.byref(on), C.byref(m)))
        return bool(on.value), (None if np.isnan(m.value) else float(m.value))

    def set_mass_reference(self, m_ref):
        """Seed the mass projection's reference mass (restart from a snapshot)."""
        self._chk(self.lib.wdm_set_mass_reference(self._h, float(m_ref)))

    def set_border(self, border):
        if border is None:
            self._chk(self.lib.wdm_set_border(self._h, None, 1))
            return
        b = np.ascontiguousarray(border, dtype=np.float64).ravel()
        if b.size != self.shape[0]:
            raise ValueError("border must be a length-Nx vector (broadcast along x)")
        self._chk(self.lib.wdm_set_border(self._h, _dptr(b), 1))

    def set_static_potential(self, V):
        if V is None:
            self._chk(self.lib.wdm_set_static_potential(self._h, None))
            return
        v = np.asarray(V, dtype=np.float64)
        if v.shape != self.shape:
            raise ValueError(f"V must have shape {self.shape}")
        v = np.ascontiguousarray(v.ravel(order="F"))
        self._chk(self.lib.wdm_set_static_potential(self._h, _dptr(v)))

    def set_hernquist(self, M, a, center=(0.0, 0.0, 0.0)):
        p = np.array([M, a, *center], dtype=np.float64)
        self._chk(self.lib.wdm_set_potential_model(self._h, A.WDM_V_HERNQUIST, _dptr(p), 5))

    def set_quadrupole(self, T0, axis, period, center=(0.0, 0.0, 0.0)):
        """V = 1/2 (x-c)^T R(t) T0 R(t)^T (x-c); T0 symmetric 3x3 (or the 6
        components xx, yy, zz, xy, xz, yz), R(t) rotation about ``axis`` by
        2 pi t / period."""
        T0 = np.asarray(T0, dtype=np.float64)
        if T0.shape == (3, 3):
            T0 = np.array([T0[0, 0], T0[1, 1], T0[2, 2], T0[0, 1], T0[0, 2], T0[1, 2]])
        p = np.array([*T0, *axis, period, *center], dtype=np.float64)
        self._chk(self.lib.wdm_set_potential_model(self._h, A.WDM_V_QUADRUPOLE, _dptr(p), 13))

    def set_kappa(self, kappa):
        p = np.array([kappa], dtype=np.float64)
        self._chk(self.lib.wdm_set_potential_model(self._h, A.WDM_V_KAPPA, _dptr(p), 1))

    def set_tidal(self, table, lo=None, hi=None, traj_t=None, traj_xyz=None):
        """Milky Way tidal potential in every kick (src/physics/tidal_fields.jl:64-146,
        applied; the reference computes and drops it).  ``table`` (tnx, tny, tnz)
        on the uniform grid from ``lo`` to ``hi`` (3 each); ``traj_t`` (ntraj,)
        kick times and ``traj_xyz`` (ntraj, 3) offsets: a kick at t uses the
        last row with traj_t <= t.  ``table=None`` clears."""
        if table is None:
            self._chk(self.lib.wdm_set_tidal_model(self._h, None, 0, 0, 0, None, None, None, None, 0))
            return
        tab = np.asarray(table, dtype=np.float64)
        if tab.ndim != 3:
            raise ValueError("table must be 3-D")
        lo = np.ascontiguousarray(lo, dtype=np.float64)
        hi = np.ascontiguousarray(hi, dtype=np.float64)
        tt = np.ascontiguousarray(np.atleast_1d(traj_t), dtype=np.float64)
        tx = np.ascontiguousarray(np.asarray(traj_xyz, dtype=np.float64).reshape(-1, 3))
        if lo.shape != (3,) or hi.shape != (3,) or tx.shape[0] != tt.size:
            raise ValueError("lo and hi need 3 entries and traj_xyz one row of 3 per traj_t entry")
        flat = np.ascontiguousarray(tab.ravel(order="F"))
        self._chk(self.lib.wdm_set_tidal_model(self._h, _dptr(flat), *tab.shape, _dptr(lo), _dptr(hi), _dptr(tt),
                                               _dptr(tx), int(tt.size)))

    def clear_potential_models(self):
        self._chk(self.lib.wdm_clear_potential_models(self._h))

    # -- state ----------------------------------------------------------------
    def set_psi(self, psi):
        psi = np.asarray(psi)
        if psi.shape != self.shape:
            raise ValueError(f"psi must have shape {self.shape}, got {psi.shape}")
        if psi.dtype == np.complex64:
            buf, dt = np.ascontiguousarray(psi.ravel(order="F")), A.WDM_C64
        else:
            buf, dt = np.ascontiguousarray(psi.astype(np.complex128).ravel(order="F")), A.WDM_C128
        self._chk(self.lib.wdm_set_psi(self._h, buf.ctypes.data_as(C.c_void_p), dt))

    def get_psi(self, dtype=np.complex128):
        out = np.empty(self.ncell, dtype=dtype)
        dt = A.WDM_C64 if np.dtype(dtype) == np.complex64 else A.WDM_C128
        self._chk(self.lib.wdm_get_psi(self._h, out.ctypes.data_as(C.c_void_p), dt))
        return out.reshape(self.shape, order="F")

    def set_psi_device(self, ptr: int, complex64: bool):
        self._chk(self.lib.wdm_set_psi_device(self._h, C.c_void_p(ptr),
                                              A.WDM_C64 if complex64 else A.WDM_C128))

    def get_psi_device(self, ptr: int, complex64: bool):
        self._chk(self.lib.wdm_get_psi_device(self._h, C.c_void_p(ptr),
                                              A.WDM_C64 if complex64 else A.WDM_C128))

    # -- operators ------------------------------------------------------------
    def poisson(self, rho=None, scale=0.0, want_phi=True):
        """scale * lap^-1 (rho - mean) (rho=None: |psi|^2).  Returns Phi (FP64)."""
        rbuf = None
        if rho is not None:
            rho = np.asarray(rho, dtype=np.float64)
            if rho.shape != self.shape:
                raise ValueError("rho shape mismatch")
            rbuf = np.ascontiguousarray(rho.ravel(order="F"))
        out = np.empty(self.ncell, dtype=np.float64) if want_phi else None
        self._chk(self.lib.wdm_poisson(self._h, _dptr(rbuf) if rbuf is not None else None,
                                       _dptr(out) if out is not None else None, float(scale)))
        return None if out is None else out.reshape(self.shape, order="F")

    def get_phi(self):
        out = np.empty(self.ncell, dtype=np.float64)
        self._chk(self.lib.wdm_get_phi(self._h, _dptr(out)))
        return out.reshape(self.shape, order="F")

    def kick(self, dt_kick, t=0.0):
        self._chk(self.lib.wdm_kick(self._h, float(dt_kick), float(t)))

    def drift(self, dt):
        self._chk(self.lib.wdm_drift(self._h, float(dt)))

    def step(self, nsteps, dt, dt_kick=None, t0=0.0, split="reference"):
        if dt_kick is None:
            dt_kick = 0.5 * dt
        mode = {"reference": A.WDM_SPLIT_REFERENCE, "strang": A.WDM_SPLIT_STRANG}[split]
        self._chk(self.lib.wdm_step(self._h, int(nsteps), float(dt), float(dt_kick), float(t0), mode))

    def step_timed(self, nsteps, dt, dt_kick=None, t0=0.0, passes=None):
        """Reference-split steps with HIP events around every pass (or only the
        passes named in ``passes``); returns (per-pass summed ms dict, total
        event ms)."""
        if dt_kick is None:
            dt_kick = 0.5 * dt
        ms = np.zeros(len(A.PASS_NAMES), dtype=np.float64)
        tot = np.zeros(1, dtype=np.float64)
        if passes is None:
            self._chk(self.lib.wdm_step_timed(self._h, int(nsteps), float(dt), float(dt_kick), float(t0),
                                              _dptr(ms), len(ms), _dptr(tot)))
            return dict(zip(A.PASS_NAMES, ms.tolist())), float(tot[0])
        mask = 0
        for name in passes:
            mask |= 1 << A.PASS_NAMES.index(name)
        self._chk(self.lib.wdm_step_timed_mask(self._h, int(nsteps), float(dt), float(dt_kick), float(t0),
                                               mask, _dptr(ms), len(ms), _dptr(tot)))
        return {n: ms[A.PASS_NAMES.index(n)] for n in passes}, float(tot[0])

    def set_ic_solitons(self, params):
        """Device-side soliton superposition (rows x0,y0,z0,r_c,rho_c,phase,kx,ky,kz)."""
        prm = np.ascontiguousarray(np.asarray(params, dtype=np.float64).reshape(-1, 9))
        self._chk(self.lib.wdm_ic_solitons(self._h, _dptr(prm), prm.shape[0]))

    def set_ic_velocity(self, rho, vx, vy, vz, remove_bulk=True):
        """psi = sqrt(rho) exp(i theta / (hbar/m)), theta from the velocity field
        by the reference's solve_vector_equation (src/physics/schrodinger.jl:8-30),
        after removing the rho-weighted mean velocity (src/ic/generation.jl:330-336)."""
        arrs = []
        for a in (rho, vx, vy, vz):
            a = np.asarray(a, dtype=np.float64)
            if a.shape != self.shape:
                raise ValueError(f"fields must have shape {self.shape}")
            arrs.append(np.ascontiguousarray(a.ravel(order="F")))
        self._chk(self.lib.wdm_ic_velocity(self._h, *[_dptr(a) for a in arrs], 1 if remove_bulk else 0))

    def reduce(self, n=A.WDM_RED_COUNT):
        out = np.zeros(A.WDM_RED_COUNT, dtype=np.float64)
        self._chk(self.lib.wdm_reduce(self._h, _dptr(out), int(n)))
        keys = ["mass", "kinetic", "w_self", "v_ext", "psi_max", "phi_max"]
        return {k: float(out[i]) for i, k in enumerate(keys[:n])}

    # Column names of the reference's Prop.csv (src/io/save.jl:46-67) for the
    # WDM_DIAG_* slots, in slot order.
    DIAG_KEYS = ("TotalMass", "rho_max", "peak_ix", "peak_iy", "peak_iz", "R",
                 "R1", "R2", "R3", "R4", "R5", "R6", "R7", "R8", "R9",
                 "MomentumX", "MomentumY", "MomentumZ", "PE_abs", "KE", "QE", "Virial")

    def diagnostics(self, n=A.WDM_DIAG_COUNT):
        """The reference loop's per-step diagnostics of the current psi
        (src/simulation_main.jl:550-558,641-656; src/visualization/runtime.jl:210-238),
        keyed by the Prop.csv column names; ``n`` < WDM_DIAG_PE_ABS + 1 skips
        the potential."""
        out = np.zeros(A.WDM_DIAG_COUNT, dtype=np.float64)
        self._chk(self.lib.wdm_diagnostics(self._h, _dptr(out), int(n)))
        d = {k: float(out[i]) for i, k in enumerate(self.DIAG_KEYS[:n])}
        for k in ("peak_ix", "peak_iy", "peak_iz"):
            if k in d:
                d[k] = int(d[k])
        return d

    def acceleration(self):
        """-grad Phi_all by central differences (grad_central(-dx, Phi_all),
        src/simulation_main.jl:652): three real arrays of the grid shape."""
        out = [np.empty(self.ncell, dtype=np.float64) for _ in range(3)]
        self._chk(self.lib.wdm_acceleration(self._h, *[_dptr(a) for a in out]))
        return tuple(a.reshape(self.shape, order="F") for a in out)

    def density_accumulate(self):
        """buffer_psi2 += |psi|^2 on the device (src/simulation_main.jl:659-660)."""
        self._chk(self.lib.wdm_density_accumulate(self._h))

    def density_average(self):
        """(buffer / count, count): the time-averaged density of the accumulated
        steps (src/io/save.jl:79-81), or (None, 0) when nothing was accumulated."""
        n = C.c_longlong(0)
        self._chk(self.lib.wdm_density_average(self._h, None, C.byref(n)))
        if n.value == 0:
            return None, 0
        out = np.empty(self.ncell, dtype=np.float64)
        self._chk(self.lib.wdm_density_average(self._h, _dptr(out), C.byref(n)))
        return out.reshape(self.shape, order="F"), int(n.value)

    def density_reset(self):
        self._chk(self.lib.wdm_density_reset(self._h))

    def density_sum(self):
        """(FP64 sum of |psi|^2 over the accumulated steps, count); (None, 0)
        when nothing was accumulated."""
        n = C.c_longlong(0)
        self._chk(self.lib.wdm_density_sum(self._h, None, C.byref(n)))
        if n.value == 0:
            return None, 0
        out = np.empty(self.ncell, dtype=np.float64)
        self._chk(self.lib.wdm_density_sum(self._h, _dptr(out), C.byref(n)))
        return out.reshape(self.shape, order="F"), int(n.value)

    def density_restore(self, total, count):
        """Replace the device sum and its count (continue a saved average)."""
        t = np.asarray(total, dtype=np.float64)
        if t.shape != self.shape:
            raise ValueError(f"sum must have shape {self.shape}")
        buf = np.ascontiguousarray(t.ravel(order="F"))
        self._chk(self.lib.wdm_density_restore(self._h, _dptr(buf), int(count)))

    def cfl(self, kappa=0.0):
        out = np.zeros(3, dtype=np.float64)
        self._chk(self.lib.wdm_cfl(self._h, float(kappa), _dptr(out)))
        return tuple(float(v) for v in out)

    def synchronize(self):
        self._chk(self.lib.wdm_synchronize(self._h))

    # -- tuning ---------------------------------------------------------------
    def autotune(self, trial_steps, dt):
        """Time every compiled pass-table variant per axis on ``trial_steps``
        steps of a copy of the state and keep the fastest; returns
        (variants[3], ms per step of each axis's passes[3])."""
        v = (C.c_int * 3)()
        ms = np.zeros(3, dtype=np.float64)
        self._chk(self.lib.wdm_autotune(self._h, int(trial_steps), float(dt), v, _dptr(ms)))
        return [int(x) for x in v], ms

    def select_variants(self, variants):
        v = (C.c_int * 3)(*[int(x) for x in variants])
        self._chk(self.lib.wdm_select_variants(self._h, v))

    # -- adjoint --------------------------------------------------------------
    def adjoint_forward(self, nsteps, dt, dt_kick=None, t0=0.0, ckpt_every=20):
        """Reference-split steps from the current psi with device checkpoints
        every ``ckpt_every`` steps (the taped forward of the adjoint)."""
        if dt_kick is None:
            dt_kick = 0.5 * dt
        self._chk(self.lib.wdm_adjoint_forward(self._h, int(nsteps), float(dt), float(dt_kick), float(t0),
                                               int(ckpt_every)))

    def adjoint_backward(self, grad_psi, dt_scale=False):
        """Given dL/dRe(psi_N) + i dL/dIm(psi_N), return (grad_psi0, grad_params)
        with grad_params = [dL/dM, dL/da, dL/dcx, dL/dcy, dL/dcz, dL/dsigma,
        dL/dTxx, dL/dTyy, dL/dTzz, dL/dTxy, dL/dTxz, dL/dTyz] (dL/dsigma is NaN
        unless ``dt_scale``; a model that is not set has zero gradients)."""
        g = np.asarray(grad_psi)
        if g.shape != self.shape:
            raise ValueError(f"grad_psi must have shape {self.shape}")
        c64 = g.dtype == np.complex64
        gbuf = np.ascontiguousarray((g if c64 else g.astype(np.complex128)).ravel(order="F"))
        out = np.empty(self.ncell, dtype=np.complex64 if c64 else np.complex128)
        prm = np.zeros(A.WDM_ADJ_NPARAM, dtype=np.float64)
        self._chk(self.lib.wdm_adjoint_backward(
            self._h, gbuf.ctypes.data_as(C.c_void_p), A.WDM_C64 if c64 else A.WDM_C128,
            out.ctypes.data_as(C.c_void_p), _dptr(prm), len(prm), A.WDM_ADJ_DT_SCALE if dt_scale else 0))
        return out.reshape(self.shape, order="F"), prm
