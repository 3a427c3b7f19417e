"""The 384^3 workload (128 solitons, rotating tidal quadrupole, self-gravity,
autoset dt) at its full size, checked through properties that do not depend
on the grid size; the CPU oracle is too slow at this size.

* Time reversal: one reference step is D(dt) K(dt_kick).  Its inverse is
  K(-dt_kick) D(-dt), because |K psi| = |psi| leaves Phi unchanged.  Running
  the steps and then their inverses must return psi0.  Bars: relative L2
  <= 1e-11 (FP64) and <= 1e-5 (FP32).
* Mass conservation at absorb_coeff = 0: relative change <= 1e-12 over 10
  FP64 steps.
* FP32 against FP64 from the same (FP32-rounded) psi0 after 5 steps: relative
  L2 <= 1e-5, the FP32 precision form's bar for psi, at a horizon where the
  dynamics do not yet amplify rounding."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NSTEPS = 4


def _rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _plan(wdm, nativelib, cfg, prec):
    p = wdm.Plan(cfg["shape"], cfg["xmax"], prec, hbar_over_m=cfg["hbar_over_m"],
                 four_pi_G=cfg["four_pi_G"], lib=nativelib)
    wdm.configs.install(p, cfg)
    return p


@pytest.fixture(scope="module")
def c2(wdm):
    return wdm.configs.build("c2")


@pytest.mark.parametrize("prec,tol", [("f64", 1e-11), ("f32", 1e-5)])
def test_c2_steps_reverse_to_psi0(wdm, nativelib, c2, prec, tol):
    print(f"c2 384^3 {prec}: {NSTEPS} steps forward and back", flush=True)
    with _plan(wdm, nativelib, c2, prec) as p:
        dt, _ = wdm.configs.autoset_dt(p, c2["steps"])
        psi0 = p.get_psi()
        p.step(NSTEPS, dt, 0.5 * dt, 0.0)
        moved = _rel_l2(p.get_psi(), psi0)
        for i in reversed(range(NSTEPS)):
            t = i * dt
            p.kick(0.0, t)        # to k-space, no rotation
            p.drift(-dt)          # undo the drift: psi = K psi_i
            p.kick(-0.5 * dt, t)  # undo the kick at the same time, Phi[K psi_i] = Phi[psi_i]
            p.drift(0.0)          # back to real space
        back = p.get_psi()
    err = _rel_l2(back, psi0)
    print(f"  moved {moved:.3e}, returned to {err:.3e}", flush=True)
    assert moved > 100 * tol, moved  # the steps did change psi
    assert err < tol, err


def test_c2_mass_conserved_fp64(wdm, nativelib, c2):
    print("c2 384^3 f64: mass over 10 steps", flush=True)
    with _plan(wdm, nativelib, c2, "f64") as p:
        dt, _ = wdm.configs.autoset_dt(p, c2["steps"])
        m0 = p.reduce(1)["mass"]
        p.step(10, dt, 0.5 * dt, 0.0)
        m1 = p.reduce(1)["mass"]
    assert m0 > 0
    assert abs(m1 - m0) / m0 < 1e-12, (m0, m1)


def test_c2_fp32_tracks_fp64_short_horizon(wdm, nativelib, c2):
    print("c2 384^3: f32 against f64 after 5 steps", flush=True)
    with _plan(wdm, nativelib, c2, "f32") as p32:
        dt, _ = wdm.configs.autoset_dt(p32, c2["steps"])
        psi0 = p32.get_psi(np.complex64)
        p32.step(5, dt, 0.5 * dt, 0.0)
        out32 = p32.get_psi()
    with _plan(wdm, nativelib, c2, "f64") as p64:
        p64.set_psi(psi0)
        p64.step(5, dt, 0.5 * dt, 0.0)
        out64 = p64.get_psi()
    err = _rel_l2(out32, out64)
    print(f"  rel L2 {err:.3e}", flush=True)
    assert err < 1e-5, err
