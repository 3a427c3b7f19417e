"""The opt-in Milky Way tidal potential against the oracle
(oracle/tidal_oracle.py, the intent of src/physics/tidal_fields.jl:64-146),
evaluated every kick through run_ssfm's host V: a smooth table whose grid
the trajectory offsets partly leave (clamping), three trajectory rows crossed
during the run, an off-centre density peak (unique, so the FP32 plan finds the
same cell).  Bars: FP64 plan <= 1e-12 relative L2, FP32 plan <= 1e-5."""
import numpy as np
import pytest

from oracle import ssfm_oracle as O
from oracle import tidal_oracle as T

pytestmark = pytest.mark.gpu

X3 = (4.0, 3.5, 3.0)
LO, HI, TN = (-8.0, -7.0, -6.5), (8.0, 7.5, 6.0), (17, 13, 11)
TRAJ_T = np.array([0.0, 0.005, 0.009])
TRAJ_XYZ = np.array([[0.5, -0.3, 0.2], [1.5, 0.4, -0.6], [5.0, -1.0, 0.8]])


def _table():
    axes = [np.linspace(LO[a], HI[a], TN[a]) for a in range(3)]
    A = np.meshgrid(*axes, indexing="ij")
    r = np.sqrt((A[0] + 3.0) ** 2 + 1.3 * (A[1] - 1.0) ** 2 + 0.8 * (A[2] - 0.5) ** 2 + 4.0)
    return axes, -60.0 / r + 0.05 * A[0] * A[1]


def _psi(shape):
    rng = np.random.default_rng(7)
    x, y, z, _, _ = O.setup_grid(*X3, *shape)
    r2 = (x[:, None, None] - 0.37) ** 2 + (y[None, :, None] + 0.21) ** 2 + (z[None, None, :] - 0.13) ** 2
    return np.asfortranarray(20.0 * np.exp(-r2 / 1.5) * np.exp(1j * rng.uniform(-0.3, 0.3, shape)))


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("shape", [(32, 24, 16), (24, 16, 48)], ids=["32x24x16", "24x16x48"])
def test_tidal_kicks_match_oracle(wdm, nativelib, shape):
    axes, table = _table()
    psi = _psi(shape)
    nsteps, dt = 6, 2e-3

    def V(xxx, yyy, zzz, ps, t):
        x, y, z = xxx[:, 0, 0], yyy[0, :, 0], zzz[0, 0, :]
        o = T.trajectory_offset(TRAJ_T, TRAJ_XYZ, t)
        return T.tidal_potential(axes, table, x, y, z, o, np.abs(ps) ** 2)

    ref = O.run_ssfm(psi, X3, nsteps, dt, V=V)
    assert _rel(O.run_ssfm(psi, X3, nsteps, dt), ref) > 1e-4  # the term matters
    print(f"tidal {shape}: FP64 and FP32 plans, {nsteps} steps", flush=True)
    for prec, tol in (("f64", 1e-12), ("f32", 1e-5)):
        with wdm.Plan(shape, X3, prec, lib=nativelib) as p:
            p.set_psi(psi)
            p.set_tidal(table, LO, HI, TRAJ_T, TRAJ_XYZ)
            p.step(nsteps, dt)
            out = p.get_psi()
            err = _rel(out, ref)
            print(f"  {prec}: rel L2 {err:.3e}", flush=True)
            assert err < tol, (prec, err)
            # cleared: the plan steps as without the term
            p.set_psi(psi)
            p.clear_potential_models()
            p.step(nsteps, dt)
            assert _rel(p.get_psi(), O.run_ssfm(psi, X3, nsteps, dt)) < tol


def test_tidal_refusals(wdm, nativelib):
    axes, table = _table()
    with wdm.Plan((16, 16, 16), X3, "f64", lib=nativelib) as p:
        p.set_psi(_psi((16, 16, 16)))
        with pytest.raises(wdm.WaveDMError, match="hi must exceed lo"):
            p.set_tidal(table, HI, LO, TRAJ_T, TRAJ_XYZ)
        with pytest.raises(wdm.WaveDMError, match="must not decrease"):
            p.set_tidal(table, LO, HI, TRAJ_T[::-1], TRAJ_XYZ)
        p.set_tidal(table, LO, HI, TRAJ_T, TRAJ_XYZ)
        with pytest.raises(wdm.WaveDMError, match="tidal"):
            p.adjoint_forward(2, 1e-3)
        p.set_tidal(None)
        p.step(1, 1e-3)
