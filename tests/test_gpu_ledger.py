"""Opt-in mass projection (wdm_set_mass_ledger) through the C ABI.

Without it, FP32 rounding makes the norm a random walk; with it the mass
deviation stays at the size of one step's rounding.  Bars: over 300 FP32
steps the projected deviation never exceeds 6 single-step increments of the
unprojected run and its RMS is below 2/3 of the unprojected RMS; states
with and without the projection differ only at FP32 rounding level
(<= 1e-5 relative L2 after 40 steps) and, for an FP64 plan, at FP64 level
(<= 1e-12); error paths are refused."""
import numpy as np
import pytest

from oracle import ssfm_oracle as O

pytestmark = pytest.mark.gpu

X = (4.0, 4.0, 4.0)
G4 = 40.0


def _psi(shape, seed=2):
    x, y, z, _, _ = O.setup_grid(*X, *shape)
    Xg, Yg, Zg = np.meshgrid(x, y, z, indexing="ij")
    rng = np.random.default_rng(seed)
    psi = np.zeros(shape, dtype=complex)
    for c in rng.uniform(-1.5, 1.5, size=(4, 3)):
        r2 = (Xg - c[0]) ** 2 + (Yg - c[1]) ** 2 + (Zg - c[2]) ** 2
        psi += 3.0 * np.exp(-r2 / 0.5 + 1j * rng.uniform(-2, 2) * Xg)
    return psi


def _run(wdm, lib, prec, ledger, nsteps, dt, split="reference", shape=(32, 32, 32), per_step=False):
    with wdm.Plan(shape, X, prec, four_pi_G=G4, lib=lib) as p:
        p.set_mass_ledger(ledger)
        p.set_psi(_psi(shape))
        m0 = p.reduce(1)["mass"]
        dev = []
        if per_step:
            for _ in range(nsteps):
                p.step(1, dt, split=split)
                dev.append(p.reduce(1)["mass"] / m0 - 1)
        else:
            p.step(nsteps, dt, split=split)
            dev.append(p.reduce(1)["mass"] / m0 - 1)
        return p.get_psi(), np.array(dev)


@pytest.mark.parametrize("split", ["reference", "strang"], ids=["kd", "kdk"])
def test_fp32_projection_bounds_the_mass_deviation(wdm, nativelib, split):
    dt, n = 2e-3, 300
    _, on = _run(wdm, nativelib, "f32", True, n, dt, split, shape=(16, 16, 16), per_step=True)
    _, off = _run(wdm, nativelib, "f32", False, n, dt, split, shape=(16, 16, 16), per_step=True)
    sigma1 = np.sqrt(np.mean(np.diff(off) ** 2))
    rms_on, rms_off = np.sqrt(np.mean(on ** 2)), np.sqrt(np.mean(off ** 2))
    print(f"{split}: FP32 16^3 {n} steps: step noise {sigma1:.2e}; |dev| max with projection "
          f"{np.abs(on).max():.2e} (rms {rms_on:.2e}), without {np.abs(off).max():.2e} (rms {rms_off:.2e})")
    assert np.abs(on[100:]).max() <= 6 * sigma1
    assert rms_on < rms_off / 1.5


def test_fp32_projection_changes_the_state_at_rounding_level(wdm, nativelib):
    psi_on, _ = _run(wdm, nativelib, "f32", True, 40, 2e-3)
    psi_off, _ = _run(wdm, nativelib, "f32", False, 40, 2e-3)
    err = np.linalg.norm(psi_on - psi_off) / np.linalg.norm(psi_off)
    assert err < 1e-5, err


def test_fp64_projection_changes_nothing_measurable(wdm, nativelib):
    psi_on, dev_on = _run(wdm, nativelib, "f64", True, 10, 2e-3)
    psi_off, _ = _run(wdm, nativelib, "f64", False, 10, 2e-3)
    assert abs(dev_on[-1]) < 1e-13
    assert np.linalg.norm(psi_on - psi_off) / np.linalg.norm(psi_off) < 1e-12


def test_projection_reference_follows_set_psi_and_kick_drift(wdm, nativelib):
    # after set_psi the new state's mass is the reference; manual kick/drift
    # pairs carry the same projection as wdm_step (FP64: exact to rounding)
    shape = (16, 16, 16)
    with wdm.Plan(shape, X, "f64", four_pi_G=G4, lib=nativelib) as p:
        p.set_mass_ledger(True)
        p.set_psi(_psi(shape))
        p.step(3, 2e-3)
        p.set_psi(2.0 * _psi(shape, seed=5))
        m0 = p.reduce(1)["mass"]
        for _ in range(5):
            p.kick(1e-3)
            p.drift(2e-3)
        m1 = p.reduce(1)["mass"]
    assert abs(m1 / m0 - 1) < 1e-13


def test_projection_error_paths(wdm, nativelib):
    from wavedm_amd._capi import WaveDMError, WDM_ERR_UNSUPPORTED
    shape = (16, 16, 16)
    with wdm.Plan(shape, X, "f32", four_pi_G=G4, lib=nativelib) as p:
        p.set_border(np.linspace(0.9, 1.0, shape[0]))
        with pytest.raises(WaveDMError) as e:
            p.set_mass_ledger(True)
        assert e.value.code == WDM_ERR_UNSUPPORTED
        p.set_border(None)
        p.set_mass_ledger(True)
        for call in (lambda: p.set_border(np.ones(shape[0])), lambda: p.set_self_gravity(False)):
            with pytest.raises(WaveDMError) as e:
                call()
            assert e.value.code == WDM_ERR_UNSUPPORTED
        p.set_psi(_psi(shape))
        with pytest.raises(WaveDMError) as e:
            p.adjoint_forward(2, 1e-3)
        assert e.value.code == WDM_ERR_UNSUPPORTED
        p.set_mass_ledger(False)
        p.adjoint_forward(2, 1e-3)


@pytest.mark.parametrize("graph", [True, False], ids=["graph", "eager"])
def test_projection_in_an_ensemble(wdm, nativelib, graph):
    # an ensemble step is one wdm_step of each member: with the projection on,
    # members equal plans stepped one step per call, bit for bit
    shapes = [(32, 32, 32), (16, 16, 16)]
    dts = [2e-3, 1.5e-3]

    def make(shape):
        p = wdm.Plan(shape, X, "f32", four_pi_G=G4, lib=nativelib)
        p.set_mass_ledger(True)
        p.set_psi(_psi(shape))
        return p

    ens = [make(s) for s in shapes]
    ref = [make(s) for s in shapes]
    try:
        m0 = [p.reduce(1)["mass"] for p in ens]
        with wdm.Ensemble(ens) as e:
            e.step(5, dts, graph=graph)
        for p, dt in zip(ref, dts):
            for k in range(5):
                p.step(1, dt, t0=k * dt)
        for a, b, m in zip(ens, ref, m0):
            assert np.array_equal(a.get_psi(np.complex64), b.get_psi(np.complex64))
            assert abs(a.reduce(1)["mass"] / m - 1) < 1e-7
    finally:
        for p in ens + ref:
            p.close()


def test_projection_enabled_between_kick_and_drift_is_inert(wdm, nativelib):
    # switched on after a kick taken without it: the drift scales by exactly 1
    shape = (16, 16, 16)
    out = []
    for on in (True, False):
        with wdm.Plan(shape, X, "f64", four_pi_G=G4, lib=nativelib) as p:
            p.set_psi(_psi(shape))
            p.kick(1e-3)
            p.set_mass_ledger(on)
            p.drift(2e-3)
            out.append(p.get_psi())
    assert np.array_equal(out[0], out[1])
