"""Device diagnostics (wdm_diagnostics, wdm_acceleration) through the C ABI
against the FP64 oracle (oracle/diag_oracle.py) evaluated on the same psi.

Bars: sums, peak, plane radius and acceleration of an FP64 plan <= 1e-11
relative (FP64 arithmetic on both sides, only the summation order differs);
the same for an FP32 plan on its own psi (the device upcasts every value to
FP64) except PE_abs and the acceleration, whose Phi comes from FP32 FFTs
(<= 1e-4 against the FP64 potential, <= 1e-11 against the device's own Phi).
Radii: within two histogram bin widths plus the gap between the cell
distances around the sorted weighted quantile (derivation in wdm_diag.inl)."""
import numpy as np
import pytest

from oracle import diag_oracle as D
from oracle import ssfm_oracle as O

pytestmark = pytest.mark.gpu

FOUR_PI_G = 4 * np.pi * 0.25
HQ = (15.0, 0.8, 0.3, -0.4, 0.2)
T0 = np.array([[3.0, 0.4, -0.2], [0.4, -2.0, 0.1], [-0.2, 0.1, 1.0]])
QCEN = (0.1, 0.0, -0.2)


def _psi(shape, X, seed):
    """Two phase-gradient blobs on a faint noisy background (no exact ties)."""
    x, y, z, _, _ = O.setup_grid(*X, *shape)
    Xg, Yg, Zg = np.meshgrid(x, y, z, indexing="ij")
    rng = np.random.default_rng(seed)
    psi = np.zeros(shape, dtype=complex)
    for amp, c, w, k in ((2.0, (0.4, -0.3, 0.2), 0.35, (1.5, -0.7, 0.4)),
                         (1.2, (-0.5, 0.4, -0.3), 0.25, (-0.8, 1.1, 0.9))):
        r2 = ((Xg - c[0] * X[0]) ** 2 + (Yg - c[1] * X[1]) ** 2 + (Zg - c[2] * X[2]) ** 2)
        psi += amp * np.exp(-r2 / (2 * (w * min(X)) ** 2) + 1j * (k[0] * Xg + k[1] * Yg + k[2] * Zg))
    psi += 1e-3 * (rng.normal(size=shape) + 1j * rng.normal(size=shape))
    return psi


def _v_host(shape, X, static, quad):
    x, y, z, _, _ = O.setup_grid(*X, *shape)
    Xg, Yg, Zg = np.meshgrid(x, y, z, indexing="ij")
    G = FOUR_PI_G / (4 * np.pi)
    M, a, cx, cy, cz = HQ
    v = -G * M / (np.sqrt((Xg - cx) ** 2 + (Yg - cy) ** 2 + (Zg - cz) ** 2) + a)
    if static is not None:
        v = v + static
    if quad:
        d = [Xg - QCEN[0], Yg - QCEN[1], Zg - QCEN[2]]
        v = v + 0.5 * sum(T0[i, j] * d[i] * d[j] for i in range(3) for j in range(3))
    return v


def _plan(wdm, lib, shape, X, prec, static, quad):
    p = wdm.Plan(shape, X, prec, four_pi_G=FOUR_PI_G, lib=lib)
    p.set_hernquist(HQ[0], HQ[1], HQ[2:])
    p.set_kappa(0.7)  # not part of Phi_all: must not change PE_abs
    if static is not None:
        p.set_static_potential(static)
    if quad:
        p.set_quadrupole(T0, (0.0, 0.0, 1.0), 0.0, QCEN)  # period 0: no rotation
    return p


def _radius_tol(psi, X, shape, q_or):
    """2 bin widths + the gap between the cell distances around each quantile."""
    x, y, z, _, _ = O.setup_grid(*X, *shape)
    rho = np.abs(psi) ** 2
    ix, iy, iz = np.unravel_index(int(np.argmax(rho.ravel(order="F"))), shape, order="F")
    Xg, Yg, Zg = np.meshgrid(x, y, z, indexing="ij")
    rc = np.sqrt((Xg - x[ix]) ** 2 + (Yg - y[iy]) ** 2 + (Zg - z[iz]) ** 2).ravel()
    far = np.sqrt(sum(max(abs(c[i] - c[0]), abs(c[-1] - c[i])) ** 2 for c, i in ((x, ix), (y, iy), (z, iz))))
    width = far / 4096
    v = np.unique(rc)
    tol = []
    for q in q_or:
        k = int(np.searchsorted(v, q))
        lo, hi = v[max(k - 2, 0)], v[min(k + 1, len(v) - 1)]
        tol.append(2 * width + (hi - lo))
    return np.array(tol)


def _check(dev, ref, psi, X, shape, pe_tol, sharp=True):
    rel = lambda a, b: abs(a - b) / max(abs(b), 1e-300)  # noqa: E731
    assert rel(dev["TotalMass"], ref["mass"]) < 1e-12
    assert rel(dev["rho_max"], ref["rho_max"]) < 1e-13
    pk = np.unravel_index(ref["peak_index"], shape, order="F")
    assert (dev["peak_ix"], dev["peak_iy"], dev["peak_iz"]) == tuple(int(v) for v in pk)
    assert rel(dev["R"], ref["r_std_mid_plane"]) < 1e-11
    mscale = max(abs(v) for v in ref["momentum"]) + 1e-300
    for c, k in enumerate(("MomentumX", "MomentumY", "MomentumZ")):
        assert abs(dev[k] - ref["momentum"][c]) < 1e-11 * mscale, (k, dev[k], ref["momentum"][c])
    assert rel(dev["KE"], ref["kinetic"]) < 1e-11
    assert rel(dev["QE"], ref["quantum"]) < 1e-11
    assert rel(dev["PE_abs"], ref["virial_potential"]) < pe_tol, (dev["PE_abs"], ref["virial_potential"])
    assert dev["Virial"] == 2 * dev["KE"] + 2 * dev["QE"] - dev["PE_abs"]
    radii = np.array([dev[f"R{i}"] for i in range(1, 10)])
    tol = _radius_tol(psi, X, shape, ref["radii"])
    err = np.abs(radii - ref["radii"])
    assert np.all(err <= tol), (radii, ref["radii"], tol)
    if sharp:  # the bound tells neighbouring deciles apart (grids of >= 10^4 cells)
        assert np.all(tol[:-1] < 0.5 * np.diff(ref["radii"])), (tol, ref["radii"])


CASES = [((32, 24, 16), (3.0, 2.5, 2.0), True, True), ((64, 32, 48), (4.0, 3.0, 3.5), False, True),
         ((16, 12, 8), (2.0, 1.5, 1.0), True, False),
         # rows longer than one wave (x neighbours across 64-column chunks, a partial chunk)
         # and a y extent that is not a multiple of the 32-row walk
         ((96, 48, 12), (4.0, 3.0, 2.0), True, True)]
IDS = ["x".join(map(str, c[0])) for c in CASES]


@pytest.mark.parametrize("shape,X,with_static,quad", CASES, ids=IDS)
def test_diagnostics_match_oracle(wdm, nativelib, shape, X, with_static, quad):
    psi = _psi(shape, X, seed=sum(shape))
    rng = np.random.default_rng(7)
    static = rng.normal(size=shape) * 0.3 if with_static else None
    x, y, z, d, _ = O.setup_grid(*X, *shape)
    phi_ref = np.real(O.compute_gravitational_potential(psi, d, FOUR_PI_G)) + _v_host(shape, X, static, quad)
    for prec in ("f64", "f32"):
        with _plan(wdm, nativelib, shape, X, prec, static, quad) as p:
            p.set_psi(psi)
            pd = p.get_psi()  # the plan's own values (rounded to FP32 for the f32 plan)
            dev = p.diagnostics()
            again = p.diagnostics()
            phi_dev = p.get_phi() + _v_host(shape, X, static, quad)
            acc = p.acceleration()
        assert all(dev[k] == again[k] or (np.isnan(dev[k]) and np.isnan(again[k])) for k in dev)  # deterministic
        ref = D.diagnostics(pd, phi_ref, x, y, z)
        _check(dev, ref, pd, X, shape, 1e-10 if prec == "f64" else 1e-4, sharp=np.prod(shape) >= 10 ** 4)
        own = D.diagnostics(pd, phi_dev, x, y, z)
        assert abs(dev["PE_abs"] - own["virial_potential"]) < 1e-11 * abs(own["virial_potential"])
        for a in range(3):
            ref_a = -D.central_diff(phi_dev, a, d[a])
            scale = np.abs(ref_a).max()
            assert np.abs(acc[a] - ref_a).max() < 1e-11 * scale, (prec, a)
            if prec == "f64":
                assert np.abs(acc[a] + D.central_diff(phi_ref, a, d[a])).max() < 1e-10 * scale


def test_diagnostics_after_steps_and_short_output(wdm, nativelib):
    shape, X = (32, 24, 16), (3.0, 2.5, 2.0)
    psi = _psi(shape, X, seed=11)
    x, y, z, d, _ = O.setup_grid(*X, *shape)
    with _plan(wdm, nativelib, shape, X, "f64", None, False) as p:
        p.set_psi(psi)
        p.step(5, 2e-3)
        full = p.diagnostics()
        short = p.diagnostics(6)  # stops before the potential: Phi is not computed
        pd = p.get_psi()
        phi = p.get_phi() + _v_host(shape, X, None, False)
    assert list(short) == list(p.DIAG_KEYS[:6])
    for k in short:
        assert short[k] == full[k]
    ref = D.diagnostics(pd, phi, x, y, z)
    _check(full, ref, pd, X, shape, 1e-11)


def test_diagnostics_error_paths_and_empty_field(wdm, nativelib):
    from wavedm_amd._capi import WaveDMError, WDM_ERR_STATE, WDM_ERR_ARG
    shape = (16, 12, 8)
    with wdm.Plan(shape, 2.0, "f64", lib=nativelib) as p:
        p.set_psi(_psi(shape, (2.0, 2.0, 2.0), seed=1))
        with pytest.raises(WaveDMError) as e:
            p.diagnostics(0)
        assert e.value.code == WDM_ERR_ARG
        p.kick(1e-3)  # psi now holds the spectrum
        for call in (p.diagnostics, p.acceleration):
            with pytest.raises(WaveDMError) as e:
                call()
            assert e.value.code == WDM_ERR_STATE
        p.drift(1e-3)
        p.diagnostics()
        # psi = 0: no weights (the reference's quantile raises), KE = 0/0 as in the reference
        p.set_psi(np.zeros(shape, dtype=complex))
        z = p.diagnostics()
    assert z["TotalMass"] == 0.0 and z["rho_max"] == 0.0
    assert all(np.isnan(z[f"R{i}"]) for i in range(1, 10))
    assert np.isnan(z["KE"]) and z["QE"] == 0.0
