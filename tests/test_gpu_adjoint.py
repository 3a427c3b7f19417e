"""Device adjoint (wdm_adjoint_forward / wdm_adjoint_backward) through the C ABI
against the FP64 CPU adjoint (oracle/adjoint_oracle.py, itself pinned by
central differences) and against central differences of the device forward.

Bars: FP64 plan <= 1e-9 relative to the oracle; FP32 plan <= 2e-3; device FP64
gradient vs device FP64 central differences <= 1e-4 (h = 1e-3 relative, the
configured check; Richardson-extrapolated for the dt scale)."""
import numpy as np
import pytest

from oracle import adjoint_oracle as AO

pytestmark = pytest.mark.gpu

X3 = (3.0, 2.5, 3.5)
HQ = (20.0, 1.3, 0.3, -0.2, 0.4)
FOUR_PI_G = 4 * np.pi * 0.3


def _problem(shape, border, kappa, dt=0.05):
    b = np.linspace(0.9, 1.0, shape[0]) if border else None
    return AO.Problem(shape, X3, dt=dt, hernquist=HQ, kappa=kappa, border=b, four_pi_G=FOUR_PI_G)


def _psi0(shape, seed=3):
    rng = np.random.default_rng(seed)
    return (rng.normal(size=shape) + 1j * rng.normal(size=shape)) * 0.5 + 1.0


def _plan(wdm, lib, pr, prec):
    p = wdm.Plan(pr.shape, X3, prec, four_pi_G=FOUR_PI_G, lib=lib)
    p.set_hernquist(HQ[0], HQ[1], HQ[2:])
    if pr.kappa:
        p.set_kappa(pr.kappa)
    if pr.border is not None:
        p.set_border(pr.border)
    return p


CASES = [((16, 12, 8), True, 0.7), ((32, 16, 8), False, 0.0), ((12, 8, 16), True, 0.0)]
CASE_IDS = ["x".join(map(str, c[0])) + ("-border" if c[1] else "") + ("-kappa" if c[2] else "") for c in CASES]


@pytest.mark.parametrize("shape,border,kappa", CASES, ids=CASE_IDS)
def test_adjoint_matches_oracle(wdm, nativelib, shape, border, kappa):
    pr = _problem(shape, border, kappa)
    psi0 = _psi0(shape)
    nsteps = 7
    loss = wdm.adjoint.ProfileLoss(shape, X3, nbins=8)
    target = loss.profile(pr.forward(psi0, nsteps)) * 1.2
    L_ref, gN = loss(pr.forward(psi0, nsteps), target)
    g_ref, p_ref = pr.backward(psi0, nsteps, gN)
    for prec, tol in (("f64", 1e-9), ("f32", 2e-3)):
        with _plan(wdm, nativelib, pr, prec) as p:
            L, g0, gp = wdm.adjoint.value_and_grad(p, psi0, nsteps, pr.dt, loss, target, ckpt_every=3)
        assert abs(L - L_ref) <= tol * abs(L_ref), (prec, L, L_ref)
        err = np.linalg.norm(g0 - g_ref) / np.linalg.norm(g_ref)
        assert err < tol, (prec, err)
        perr = np.abs(gp - p_ref) / np.maximum(np.abs(p_ref), 1e-30)
        assert np.all(perr < tol * 10 if prec == "f32" else perr < tol), (prec, gp, p_ref)


QUAD = ((30.0, -20.0, 10.0, 4.0, -2.0, 1.0), (0.2, 0.5, 1.0), 0.4, (0.1, 0.0, -0.2))


@pytest.mark.parametrize("shape", [(16, 12, 8), (12, 8, 16)], ids=["16x12x8", "12x8x16"])
def test_quadrupole_gradients_match_oracle(wdm, nativelib, shape):
    # the rotating tidal quadrupole's T0 (grad_params 6..11) with the Hernquist
    # host, kicks at t0 + n dt; dt scale off (the potential is time-dependent)
    t0 = 0.03
    pr = AO.Problem(shape, X3, dt=0.05, hernquist=HQ, four_pi_G=FOUR_PI_G, quadrupole=QUAD, t0=t0)
    psi0 = _psi0(shape, seed=7)
    nsteps = 6
    loss = wdm.adjoint.ProfileLoss(shape, X3, nbins=8)
    target = loss.profile(pr.forward(psi0, nsteps)) * 1.2
    L_ref, gN = loss(pr.forward(psi0, nsteps), target)
    g_ref, p_ref = pr.backward(psi0, nsteps, gN)
    sel = [0, 1, 2, 3, 4, 6, 7, 8, 9, 10, 11]
    assert np.all(np.abs(p_ref[6:]) > 0)
    for prec, tol in (("f64", 1e-9), ("f32", 2e-3)):
        with _plan(wdm, nativelib, pr, prec) as p:
            T0v, axis, period, cen = QUAD
            p.set_quadrupole(T0v, axis, period, cen)
            L, g0, gp = wdm.adjoint.value_and_grad(p, psi0, nsteps, pr.dt, loss, target, t0=t0, ckpt_every=4,
                                                   dt_scale=False)
        assert abs(L - L_ref) <= tol * abs(L_ref), (prec, L, L_ref)
        err = np.linalg.norm(g0 - g_ref) / np.linalg.norm(g_ref)
        assert err < tol, (prec, err)
        perr = np.abs(gp[sel] - p_ref[sel]) / np.abs(p_ref[sel])
        assert np.all(perr < (tol * 10 if prec == "f32" else tol)), (prec, gp, p_ref)
        assert np.isnan(gp[5])


def test_checkpoint_interval_does_not_change_the_gradient(wdm, nativelib):
    pr = _problem((16, 12, 8), True, 0.7)
    psi0 = _psi0(pr.shape, seed=5)
    loss = wdm.adjoint.ProfileLoss(pr.shape, X3, nbins=8)
    target = loss.profile(pr.forward(psi0, 6)) * 0.8
    res = []
    with _plan(wdm, nativelib, pr, "f64") as p:
        for ck in (1, 4, 6, 50):
            res.append(wdm.adjoint.value_and_grad(p, psi0, 6, pr.dt, loss, target, ckpt_every=ck))
    for L, g, gp in res[1:]:
        assert L == res[0][0]
        assert np.array_equal(g, res[0][1]) and np.array_equal(gp, res[0][2])


def test_device_gradient_matches_device_central_differences(wdm, nativelib):
    # the configured check: h = 1e-3 relative central differences of the
    # device forward, FP64 plan, random psi0 direction and every parameter
    shape = (32, 32, 32)
    pr = _problem(shape, False, 0.0, dt=0.02)
    psi0 = _psi0(shape, seed=7)
    nsteps = 10
    loss = wdm.adjoint.ProfileLoss(shape, X3, nbins=12)
    with _plan(wdm, nativelib, pr, "f64") as p:
        p.set_psi(psi0)
        p.step(nsteps, pr.dt)
        target = loss.profile(p.get_psi()) * 1.1
        L, g0, gp = wdm.adjoint.value_and_grad(p, psi0, nsteps, pr.dt, loss, target, ckpt_every=4)

        def fwd(psi, hq=HQ, sigma=1.0):
            p.set_hernquist(hq[0], hq[1], hq[2:])
            p.set_psi(psi)
            p.step(nsteps, sigma * pr.dt)
            p.clear_potential_models()
            return loss(p.get_psi(), target)[0]

        p.clear_potential_models()
        rng = np.random.default_rng(11)
        d = rng.normal(size=shape) + 1j * rng.normal(size=shape)
        h = 1e-3 * np.linalg.norm(psi0) / np.linalg.norm(d)
        fd = (fwd(psi0 + h * d) - fwd(psi0 - h * d)) / (2 * h)
        ad = float(np.sum(np.real(np.conj(g0) * d)))
        assert abs(fd - ad) <= 1e-4 * abs(fd), ("psi0", fd, ad)
        hq = np.array(HQ)
        for j in range(5):
            e = np.zeros(5)
            e[j] = 1e-3 * max(abs(hq[j]), 1.0)
            fd = (fwd(psi0, tuple(hq + e)) - fwd(psi0, tuple(hq - e))) / (2 * e[j])
            assert abs(fd - gp[j]) <= 1e-4 * abs(fd), (AO.PARAM_NAMES[j], fd, gp[j])
        # sigma multiplies drift phases of up to ~80 rad over the run, so the
        # O(h^2) truncation of a plain h = 1e-3 central difference is ~1e-3;
        # Richardson-extrapolate the h and h/2 differences (O(h^4))
        d1 = (fwd(psi0, sigma=1 + 1e-3) - fwd(psi0, sigma=1 - 1e-3)) / 2e-3
        d2 = (fwd(psi0, sigma=1 + 5e-4) - fwd(psi0, sigma=1 - 5e-4)) / 1e-3
        fd = (4 * d2 - d1) / 3
        assert abs(fd - gp[5]) <= 1e-4 * abs(fd), ("sigma", fd, d1, gp[5])


def test_adjoint_error_paths(wdm, nativelib):
    pr = _problem((16, 12, 8), False, 0.0)
    with _plan(wdm, nativelib, pr, "f32") as p:
        g = np.zeros(pr.shape, np.complex128)
        with pytest.raises(wdm.WaveDMError) as e:
            p.adjoint_backward(g)
        assert e.value.code == wdm._capi.WDM_ERR_STATE
        with pytest.raises(wdm.WaveDMError):
            p.adjoint_forward(0, 0.1)
        p.set_psi(_psi0(pr.shape))
        p.set_quadrupole(np.diag([1.0, 2.0, -3.0]), (0, 0, 1), 1.0)
        p.adjoint_forward(3, 0.01)
        with pytest.raises(wdm.WaveDMError) as e:
            p.adjoint_backward(g, dt_scale=True)
        assert e.value.code == wdm._capi.WDM_ERR_UNSUPPORTED
        g0, gp = p.adjoint_backward(g, dt_scale=False)
        assert np.all(g0 == 0) and np.isnan(gp[5])


def test_adjoint_tape_refuses_changed_plan(wdm, nativelib):
    """The backward sweep recomputes its segments with the plan's current
    models and restores psi_N at the end: a parameter change, a new psi or a
    step between the taped forward and the backward sweep is refused with
    WDM_ERR_STATE instead of returning the gradient of another trajectory.
    A second backward on an unchanged plan repeats the first exactly."""
    pr = _problem((16, 12, 8), False, 0.0)
    psi0 = _psi0(pr.shape)
    rng = np.random.default_rng(5)
    g = rng.normal(size=pr.shape) + 1j * rng.normal(size=pr.shape)
    with _plan(wdm, nativelib, pr, "f64") as p:
        p.set_psi(psi0)
        p.adjoint_forward(4, pr.dt, ckpt_every=2)
        g0a, gpa = p.adjoint_backward(g)
        g0b, gpb = p.adjoint_backward(g)   # same tape, unchanged plan
        assert np.array_equal(g0a, g0b) and np.array_equal(gpa, gpb, equal_nan=True)
        for change in ("hernquist", "coeffs", "set_psi", "step"):
            p.set_psi(psi0)
            p.adjoint_forward(4, pr.dt, ckpt_every=2)
            if change == "hernquist":
                p.set_hernquist(HQ[0] * 1.1, HQ[1], HQ[2:])
            elif change == "coeffs":
                p.set_coeffs(1.0, FOUR_PI_G * 1.01)
            elif change == "set_psi":
                p.set_psi(psi0 * 1.01)
            else:
                p.step(1, pr.dt)
            with pytest.raises(wdm.WaveDMError) as e:
                p.adjoint_backward(g)
            assert e.value.code == wdm._capi.WDM_ERR_STATE, change
            assert "tape" in str(e.value), change
            # restore the configuration for the next case
            p.clear_potential_models()
            p.set_hernquist(HQ[0], HQ[1], HQ[2:])
            p.set_coeffs(1.0, FOUR_PI_G)
        # after re-taping on the restored plan the gradient is the first one again
        p.set_psi(psi0)
        p.adjoint_forward(4, pr.dt, ckpt_every=2)
        g0c, gpc = p.adjoint_backward(g)
        assert np.array_equal(g0a, g0c) and np.array_equal(gpa, gpc, equal_nan=True)
