"""Pins of the FP64 CPU adjoint (oracle/adjoint_oracle.py) by central finite
differences of its own forward map, and of the package's profile loss against
the oracle's.  The reference defines no adjoint, so the FD derivative of the
reference step (restated in ssfm_oracle) is the only truth available."""
import numpy as np
import pytest

from oracle import adjoint_oracle as AO
from oracle import ssfm_oracle as O


def _setup(shape=(8, 6, 10), border=True, kappa=0.7, seed=1):
    rng = np.random.default_rng(seed)
    X = (3.0, 2.5, 3.5)
    pr = AO.Problem(shape, X, dt=0.05, hernquist=(20.0, 1.3, 0.3, -0.2, 0.4), kappa=kappa,
                    border=np.linspace(0.9, 1.0, shape[0]) if border else None, four_pi_G=4 * np.pi * 0.3)
    psi0 = (rng.normal(size=shape) + 1j * rng.normal(size=shape)) * 0.5 + 1.0
    idx = AO.radial_bins(pr, nbins=6)
    return pr, psi0, idx, rng


def _target(pr, psi0, idx, nsteps):
    psiN = pr.forward(psi0, nsteps)
    sel = idx >= 0
    cnt = np.maximum(np.bincount(idx[sel], minlength=6), 1)
    return np.bincount(idx[sel], weights=(np.abs(psiN) ** 2)[sel], minlength=6) / cnt * 1.3


def test_forward_equals_reference_step():
    # the adjoint's forward map is the oracle's reference-split loop
    pr, psi0, _, _ = _setup(border=False, kappa=0.0)
    M, a, cx, cy, cz = pr.hernquist

    def V(xxx, yyy, zzz, ps, t):
        return -pr.G * M / (np.sqrt((xxx - cx) ** 2 + (yyy - cy) ** 2 + (zzz - cz) ** 2) + a)

    ref = O.run_ssfm(psi0, pr.Xmax, 4, pr.dt, hbar_over_m=1.0, four_pi_G=pr.four_pi_G, V=V)
    got = pr.forward(psi0, 4)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-13


@pytest.mark.parametrize("border,kappa", [(True, 0.7), (False, 0.0)])
def test_adjoint_matches_central_differences(border, kappa):
    pr, psi0, idx, rng = _setup(border=border, kappa=kappa)
    nsteps = 5
    target = _target(pr, psi0, idx, nsteps)
    L, gN = AO.profile_loss(pr.forward(psi0, nsteps), idx, target, 6)
    g0, gp = pr.backward(psi0, nsteps, gN)
    loss = lambda **kw: AO.profile_loss(pr.forward(kw.pop("psi", psi0), nsteps, **kw), idx, target, 6)[0]  # noqa
    d = rng.normal(size=psi0.shape) + 1j * rng.normal(size=psi0.shape)
    h = 1e-6
    fd = (loss(psi=psi0 + h * d) - loss(psi=psi0 - h * d)) / (2 * h)
    ad = float(np.sum(np.real(np.conj(g0) * d)))
    assert abs(fd - ad) < 1e-6 * abs(fd), (fd, ad)
    hq = np.array(pr.hernquist)
    for j in range(5):
        e = np.zeros(5)
        e[j] = 1e-6 * max(1.0, abs(hq[j]))
        fd = (loss(hq=tuple(hq + e)) - loss(hq=tuple(hq - e))) / (2 * e[j])
        assert abs(fd - gp[j]) < 1e-5 * abs(fd), (AO.PARAM_NAMES[j], fd, gp[j])
    fd = (loss(sigma=1 + 1e-6) - loss(sigma=1 - 1e-6)) / 2e-6
    assert abs(fd - gp[5]) < 1e-6 * abs(fd), ("sigma", fd, gp[5])


def test_plant_dropped_poisson_term_is_caught():
    # the FD pin must notice a missing linearised-gravity term
    pr, psi0, idx, rng = _setup(border=False, kappa=0.0)
    nsteps = 4
    target = _target(pr, psi0, idx, nsteps)
    _, gN = AO.profile_loss(pr.forward(psi0, nsteps), idx, target, 6)
    good, _ = pr.backward(psi0, nsteps, gN)
    pr.op_G = lambda s: 0.0 * s
    bad, _ = pr.backward(psi0, nsteps, gN)
    assert np.linalg.norm(bad - good) > 1e-3 * np.linalg.norm(good)


def test_package_profile_loss_matches_oracle(wdm):
    pr, psi0, _, _ = _setup(shape=(16, 12, 8))
    pl = wdm.adjoint.ProfileLoss(pr.shape, pr.Xmax, nbins=10)
    idx = AO.radial_bins(pr, nbins=10)
    assert np.array_equal(pl.idx, idx)
    t = np.linspace(1.0, 2.0, 10)
    L1, g1 = pl(psi0, t)
    L2, g2 = AO.profile_loss(psi0, idx, t, 10)
    assert abs(L1 - L2) <= 1e-14 * abs(L2)
    assert np.allclose(g1, g2, rtol=1e-14, atol=0)


QUAD = ((30.0, -20.0, 10.0, 4.0, -2.0, 1.0), (0.2, 0.5, 1.0), 0.4, (0.1, 0.0, -0.2))


def _setup_quad(t0=0.03):
    pr, psi0, idx, rng = _setup(border=False, kappa=0.0)
    pr = AO.Problem(pr.shape, pr.Xmax, dt=pr.dt, hernquist=pr.hernquist, four_pi_G=pr.four_pi_G,
                    quadrupole=QUAD, t0=t0)
    return pr, psi0, idx, rng


def test_forward_with_rotating_quadrupole_equals_reference_step():
    # kick n at t0 + n dt with V_Q(t) = 1/2 (x-c)^T R T0 R^T (x-c), as run_ssfm's host V
    pr, psi0, _, _ = _setup_quad()
    M, a, cx, cy, cz = pr.hernquist
    T0v, axis, period, cen = QUAD
    T0 = np.array([[T0v[0], T0v[3], T0v[4]], [T0v[3], T0v[1], T0v[5]], [T0v[4], T0v[5], T0v[2]]])

    def V(xxx, yyy, zzz, ps, t):
        R = AO.rotation(axis, 2 * np.pi * t / period)
        T = R @ T0 @ R.T
        d = [xxx - cen[0], yyy - cen[1], zzz - cen[2]]
        vq = 0.5 * sum(T[i, j] * d[i] * d[j] for i in range(3) for j in range(3))
        return vq - pr.G * M / (np.sqrt((xxx - cx) ** 2 + (yyy - cy) ** 2 + (zzz - cz) ** 2) + a)

    ref = O.run_ssfm(psi0, pr.Xmax, 4, pr.dt, hbar_over_m=1.0, four_pi_G=pr.four_pi_G, V=V, t0=pr.t0)
    got = pr.forward(psi0, 4)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-13
    # the rotation matters over these kicks
    assert np.linalg.norm(AO.Problem(pr.shape, pr.Xmax, dt=pr.dt, hernquist=pr.hernquist,
                                     four_pi_G=pr.four_pi_G, quadrupole=(T0v, axis, 0.0, cen),
                                     t0=pr.t0).forward(psi0, 4) - ref) > 1e-4 * np.linalg.norm(ref)


def test_quadrupole_gradients_match_central_differences():
    pr, psi0, idx, rng = _setup_quad()
    nsteps = 5
    target = _target(pr, psi0, idx, nsteps)
    L, gN = AO.profile_loss(pr.forward(psi0, nsteps), idx, target, 6)
    g0, gp = pr.backward(psi0, nsteps, gN)
    T0 = np.array(QUAD[0])
    for j in range(6):
        e = np.zeros(6)
        e[j] = 1e-5 * max(1.0, abs(T0[j]))
        lp = AO.profile_loss(pr.forward(psi0, nsteps, T0=tuple(T0 + e)), idx, target, 6)[0]
        lm = AO.profile_loss(pr.forward(psi0, nsteps, T0=tuple(T0 - e)), idx, target, 6)[0]
        fd = (lp - lm) / (2 * e[j])
        assert abs(fd - gp[6 + j]) < 1e-5 * abs(fd), (AO.PARAM_NAMES[6 + j], fd, gp[6 + j])
    # the Hernquist gradients are unchanged by the extra model
    hq = np.array(pr.hernquist)
    e = np.zeros(5)
    e[0] = 1e-6 * hq[0]
    lp = AO.profile_loss(pr.forward(psi0, nsteps, hq=tuple(hq + e)), idx, target, 6)[0]
    lm = AO.profile_loss(pr.forward(psi0, nsteps, hq=tuple(hq - e)), idx, target, 6)[0]
    assert abs((lp - lm) / (2 * e[0]) - gp[0]) < 1e-5 * abs(gp[0])
