"""Pins for the FP64 CPU oracle (oracle/ssfm_oracle.py).

Each test checks the oracle against something it does not itself restate:
the reference's own unit-test assertions (test/unit_test.jl,
test/test_parallel.jl), brute-force DFT sums, closed-form eigenmodes, exact
invariants (norm, reversibility, lattice symmetry) and the harmonic-oscillator
ground state.  A plausible slip (a sign, a missing 1/2, a transposed axis, the
wrong k-period) fails at least one of them.
"""
import math

import numpy as np
import pytest

from oracle import ssfm_oracle as O


# ---------------------------------------------------------------- geometry
def test_setup_grid_reference_unit_test():
    # test/unit_test.jl:1-24
    X = 2.0
    x, y, z, d, vol = O.setup_grid(X, X, X, 64, 64, 64)
    assert math.isclose(x.min(), -X) and math.isclose(x.max(), X)
    assert len(x) == len(y) == len(z) == 64
    assert len(d) == 3
    for k in range(3):
        assert math.isclose(d[k], 2 * X / 63, rel_tol=1e-14)
    assert math.isclose(vol, np.prod(d))
    x8, _, _, d8, _ = O.setup_grid(1.0, 1.0, 1.0, 8, 8, 8)
    assert len(x8) == 8 and math.isclose(d8[0], 2.0 / 7)


def test_setup_coordinates_reference_unit_test():
    # test/unit_test.jl:26-52
    X, n = 2.0, 9
    x = O.linrange(-X, X, n)
    xxx, yyy, zzz, r = O.setup_coordinates(x, x, x)
    assert xxx.shape == yyy.shape == zzz.shape == r.shape == (n, n, n)
    c = n // 2
    assert abs(xxx[c, c, c]) < 1e-15 and abs(yyy[c, c, c]) < 1e-15
    assert abs(r[c, c, c]) < 1e-10 and np.all(r >= 0)
    assert np.array_equal(xxx[:, 0, 0], x)
    assert np.array_equal(yyy[0, :, 0], x)
    assert np.array_equal(zzz[0, 0, :], x)


def test_linrange_endpoints_and_symmetry():
    for n in (2, 3, 8, 63, 64, 384):
        v = O.linrange(-3.7, 3.7, n)
        assert v[0] == -3.7 and v[-1] == 3.7
        np.testing.assert_allclose(v + v[::-1], 0.0, atol=4e-16 * 3.7)


# ---------------------------------------------------------------- FFT convention
def _dft_matrix(n, sign):
    j = np.arange(n)
    return np.exp(sign * 2j * np.pi * np.outer(j, j) / n)


@pytest.mark.parametrize("shape", [(2, 3, 4), (5, 6, 8), (12, 1, 3), (16, 4, 1)])
def test_fft_matches_bruteforce_dft(shape):
    rng = np.random.default_rng(1)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    Fx, Fy, Fz = (_dft_matrix(n, -1) for n in shape)
    ref = np.einsum("ia,jb,kc,abc->ijk", Fx, Fy, Fz, a)
    np.testing.assert_allclose(O._NUMPY.fftn(a), ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    inv = np.einsum("ia,jb,kc,abc->ijk", *(_dft_matrix(n, +1) for n in shape), a) / a.size
    np.testing.assert_allclose(O._NUMPY.ifftn(a), inv, rtol=0, atol=1e-12 * np.abs(inv).max())


# ---------------------------------------------------------------- drift factor
@pytest.mark.parametrize("n", [8, 12, 16])
def test_shifted_frequency_is_integer_lattice_for_even_n(n):
    X = 3.3
    f = O._shifted_freq_1d(n, X)
    np.testing.assert_allclose(f * 2 * X, O.fft_order_int(n), atol=1e-12)


def test_shifted_frequency_odd_n_quirk():
    # LinRange(-N/4X, N/4X-1/2X, N) for odd N sits on a half-integer lattice;
    # fftshift (circshift by N div 2) then starts at +1/2: [0.5,1.5,-2.5,-1.5,-0.5].
    f = O._shifted_freq_1d(5, 1.0) * 2.0
    np.testing.assert_allclose(f, [0.5, 1.5, -2.5, -1.5, -0.5], atol=1e-14)


@pytest.mark.parametrize("hbar_over_m", [1.0, 19.60695])
def test_drift_plane_wave_eigenmodes(hbar_over_m):
    nx, ny, nz = 8, 6, 4
    X, Y, Z = 2.0, 1.5, 3.0
    dt = 0.013
    lp = O.setup_fft_operators(X, Y, Z, nx, ny, nz, dt, hbar_over_m)
    j = np.arange(nx)[:, None, None]
    k = np.arange(ny)[None, :, None]
    l = np.arange(nz)[None, None, :]
    for (a, b, c) in [(0, 0, 0), (1, 0, 0), (3, 2, 1), (4, 3, 2), (7, 5, 3)]:
        psi = np.exp(2j * np.pi * (a * j / nx + b * k / ny + c * l / nz))
        out = O.apply_drift_step(O._NUMPY.fftn(psi), lp, None)
        ma, mb, mc = (O.fft_order_int(nx)[a], O.fft_order_int(ny)[b], O.fft_order_int(nz)[c])
        k2 = (2 * np.pi * ma / (2 * X)) ** 2 + (2 * np.pi * mb / (2 * Y)) ** 2 + (2 * np.pi * mc / (2 * Z)) ** 2
        expect = psi * np.exp(-1j * hbar_over_m * k2 * dt / 2)
        np.testing.assert_allclose(out, expect, atol=1e-13)


# ---------------------------------------------------------------- Poisson
def test_poisson_single_mode_closed_form():
    n = (8, 12, 16)
    d = (0.3, 0.25, 0.2)
    Narg = (7, 11, 15)  # gravity.jl:23 passes N-1
    j = np.arange(n[0])[:, None, None]
    k = np.arange(n[1])[None, :, None]
    l = np.arange(n[2])[None, None, :]
    for (a, b, c) in [(1, 0, 0), (0, 2, 0), (1, 3, 5), (4, 6, 8)]:
        rho = np.cos(2 * np.pi * (a * j / n[0] + b * k / n[1] + c * l / n[2]))
        phi = O.fft_poisson(d, Narg, rho)
        m = [O.fft_order_int(n[i])[(a, b, c)[i]] for i in range(3)]
        k2 = sum((2 * np.pi * m[i] / (Narg[i] * d[i])) ** 2 for i in range(3))
        np.testing.assert_allclose(phi, -rho / k2, atol=1e-13)
    # k = 0 -> 0: a constant source gives zero potential
    np.testing.assert_allclose(O.fft_poisson(d, Narg, np.full(n, 3.0)), 0.0, atol=1e-13)


def test_poisson_linear_and_mean_free():
    rng = np.random.default_rng(3)
    a = rng.standard_normal((8, 8, 8))
    b = rng.standard_normal((8, 8, 8))
    d, N = (0.5, 0.5, 0.5), (7, 7, 7)
    pa, pb = O.fft_poisson(d, N, a), O.fft_poisson(d, N, b)
    np.testing.assert_allclose(O.fft_poisson(d, N, 2 * a - 3 * b), 2 * pa - 3 * pb, atol=1e-12)
    np.testing.assert_allclose(O.fft_poisson(d, N, a + 7.0), pa, atol=1e-12)
    assert abs(pa.mean()) < 1e-13


def test_poisson_reference_sanity_test_parallel():
    # test/test_parallel.jl:70-97: 32^3 Gaussian, Delta=1, N=(32,32,32)
    x = O.linrange(-np.pi, np.pi, 32)
    rho = np.exp(-(x[:, None, None] ** 2 + x[None, :, None] ** 2 + x[None, None, :] ** 2))
    phi = O.parallel_poisson((1.0, 1.0, 1.0), (32, 32, 32), rho)
    assert phi.shape == (32, 32, 32) and np.iscomplexobj(phi)
    assert np.all(np.isfinite(phi))
    assert 1e-6 < np.abs(phi).max() < 100.0


def test_poisson_laplacian_inverts_spectral_laplacian():
    # Apply the lattice Laplacian (multiply by -k^2) to the solution: recovers rho - mean.
    rng = np.random.default_rng(5)
    rho = rng.standard_normal((8, 6, 10))
    d, N = (0.4, 0.5, 0.3), (8, 6, 10)
    phi = O.fft_poisson(d, N, rho)
    k2 = sum(np.meshgrid(*[(2 * np.pi * O.fft_order_int(rho.shape[i]) / (N[i] * d[i])) ** 2
                           for i in range(3)], indexing="ij"))
    lap = np.fft.ifftn(-k2 * np.fft.fftn(phi))
    np.testing.assert_allclose(lap, rho - rho.mean(), atol=1e-12)


@pytest.mark.parametrize("shape,X", [((8, 12, 16), (3.0, 2.5, 2.0)), ((16, 16, 10), (10.0, 10.0, 4.0))])
def test_poisson_lattice_is_the_drift_operators_laplacian(shape, X):
    """Ties the oracle's Poisson lattice to reference text rather than to its own
    k formula.  The Laplacian is read off the reference's drift operator,
    written out here from src/solver/fft_operators.jl:9-13 (f on
    LinRange(-N/(4X), N/(4X) - 1/(2X), N), Lap = sum (2 pi i f)^2,
    linear_phase = fftshift(exp(i Lap dt/2)); dt small enough that the phase
    does not wrap, so Lap = 2 arg(linear_phase)/dt).  The oracle's potential of
    src/solver/gravity.jl:13-23 (4 pi * parallel_poisson(Delta, [N-1...], rho -
    mean)) must then satisfy Lap * F(Phi) = 4 pi F(rho - mean) at every k != 0
    on the reference grid (Delta = 2X/(N-1), even N).  A Poisson period of
    N*Delta instead of (N-1)*Delta = 2X fails by a factor ((N-1)/N)^2."""
    lap_axes = []
    for n, Xa in zip(shape, X):
        f = O.linrange(-n / 4.0 / Xa, n / 4.0 / Xa - 1.0 / 2.0 / Xa, n)  # LinRange (pinned above)
        lap_axes.append(np.roll(f, n // 2))  # Julia fftshift = circshift by n div 2
    fx, fy, fz = lap_axes
    Lap = ((2j * np.pi * fx[:, None, None]) ** 2 + (2j * np.pi * fy[None, :, None]) ** 2
           + (2j * np.pi * fz[None, None, :]) ** 2).real
    dt = 1e-3 / np.abs(Lap).max()
    linear_phase = np.exp(1j * Lap * dt / 2)
    Lap_read = 2.0 * np.angle(linear_phase) / dt
    np.testing.assert_allclose(Lap_read, Lap, rtol=1e-9, atol=1e-12 * np.abs(Lap).max())
    # the oracle's drift operator is the same array (its own restatement)
    np.testing.assert_allclose(O.setup_fft_operators(*X, *shape, dt), linear_phase, atol=1e-14)
    x, y, z, delta, _ = O.setup_grid(*X, *shape)
    rng = np.random.default_rng(11)
    psi = rng.standard_normal(shape) + 1j * rng.standard_normal(shape) + 1.5
    phi = O.compute_gravitational_potential(psi, delta)
    rho = np.abs(psi) ** 2
    lhs = Lap_read * np.fft.fftn(phi)
    rhs = 4.0 * np.pi * np.fft.fftn(rho - rho.mean())
    nz = np.ones(shape, dtype=bool)
    nz[0, 0, 0] = False  # FFT order: k = 0 first on every axis
    assert np.abs(Lap_read[0, 0, 0]) <= 1e-12 * np.abs(Lap).max()
    assert np.all(np.abs(Lap_read[nz]) > 1e-6 * np.abs(Lap).max())
    np.testing.assert_allclose(lhs[nz], rhs[nz], rtol=1e-9, atol=1e-9 * np.abs(rhs).max())
    assert abs(np.fft.fftn(phi)[0, 0, 0]) < 1e-9 * np.abs(rhs).max()


# ---------------------------------------------------------------- kick
def test_kick_preserves_modulus_and_sets_phase():
    rng = np.random.default_rng(7)
    psi = rng.standard_normal((6, 6, 6)) + 1j * rng.standard_normal((6, 6, 6))
    phi = rng.standard_normal((6, 6, 6))
    V = rng.standard_normal((6, 6, 6))
    hbar_over_m, dtk = 2.5, 0.01
    spec = O.apply_kick_step(psi, phi, V, dtk, hbar_over_m)
    kicked = np.fft.ifftn(spec)
    np.testing.assert_allclose(np.abs(kicked), np.abs(psi), rtol=1e-13)
    np.testing.assert_allclose(kicked / psi, np.exp(-1j * dtk * (phi + V) / hbar_over_m), rtol=1e-13)


# ---------------------------------------------------------------- absorber / timestep
def test_absorber_alpha_zero_is_one_and_1d():
    x = O.linrange(-4, 4, 16)
    b = O.setup_absorption_boundary(4, 4, 4, x, x, x, 0.0, 1e-3)
    assert b.shape == (16,) and np.all(b == 1.0)
    b = O.setup_absorption_boundary(4, 4, 4, x, x, x, 10.0, 1e-3)
    assert b.shape == (16,) and np.all(b <= 1.0) and b[8] > b[0]


def test_compute_timestep_float32_and_autoset():
    psi = np.ones((4, 4, 4))
    phi = np.full((4, 4, 4), 2.0)
    t, dt, Nt, lims = O.compute_timestep(0.1, phi, 0, psi, 1.0, 30, False, 0.9)
    assert Nt == 30 and dt == float(np.float32(1.0 / 30)) and len(t) == 30
    assert math.isclose(lims[0], 2 * 0.01 / math.pi) and math.isclose(lims[1], math.pi)
    assert lims[2] == math.inf
    t, dt, Nt, lims = O.compute_timestep(0.1, phi, 0, psi, 1.0, 30, True, 0.9)
    assert Nt == math.ceil(1.0 / (0.9 * lims[0]))


# ---------------------------------------------------------------- whole step invariants
def _seeded_psi(n, seed=11, X=4.0):
    rng = np.random.default_rng(seed)
    x = O.linrange(-X, X, n)
    r2 = x[:, None, None] ** 2 + x[None, :, None] ** 2 + x[None, None, :] ** 2
    amp = 30.0 * np.exp(-r2 / 2.0)
    ph = rng.uniform(-0.3, 0.3, (n, n, n))
    return np.asfortranarray(amp * np.exp(1j * ph))


def test_mass_conserved_over_100_steps():
    psi0 = _seeded_psi(16)
    _, _, _, _, dV = O.setup_grid(4, 4, 4, 16, 16, 16)
    m0 = O.mass(psi0, dV)
    psi = O.run_ssfm(psi0, (4, 4, 4), 100, 2e-3, four_pi_G=4 * np.pi)
    assert abs(O.mass(psi, dV) - m0) / m0 < 1e-13


def test_step_is_exactly_reversible():
    psi0 = _seeded_psi(12, seed=2)
    X, dt, dtk = (4.0, 4.0, 4.0), 3e-3, 1.5e-3
    n = 12
    _, _, _, delta, _ = O.setup_grid(*X, n, n, n)
    psi1 = O.run_ssfm(psi0, X, 1, dt, dtk)
    lp = O.setup_fft_operators(*X, n, n, n, dt)
    back = np.fft.ifftn(np.fft.fftn(psi1) / lp)
    phi = O.compute_gravitational_potential(back, delta)
    rec = np.exp(+1j * dtk * phi) * back
    np.testing.assert_allclose(rec, psi0, atol=1e-12 * np.abs(psi0).max())


def test_axis_permutation_and_reflection_symmetry():
    psi0 = _seeded_psi(10, seed=4)
    X = (4.0, 4.0, 4.0)
    out = O.run_ssfm(psi0, X, 3, 2e-3)
    perm = O.run_ssfm(np.transpose(psi0, (2, 0, 1)), X, 3, 2e-3)
    np.testing.assert_allclose(perm, np.transpose(out, (2, 0, 1)), atol=1e-12 * np.abs(out).max())
    refl = O.run_ssfm(psi0[::-1, :, ::-1], X, 3, 2e-3)
    np.testing.assert_allclose(refl, out[::-1, :, ::-1], atol=1e-12 * np.abs(out).max())


@pytest.mark.parametrize("split,order", [(O.SPLIT_STRANG, 2)])
def test_harmonic_oscillator_ground_state(split, order):
    # 3-D analogue of test/simulation_SPE1D.jl:8-10,202: V = r^2/2, hbar/m = 1,
    # no gravity; psi0 = exp(-r^2/2) is stationary with energy 3/2.
    n, X, T = 24, 7.0, 0.6
    x, _, _, _, dV = O.setup_grid(X, X, X, n, n, n, consistent=True)
    xxx, yyy, zzz, r = O.setup_coordinates(x, x, x)
    psi0 = np.exp(-r ** 2 / 2).astype(complex)
    V = 0.5 * r ** 2
    errs = []
    for nsteps in (20, 40):
        dt = T / nsteps
        psi = O.run_ssfm(psi0, (X, X, X), nsteps, dt, dt, V=V, self_gravity=False,
                         consistent_grid=True, split=split)
        exact = psi0 * np.exp(-1.5j * T)
        errs.append(np.linalg.norm(psi - exact) / np.linalg.norm(exact))
    assert errs[1] < 2e-3
    rate = math.log2(errs[0] / errs[1])
    assert rate > order - 0.3, (errs, rate)


def test_lie_split_harmonic_oscillator_phase():
    # Reference-style Lie split (dt_kick = dt, as test/simulation_SPE1D.jl:127):
    # the overlap phase tracks -3T/2 to first order in dt.
    n, X, T, nsteps = 24, 7.0, 0.5, 100
    x, _, _, _, _ = O.setup_grid(X, X, X, n, n, n, consistent=True)
    _, _, _, r = O.setup_coordinates(x, x, x)
    psi0 = np.exp(-r ** 2 / 2).astype(complex)
    psi = O.run_ssfm(psi0, (X, X, X), nsteps, T / nsteps, T / nsteps, V=0.5 * r ** 2,
                     self_gravity=False, consistent_grid=True)
    ov = np.vdot(psi0, psi) / np.vdot(psi0, psi0)
    assert abs(abs(ov) - 1) < 1e-3
    assert abs(np.angle(ov) + 1.5 * T) < 0.01


def test_nls_bright_soliton_kappa_term():
    # kappa |psi|^2 with kappa = -1 (the V of test/simulation_SPE1D.jl:219) on an
    # embedded 1-D line: psi = a sech(a x) exp(i a^2 t / 2) is exact for
    # i psi_t = -psi_xx/2 - |psi|^2 psi.
    n, X, T, a = 256, 20.0, 0.5, 1.3
    x, _, _, _, _ = O.setup_grid(X, 1.0, 1.0, n, 1, 1, consistent=True)
    psi0 = (a / np.cosh(a * x)).astype(complex)[:, None, None]
    Vf = lambda xxx, yyy, zzz, ps, t: -np.abs(ps) ** 2  # noqa: E731
    psi = O.run_ssfm(psi0, (X, 1.0, 1.0), 200, T / 200, V=Vf, self_gravity=False,
                     consistent_grid=True, split=O.SPLIT_STRANG)
    exact = psi0 * np.exp(0.5j * a * a * T)
    assert np.linalg.norm(psi - exact) / np.linalg.norm(exact) < 1e-3
