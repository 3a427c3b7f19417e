"""Pins of the diagnostics oracle (oracle/diag_oracle.py) against library
routines and closed forms, each chosen so that a plausible slip (h without
the first weight, the bias-corrected std, a sign or axis swap in the
momentum, the wrong mid plane, a missing 1/2) fails one of them."""
import numpy as np
import pytest

from oracle import diag_oracle as D
from oracle import ssfm_oracle as O


def test_equal_weight_quantile_is_numpy_linear_quantile():
    # StatsBase's non-frequency rule with unit weights: h = p (n - 1) + 1,
    # linear between order statistics -- numpy's default (type 7) quantile
    rng = np.random.default_rng(0)
    for n in (2, 3, 10, 257):
        v = rng.normal(size=n)
        ps = np.linspace(0.0, 1.0, 11)
        np.testing.assert_allclose(D.weighted_quantile(v, np.ones(n), ps), np.quantile(v, ps), rtol=0, atol=1e-14)
        # a common scale of the weights changes nothing
        np.testing.assert_allclose(D.weighted_quantile(v, 3.7 * np.ones(n), ps), np.quantile(v, ps), atol=1e-13)


def test_weighted_quantile_drops_zero_weights_and_hand_case():
    v = np.array([5.0, 1.0, 3.0, 2.0, 4.0])
    w = np.array([0.0, 1.0, 1.0, 0.0, 1.0])
    np.testing.assert_allclose(D.weighted_quantile(v, w, [0.25, 0.5, 1.0]), np.quantile([1.0, 3.0, 4.0], [0.25, 0.5, 1.0]))
    # weights (1, 2, 1) on (0, 1, 2): W = 4, w1 = 1, S = (1, 3, 4).  p = 0.5: h = 2.5,
    # first S > h is S_2 = 3 -> 0 + (2.5 - 1)/(3 - 1) (1 - 0) = 0.75
    assert D.weighted_quantile([0.0, 1.0, 2.0], [1.0, 2.0, 1.0], [0.5])[0] == pytest.approx(0.75, abs=1e-15)
    # p = 0.9: h = 3.7 -> 1 + (3.7 - 3)/(4 - 3) (2 - 1) = 1.7
    assert D.weighted_quantile([0.0, 1.0, 2.0], [1.0, 2.0, 1.0], [0.9])[0] == pytest.approx(1.7, abs=1e-15)


def test_weighted_std_is_the_uncorrected_estimate():
    rng = np.random.default_rng(1)
    v, w = rng.normal(size=50), rng.uniform(0.1, 2.0, size=50)
    assert D.weighted_std(v, w) == pytest.approx(np.sqrt(np.cov(v, aweights=w, bias=True)), rel=1e-13)
    # two points: |v1 - v2| sqrt(w1 w2) / (w1 + w2)
    assert D.weighted_std([1.0, 4.0], [1.0, 3.0]) == pytest.approx(3.0 * np.sqrt(3.0) / 4.0, rel=1e-15)


def _grid(shape, X=(3.0, 2.0, 2.5)):
    x, y, z, d, dV = O.setup_grid(*X, *shape)
    return x, y, z, d, dV


def test_plane_wave_momentum_and_kinetic_closed_form():
    # psi = A exp(i k y) with k on the lattice: the central difference gives
    # rho v_y = A^2 sin(k dy)/dy exactly, the other components 0, QE = 0
    shape = (8, 10, 6)
    x, y, z, d, dV = _grid(shape)
    A, m = 0.7, 2
    k = 2 * np.pi * m / (shape[1] * d[1])
    j = np.arange(shape[1])
    psi = A * np.exp(1j * k * d[1] * j)[None, :, None] * np.ones(shape)
    out = D.diagnostics(psi, None, x, y, z)
    rv = A * A * np.sin(k * d[1]) / d[1]
    assert out["momentum"][1] == pytest.approx(rv * psi.size * dV, rel=1e-12)
    assert abs(out["momentum"][0]) < 1e-13 and abs(out["momentum"][2]) < 1e-13
    nint = np.prod([n - 2 for n in shape])
    assert out["kinetic"] == pytest.approx(0.5 * nint * rv * rv / (A * A) * dV, rel=1e-12)
    assert abs(out["quantum"]) < 1e-24
    assert out["mass"] == pytest.approx(A * A * psi.size * dV, rel=1e-13)


def test_quantum_energy_converges_to_the_continuum_integral():
    # real Gaussian sqrt(rho) = exp(-r^2 / (2 s^2)): 1/2 int |grad sqrt(rho)|^2
    # = 1/2 * (3 / (2 s^2)) * (pi s^2)^(3/2); central differences are O(dx^2)
    s = 0.5
    errs = []
    for n in (24, 48):
        x, y, z, d, dV = _grid((n, n, n), X=(3.0, 3.0, 3.0))
        X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
        psi = np.exp(-(X ** 2 + Y ** 2 + Z ** 2) / (2 * s * s)).astype(complex)
        out = D.diagnostics(psi, None, x, y, z)
        exact = 0.5 * 3.0 / (2 * s * s) * (np.pi * s * s) ** 1.5
        errs.append(abs(out["quantum"] / exact - 1))
        assert out["kinetic"] == 0.0 and out["momentum"] == (0.0, 0.0, 0.0)
    assert errs[1] < 0.04 and errs[1] < errs[0] / 3.5, errs  # second order: ratio -> 4


def test_potential_term_and_virial_combination():
    shape = (6, 7, 8)
    x, y, z, d, dV = _grid(shape)
    rng = np.random.default_rng(2)
    psi = rng.normal(size=shape) + 1j * rng.normal(size=shape) + 2.0
    c = -1.25
    out = D.diagnostics(psi, np.full(shape, c), x, y, z)
    rho = np.abs(psi) ** 2
    assert out["virial_potential"] == pytest.approx(-0.5 * c * rho[1:-1, 1:-1, 1:-1].sum() * dV, rel=1e-13)
    assert out["virial"] == pytest.approx(2 * out["kinetic"] + 2 * out["quantum"] - out["virial_potential"], rel=1e-15)


def test_mid_plane_radius_uses_plane_nz_div_2():
    # weight only on two cells of the 0-based plane nz//2 - 1 (Julia div(end, 2))
    shape = (8, 8, 8)
    x, y, z, d, dV = _grid(shape, X=(2.0, 2.0, 2.0))
    psi = np.full(shape, 1e-3 + 0j)
    k = shape[2] // 2 - 1
    psi[1, 2, k], psi[6, 3, k] = 1.0, np.sqrt(3.0)
    out = D.diagnostics(psi, None, x, y, z)
    r1 = np.sqrt(x[1] ** 2 + y[2] ** 2 + z[k] ** 2)
    r2 = np.sqrt(x[6] ** 2 + y[3] ** 2 + z[k] ** 2)
    w = np.full((8, 8), 1e-6)
    w[1, 2], w[6, 3] = 1.0, 3.0
    rr = np.sqrt(x[:, None] ** 2 + y[None, :] ** 2 + z[k] ** 2)
    mu = (w * rr).sum() / w.sum()
    assert out["r_std_mid_plane"] == pytest.approx(np.sqrt((w * (rr - mu) ** 2).sum() / w.sum()), rel=1e-12)
    # dominated by the two bright cells: |r1 - r2| sqrt(3)/4 to O(1e-6 * 64)
    assert out["r_std_mid_plane"] == pytest.approx(abs(r1 - r2) * np.sqrt(3.0) / 4.0, rel=1e-3)


def test_peak_is_the_first_maximum_in_column_order():
    shape = (5, 4, 3)
    x, y, z, d, dV = _grid(shape)
    psi = np.full(shape, 0.1 + 0j)
    psi[3, 1, 2] = 2.0
    psi[1, 2, 2] = 2.0  # same rho at x-fastest index 51, after 48
    out = D.diagnostics(psi, None, x, y, z)
    i1 = 3 + 5 * 1 + 20 * 2
    i2 = 1 + 5 * 2 + 20 * 2
    assert out["peak_index"] == min(i1, i2)
    assert out["rho_max"] == 4.0


def test_radii_of_a_point_and_shell_distribution():
    # rho = 1 at the peak cell (weight 1 at distance 0) and 0.25 on its six
    # face neighbours at distance dx (equal spacing): sorted (v, w) = (0, 1),
    # 6 x (dx, 0.25).  W = 2.5, h = p * 1.5 + 1, S = 1, 1.25, ..., 2.5
    shape = (9, 9, 9)
    x, y, z, d, dV = _grid(shape, X=(2.0, 2.0, 2.0))
    psi = np.zeros(shape, dtype=complex)
    c = (4, 4, 4)
    psi[c] = 1.0
    for a in range(3):
        for s in (-1, 1):
            i = list(c)
            i[a] += s
            psi[tuple(i)] = 0.5
    out = D.diagnostics(psi, None, x, y, z)
    ps = np.arange(1, 10) / 10
    h = ps * 1.5 + 1.0
    # first S_k > h among S = 1 + 0.25 k: the value just above 0 is the first
    # shell element (S = 1.25): interpolate from 0 when h < 1.25, else dx
    exp = np.where(h < 1.25, (h - 1.0) / 0.25 * d[0], d[0])
    np.testing.assert_allclose(out["radii"], exp, rtol=1e-12)
