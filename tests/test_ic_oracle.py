"""Pins of the velocity -> phase oracle (oracle/ic_oracle.py) by closed forms:
the trapezoid rule is exact for velocities linear along the line, so a
gradient field v = grad(phi) with phi quadratic gives theta = phi - phi(first
cell of every axis) exactly; constant velocities give plane-wave phases; the
bulk removal leaves zero mass-weighted mean velocity."""
import numpy as np

from oracle import ic_oracle as I


def test_cumul_integrate_is_the_trapezoid_rule():
    x = np.array([0.0, 0.5, 1.5, 3.0])
    y = np.array([1.0, 3.0, 2.0, 4.0])
    # 1/2 (0.5*4), + 1/2 (1*5), + 1/2 (1.5*6)
    np.testing.assert_allclose(I.cumul_integrate(x, y), [0.0, 1.0, 3.5, 8.0], rtol=1e-15)
    xs = np.linspace(0, 2, 9)
    np.testing.assert_allclose(I.cumul_integrate(xs, 3 * xs + 1), 1.5 * xs ** 2 + xs, rtol=1e-14, atol=1e-15)


def test_axis_form_equals_the_line_loop():
    rng = np.random.default_rng(0)
    v = rng.normal(size=(5, 4, 3))
    for axis, d in ((0, 0.3), (1, 0.7), (2, 1.1)):
        got = I.cumul_integrate_axis(v, d, axis)
        ref = np.apply_along_axis(lambda y: I.cumul_integrate(np.arange(len(y)) * d, y), axis, v)
        np.testing.assert_allclose(got, ref, rtol=1e-14, atol=1e-15)


def test_gradient_field_gives_its_potential():
    n, d = (7, 6, 5), (0.4, 0.5, 0.3)
    x, y, z = (np.arange(k) * h for k, h in zip(n, d))
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    # phi = a x^2 + b y + c z^2 + e x (separable): grad along each axis is linear on its line
    a, b, c, e = 0.7, -1.3, 0.4, 2.0
    phi = a * X ** 2 + b * Y + c * Z ** 2 + e * X
    theta = I.solve_vector_equation(2 * a * X + e, b + 0 * Y, 2 * c * Z, *d)
    np.testing.assert_allclose(theta, phi, rtol=1e-13, atol=1e-13)


def test_bulk_removal_and_assembly():
    rng = np.random.default_rng(1)
    shape, d = (6, 5, 4), (0.2, 0.3, 0.4)
    rho = rng.uniform(0.1, 2.0, size=shape)
    v = [rng.normal(size=shape) + 3.0 for _ in range(3)]
    w = I.remove_bulk_velocity(rho, *v)
    for c in w:
        assert abs((c * rho).sum()) < 1e-12 * np.abs(c * rho).sum()
    # constant velocity (no bulk removal) -> plane wave exp(i (v . x) / (hbar/m))
    vc = (0.8, -0.5, 0.3)
    psi = I.ic_from_velocity(rho, *(np.full(shape, c) for c in vc), d, hbar_over_m=2.0, remove_bulk=False)
    X, Y, Z = np.meshgrid(*(np.arange(k) * h for k, h in zip(shape, d)), indexing="ij")
    np.testing.assert_allclose(psi, np.sqrt(rho) * np.exp(1j * (vc[0] * X + vc[1] * Y + vc[2] * Z) / 2.0),
                               rtol=1e-13, atol=1e-13)
    assert np.allclose(np.abs(I.ic_from_velocity(rho, *v, d)) ** 2, rho, rtol=1e-14)
