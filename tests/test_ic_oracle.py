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


# ---------------------------------------------------------------- acceleration -> velocity
def test_acceleration_field_of_a_single_mode():
    """compute_acceleration_field's arithmetic (src/ic/generation.jl:264-272) on
    rho = cos(k . x) + 2, a mode periodic on the N-cell array (Poisson period
    N Delta): Phi = -4 pi cos(k.x)/k^2, and the central difference of a mode
    along axis a is sin(k_a Delta)/Delta times its derivative factor, so
    a_a = -(4 pi / k^2) sin(k_a Delta)/Delta sin(k . x) exactly.  The oracle's
    entry point is that arithmetic on the reference's [N-1] lattice."""
    from oracle import diag_oracle as D
    from oracle import ssfm_oracle as O
    shape, X = (16, 12, 10), (3.0, 2.5, 2.0)
    _, _, _, d, _ = O.setup_grid(*X, *shape)
    m = (2, 3, 1)
    idx = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    phase = sum(2 * np.pi * m[a] * idx[a] / shape[a] for a in range(3))
    rho = np.cos(phase) + 2.0
    kN = [2 * np.pi * m[a] / (shape[a] * d[a]) for a in range(3)]
    k2N = sum(v * v for v in kN)
    phi = 4 * np.pi * O.parallel_poisson(d, shape, rho - rho.mean()).real
    np.testing.assert_allclose(phi, -4 * np.pi * np.cos(phase) / k2N, atol=1e-12 * 4 * np.pi / k2N)
    for a in range(3):
        ga = -D.central_diff(phi, a, d[a])
        expect = -(4 * np.pi / k2N) * np.sin(kN[a] * d[a]) / d[a] * np.sin(phase)
        np.testing.assert_allclose(ga, expect, atol=1e-11 * 4 * np.pi / k2N / min(d))
    phi2, ax, ay, az = I.compute_acceleration_field(rho, d)
    ref = 4 * np.pi * O.parallel_poisson(d, tuple(n - 1 for n in shape), rho - rho.mean()).real
    np.testing.assert_array_equal(phi2, ref)
    for a, g in enumerate((ax, ay, az)):
        np.testing.assert_array_equal(g, -D.central_diff(ref, a, d[a]))


def test_a_cdm_interior_magnitude_and_zero_centre():
    """generate_velocity_field (src/ic/generation.jl:289-291): |a| on
    [2:end-1]^3, 0 on every face, and the 1-based cell (div(N,2), ...)
    multiplied by 0, written out by hand on a 6 x 5 x 4 grid."""
    rng = np.random.default_rng(3)
    ax, ay, az = (rng.normal(size=(6, 5, 4)) for _ in range(3))
    out = I.a_cdm(ax, ay, az)
    for i in range(6):
        for j in range(5):
            for k in range(4):
                inner = 0 < i < 5 and 0 < j < 4 and 0 < k < 3
                centre = (i, j, k) == (2, 1, 1)  # div(6,2)=3, div(5,2)=2, div(4,2)=2, minus 1
                want = 0.0 if (not inner or centre) else np.sqrt(ax[i, j, k] ** 2 + ay[i, j, k] ** 2 + az[i, j, k] ** 2)
                assert out[i, j, k] == want, (i, j, k)


def test_center_velocity_zeroing_ranges():
    """adjust_velocity_field (src/ic/generation.jl:316-326): the 1-based range
    div(N,2)-div(s,2)+1 : div(N,2)-div(s,2)+s per axis, here N = (8, 6, 5),
    s = 3: 1-based 4:6, 3:5, 2:4."""
    v = np.arange(8 * 6 * 5, dtype=np.float64).reshape(8, 6, 5) + 1.0
    vx, vy, vz = I.zero_center_velocity(v, 2 * v, 3 * v, 3)
    mask = np.zeros(v.shape, bool)
    mask[3:6, 2:5, 1:4] = True
    for got, f in ((vx, 1), (vy, 2), (vz, 3)):
        assert np.all(got[mask] == 0) and np.array_equal(got[~mask], f * v[~mask])
    same = I.zero_center_velocity(v, v, v, 0)
    assert all(np.array_equal(a, v) for a in same)
