"""Pins of the tidal-potential oracle (oracle/tidal_oracle.py) by what the
mathematics fixes: multilinear interpolation is exact on affine tables and
returns the table at its nodes, and clamps outside the grid; the gradient
cancellation leaves a zero central difference at the centre, keeps the
rho-weighted mean of the field, and turns a field linear in the cell index
(a uniform tidal force) into a constant; the trajectory lookup picks the last
sample at or before the kick."""
import numpy as np

from oracle import tidal_oracle as T


def _axes():
    return [np.linspace(-4.0, 4.0, 9), np.linspace(-3.0, 5.0, 5), np.linspace(-2.0, 2.0, 7)]


def test_interpolation_exact_on_affine_tables_and_nodes():
    axes = _axes()
    A = np.meshgrid(*axes, indexing="ij")
    table = 1.5 + 0.7 * A[0] - 2.0 * A[1] + 0.25 * A[2]
    rng = np.random.default_rng(1)
    pts = np.stack([rng.uniform(ax[0], ax[-1], 200) for ax in axes], axis=-1)
    np.testing.assert_allclose(T.interpolate(axes, table, pts),
                               1.5 + 0.7 * pts[:, 0] - 2.0 * pts[:, 1] + 0.25 * pts[:, 2], rtol=1e-13, atol=1e-13)
    # at the nodes: the table itself, for any table
    t2 = rng.normal(size=table.shape)
    nodes = np.stack([a.ravel() for a in A], axis=-1)
    np.testing.assert_allclose(T.interpolate(axes, t2, nodes), t2.ravel(), rtol=0, atol=1e-14)
    # outside the grid: the value at the clamped point
    out = np.array([[9.0, -7.0, 0.3], [-5.0, 0.5, 3.0]])
    clamped = np.array([[4.0, -3.0, 0.3], [-4.0, 0.5, 2.0]])
    np.testing.assert_array_equal(T.interpolate(axes, t2, out), T.interpolate(axes, t2, clamped))


def test_gradient_cancellation_properties():
    rng = np.random.default_rng(2)
    field = rng.normal(size=(10, 8, 6))
    rho = rng.uniform(0.1, 1.0, size=field.shape)
    c = (4, 3, 2)
    out = T.cancel_field_gradient(field, c, rho)
    for a in range(3):
        up, dn = list(c), list(c)
        up[a] += 1
        dn[a] -= 1
        assert abs(out[tuple(up)] - out[tuple(dn)]) < 1e-13
    assert np.isclose(np.sum(rho * out), np.sum(rho * field), rtol=1e-13)


def test_uniform_tidal_force_is_removed():
    """pot = a + b . i (a uniform force in index units): centred and with its
    gradient cancelled it is one constant on the whole grid."""
    n = (12, 9, 7)
    i, j, k = np.meshgrid(*[np.arange(m) for m in n], indexing="ij")
    axes = [np.arange(m, dtype=np.float64) for m in n]
    table = 3.0 + 0.4 * i - 1.1 * j + 0.25 * k
    rho = np.exp(-((i - 5) ** 2 + (j - 4) ** 2 + (k - 3) ** 2) / 6.0)
    pot = T.tidal_potential(axes, table, axes[0], axes[1], axes[2], (0.0, 0.0, 0.0), rho)
    assert np.ptp(pot) < 1e-12
    # the constant keeps the centred field's rho-weighted mean
    centred = table - table[5, 4, 3]
    assert np.isclose(pot.mean(), np.sum(rho * centred) / np.sum(rho), rtol=1e-12)


def test_trajectory_offset_is_the_last_sample_before_the_kick():
    times = np.array([0.0, 0.5, 1.0, 1.5])
    xyz = np.arange(12.0).reshape(4, 3)
    assert np.array_equal(T.trajectory_offset(times, xyz, 0.7), xyz[1])
    assert np.array_equal(T.trajectory_offset(times, xyz, 1.0), xyz[2])
    assert np.array_equal(T.trajectory_offset(times, xyz, -1.0), xyz[0])
    assert np.array_equal(T.trajectory_offset(times, xyz, 9.0), xyz[3])
