"""Restart and the device time-averaged density through the C ABI.

A run saved after k steps, restored into a new plan and continued equals the
uninterrupted run bit for bit (FP64 and FP32).  The averaged density equals
the mean of |psi|^2 of the accumulated states (oracle: numpy on get_psi) to
1e-13 relative."""
import numpy as np
import pytest

from oracle import ssfm_oracle as O

pytestmark = pytest.mark.gpu

X = (3.0, 2.5, 2.0)


def _psi(shape, seed=4):
    x, y, z, _, _ = O.setup_grid(*X, *shape)
    Xg, Yg, Zg = np.meshgrid(x, y, z, indexing="ij")
    rng = np.random.default_rng(seed)
    return (np.exp(-(Xg ** 2 + Yg ** 2 + Zg ** 2) + 1j * (0.7 * Xg - 0.4 * Zg))
            + 0.01 * (rng.normal(size=shape) + 1j * rng.normal(size=shape)))


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_restart_continues_bit_for_bit(wdm, nativelib, tmp_path, prec):
    shape, dt = (32, 24, 16), 2e-3
    psi0 = _psi(shape)

    def install(p):
        p.set_hernquist(8.0, 0.6, (0.2, -0.1, 0.0))

    with wdm.Plan(shape, X, prec, four_pi_G=3.0, lib=nativelib) as a:
        install(a)
        a.set_psi(psi0)
        a.step(7, dt)
        ref = a.get_psi()
    with wdm.Plan(shape, X, prec, four_pi_G=3.0, lib=nativelib) as b:
        install(b)
        b.set_psi(psi0)
        b.step(3, dt)
        wdm.snapshot.save(tmp_path / "s.npz", b, t=3 * dt, step=3, dt=dt)
    snap = wdm.snapshot.load(tmp_path / "s.npz")
    with wdm.snapshot.plan_from(snap, lib=nativelib) as c:
        install(c)
        c.step(7 - snap["step"], snap["dt"], t0=snap["t"])
        out = c.get_psi()
    assert np.array_equal(out, ref)


def test_density_average_matches_host_mean(wdm, nativelib):
    shape, dt = (16, 12, 8), 2e-3
    acc = np.zeros(shape)
    with wdm.Plan(shape, X, "f64", lib=nativelib) as p:
        assert p.density_average() == (None, 0)
        p.set_psi(_psi(shape, 9))
        for _ in range(5):
            p.step(1, dt)
            p.density_accumulate()
            acc += np.abs(p.get_psi()) ** 2
        avg, n = p.density_average()
        p.density_reset()
        assert p.density_average() == (None, 0)
    assert n == 5
    ref = acc / 5
    assert np.abs(avg - ref).max() <= 1e-13 * ref.max()


def test_restart_with_mass_projection_and_average(wdm, nativelib, tmp_path):
    """FP32 with the mass projection on and a running density average: the
    snapshot carries the projection's reference mass and the FP64 density
    sum, so the restored run's second call equals the saved run's second call
    bit for bit, and the average continues over both calls."""
    shape, dt = (32, 24, 16), 2e-3
    psi0 = _psi(shape)

    def first_call(p):
        p.set_hernquist(8.0, 0.6, (0.2, -0.1, 0.0))
        p.set_psi(psi0)
        p.set_mass_ledger(True)
        p.step(3, dt)
        p.density_accumulate()

    with wdm.Plan(shape, X, "f32", four_pi_G=3.0, lib=nativelib) as a:
        first_call(a)
        m_ref = a.mass_ledger_state()
        a.step(4, dt, t0=3 * dt)
        a.density_accumulate()
        ref = a.get_psi()
        ref_avg, ref_n = a.density_average()
    with wdm.Plan(shape, X, "f32", four_pi_G=3.0, lib=nativelib) as b:
        first_call(b)
        wdm.snapshot.save(tmp_path / "s.npz", b, t=3 * dt, step=3, dt=dt, average=True)
    snap = wdm.snapshot.load(tmp_path / "s.npz")
    assert snap["mass_ledger"] is True and snap["mass_reference"] == m_ref[1] and m_ref[0]
    assert snap["average_count"] == 1 and snap["psi2_sum"].dtype == np.float64
    with wdm.snapshot.plan_from(snap, lib=nativelib) as c:
        assert c.mass_ledger_state() == (True, snap["mass_reference"])
        c.set_hernquist(8.0, 0.6, (0.2, -0.1, 0.0))
        c.step(4, snap["dt"], t0=snap["t"])
        c.density_accumulate()
        out = c.get_psi()
        avg, n = c.density_average()
    assert np.array_equal(out, ref)
    assert n == ref_n == 2 and np.array_equal(avg, ref_avg)


def test_mass_reference_errors(wdm, nativelib):
    with wdm.Plan((16, 12, 8), X, "f32", lib=nativelib) as p:
        p.set_psi(_psi((16, 12, 8)))
        assert p.mass_ledger_state() == (False, None)
        with pytest.raises(wdm.WaveDMError) as e:
            p.set_mass_reference(1.0)
        assert e.value.code == wdm._capi.WDM_ERR_STATE
        p.set_mass_ledger(True)
        assert p.mass_ledger_state() == (True, None)   # seeded by the next step
        with pytest.raises(wdm.WaveDMError):
            p.set_mass_reference(-1.0)
        p.step(1, 1e-3)
        on, m = p.mass_ledger_state()
        assert on and m > 0
        p.set_mass_reference(2.0 * m)
        assert p.mass_ledger_state() == (True, 2.0 * m)
