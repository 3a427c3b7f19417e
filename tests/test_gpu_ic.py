"""Device initial condition from a density and a velocity field
(wdm_ic_velocity) against the oracle (oracle/ic_oracle.py, pinned by closed
forms).  Bars: FP64 plan <= 1e-12 relative max error on psi, FP32 <= 2e-7."""
import numpy as np
import pytest

from oracle import ic_oracle as I
from oracle import ssfm_oracle as O

pytestmark = pytest.mark.gpu


def _fields(shape, X, seed):
    x, y, z, d, _ = O.setup_grid(*X, *shape)
    Xg, Yg, Zg = np.meshgrid(x, y, z, indexing="ij")
    rng = np.random.default_rng(seed)
    rho = np.exp(-(Xg ** 2 + Yg ** 2 + Zg ** 2) / 2.0) * 5.0 + 1e-3
    v = [rng.normal(size=shape) * 0.3 + s for s in (1.0, -0.5, 0.25)]
    # a rotating component: v = omega x r
    v[0] += -0.8 * Yg
    v[1] += 0.8 * Xg
    return rho, v, d


@pytest.mark.parametrize("shape,remove_bulk,hm", [((32, 24, 16), True, 1.0), ((16, 12, 8), False, 0.35),
                                                   ((64, 8, 12), True, 2.5)],
                         ids=["32x24x16-bulk", "16x12x8", "64x8x12-bulk"])
def test_ic_velocity_matches_oracle(wdm, nativelib, shape, remove_bulk, hm):
    X = (3.0, 2.5, 2.0)
    rho, v, d = _fields(shape, X, sum(shape))
    ref = I.ic_from_velocity(rho, *v, d, hbar_over_m=hm, remove_bulk=remove_bulk)
    for prec, tol in (("f64", 1e-12), ("f32", 2e-7)):
        with wdm.Plan(shape, X, prec, hbar_over_m=hm, lib=nativelib) as p:
            p.set_ic_velocity(rho, *v, remove_bulk=remove_bulk)
            out = p.get_psi()
        err = np.abs(out - ref).max() / np.abs(ref).max()
        assert err < tol, (prec, err)


def test_ic_velocity_rejects_zero_density(wdm, nativelib):
    from wavedm_amd._capi import WaveDMError, WDM_ERR_ARG
    shape = (16, 12, 8)
    z = np.zeros(shape)
    with wdm.Plan(shape, 2.0, "f64", lib=nativelib) as p:
        with pytest.raises(WaveDMError) as e:
            p.set_ic_velocity(z, z, z, z, remove_bulk=True)
        assert e.value.code == WDM_ERR_ARG
        p.set_ic_velocity(z, z, z, z, remove_bulk=False)  # no mean to take: psi = 0
        assert np.all(p.get_psi() == 0)


@pytest.mark.parametrize("shape,prec", [((32, 24, 16), "f64"), ((48, 32, 24), "f32"), ((24, 16, 12), "f64")])
def test_ic_acceleration_matches_oracle(wdm, nativelib, shape, prec):
    """wdm_ic_acceleration (compute_acceleration_field + a_CDM,
    src/ic/generation.jl:246-291) against the oracle: Phi, a and a_CDM.
    Bars: FP64 <= 1e-12 relative to each field's maximum, FP32 <= 2e-6."""
    X = (3.0, 2.5, 2.0)
    rho, _, d = _fields(shape, X, 7)
    ref = dict(zip(("phi", "ax", "ay", "az"), I.compute_acceleration_field(rho, d, 4 * np.pi * 0.6)))
    ref["a_cdm"] = I.a_cdm(ref["ax"], ref["ay"], ref["az"])
    with wdm.Plan(shape, X, prec, four_pi_G=4 * np.pi * 0.6, lib=nativelib) as p:
        p.set_psi(np.sqrt(rho) + 0j)
        got = wdm.ic.compute_acceleration_field(p, rho)
        # psi is kept, and the next kick's Phi is that of psi again
        assert np.allclose(p.get_psi(), np.sqrt(rho), rtol=0, atol=1e-6 * np.sqrt(rho).max())
    tol = 1e-12 if prec == "f64" else 2e-6
    for k, v in ref.items():
        err = np.abs(got[k] - v).max() / np.abs(v).max()
        print(f"{prec} {shape} {k}: {err:.3e}", flush=True)
        assert err <= tol, (k, err)
    assert got["a_cdm"][shape[0] // 2 - 1, shape[1] // 2 - 1, shape[2] // 2 - 1] == 0.0
    assert not got["a_cdm"][0].any() and not got["a_cdm"][:, :, -1].any()


def test_velocity_ic_chain_matches_oracle(wdm, nativelib):
    """The whole chain with a caller-supplied velocity formula (here a
    rotation v = omega x r scaled by a_CDM): device accelerations, centre
    zeroing (bulk_center_size 4), bulk removal, line integrals, psi; against
    the oracle chain.  FP64 <= 1e-12."""
    shape, X = (32, 24, 16), (3.0, 2.5, 2.0)
    rho, _, d = _fields(shape, X, 9)

    def velocity(Xg, Yg, Zg, a):
        return -0.3 * Yg * a, 0.3 * Xg * a, 0.05 * a

    x, y, z, _, _ = O.setup_grid(*X, *shape)
    Xg, Yg, Zg = np.meshgrid(x, y, z, indexing="ij")
    _, ax, ay, az = I.compute_acceleration_field(rho, d, 4 * np.pi)
    v = I.zero_center_velocity(*velocity(Xg, Yg, Zg, I.a_cdm(ax, ay, az)), 4)
    ref = I.ic_from_velocity(rho, *v, d)
    with wdm.Plan(shape, X, "f64", lib=nativelib) as p:
        wdm.ic.velocity_ic(p, rho, velocity, bulk_center_size=4)
        got = p.get_psi()
        with pytest.raises(NotImplementedError):
            wdm.ic.velocity_ic(p, rho, velocity, bulk_perturb=True)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"velocity IC chain: {err:.3e}", flush=True)
    assert err <= 1e-12
