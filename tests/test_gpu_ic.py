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
