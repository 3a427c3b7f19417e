"""Batched ensemble (wdm_ensemble_*) and the autotuner's pass-table variants
(wdm_autotune / wdm_select_variants) through the C ABI.

Bars: an ensemble member equals the same plan stepped alone, bit for bit
(graph and eager); every variant of every axis length matches the FP64
oracle to 1e-12 (FP64 plan) / 2e-5 (FP32 plan) after two steps."""
import numpy as np
import pytest

from oracle import ssfm_oracle as O

pytestmark = pytest.mark.gpu

X3 = (4.0, 3.5, 3.0)


def blob(shape, seed):
    rng = np.random.default_rng(seed)
    x, y, z, _, _ = O.setup_grid(*X3, *shape)
    r2 = x[:, None, None] ** 2 + y[None, :, None] ** 2 + z[None, None, :] ** 2
    return np.asfortranarray(30.0 * np.exp(-r2 / 2) * np.exp(1j * rng.uniform(-0.3, 0.3, shape)))


MEMBERS = [((32, 16, 8), "f32", None), ((16, 16, 16), "f64", "hernquist"), ((64, 8, 8), "f32", "border"),
           ((8, 24, 12), "f32", None)]


def _make(wdm, lib, shape, prec, extra, seed):
    p = wdm.Plan(shape, X3, prec, lib=lib)
    p.set_psi(blob(shape, seed))
    if extra == "hernquist":
        p.set_hernquist(50.0, 1.0, (0.5, 0.0, 0.0))
    elif extra == "border":
        p.set_border(np.linspace(0.95, 1.0, shape[0]))
    return p


@pytest.mark.parametrize("graph", [True, False], ids=["graph", "eager"])
def test_ensemble_equals_members_stepped_alone(wdm, nativelib, graph):
    dts = [0.01, 0.02, 0.015, 0.01]
    ens_plans = [_make(wdm, nativelib, s, pr, ex, i) for i, (s, pr, ex) in enumerate(MEMBERS)]
    ref_plans = [_make(wdm, nativelib, s, pr, ex, i) for i, (s, pr, ex) in enumerate(MEMBERS)]
    try:
        with wdm.Ensemble(ens_plans) as e:
            e.step(3, dts, graph=graph)
            e.step(2, dts, graph=graph)   # replays the captured step
        for p, dt in zip(ref_plans, dts):
            p.step(5, dt)
        for a, b in zip(ens_plans, ref_plans):
            dt = np.complex64 if a.precision == "f32" else np.complex128
            assert np.array_equal(a.get_psi(dt), b.get_psi(dt))
    finally:
        for p in ens_plans + ref_plans:
            p.close()


# members that share a shape, precision and potential kind step in batches
# (one launch per pass for all of them); the rest alone
BATCHED = [((32, 16, 8), "f32", None, 0.010), ((32, 16, 8), "f32", None, 0.020), ((32, 16, 8), "f32", "hernquist", 0.01),
           ((16, 12, 8), "f64", "hernquist", 0.012), ((16, 12, 8), "f64", "hernquist2", 0.008),
           ((24, 8, 16), "f32", "border", 0.015), ((24, 8, 16), "f32", "border", 0.01),
           ((16, 16, 16), "f32", "nograv", 0.01), ((16, 16, 16), "f32", "nograv", 0.02),
           ((48, 1, 16), "f64", None, 0.01), ((48, 1, 16), "f64", None, 0.02), ((32, 16, 8), "f32", "ledger", 0.01)]


def _make_b(wdm, lib, shape, prec, extra, seed):
    if extra == "nograv":
        p = wdm.Plan(shape, X3, prec, self_gravity=False, lib=lib)
        p.set_psi(blob(shape, seed))
        return p
    p = _make(wdm, lib, shape, prec, extra if extra != "hernquist2" else None, seed)
    if extra == "hernquist2":
        p.set_hernquist(30.0, 0.7, (-0.4, 0.2, 0.1))
    if extra == "ledger":
        p.set_mass_ledger(True)
    return p


@pytest.mark.parametrize("graph", [True, False], ids=["graph", "eager"])
def test_batched_members_equal_members_stepped_alone(wdm, nativelib, graph):
    ens_plans = [_make_b(wdm, nativelib, s, pr, ex, i) for i, (s, pr, ex, _) in enumerate(BATCHED)]
    ref_plans = [_make_b(wdm, nativelib, s, pr, ex, i) for i, (s, pr, ex, _) in enumerate(BATCHED)]
    dts = [m[3] for m in BATCHED]
    try:
        with wdm.Ensemble(ens_plans) as e:
            e.step(3, dts, graph=graph)
            e.step(2, dts, graph=graph)
            # five batches: the two plain (32,16,8), the Hernquist (16,12,8) pair,
            # the absorber pair, the pair without gravity, the (48,1,16) pair
            assert e.layout() == (5, 2)
        for p, dt in zip(ref_plans, dts):
            for _ in range(5):
                p.step(1, dt)
        for a, b, m in zip(ens_plans, ref_plans, BATCHED):
            dt = np.complex64 if a.precision == "f32" else np.complex128
            assert np.array_equal(a.get_psi(dt), b.get_psi(dt)), m
    finally:
        for p in ens_plans + ref_plans:
            p.close()


def test_graph_recaptures_on_changes(wdm, nativelib):
    shape = (32, 16, 8)
    a, b = _make(wdm, nativelib, shape, "f32", None, 1), _make(wdm, nativelib, shape, "f32", None, 1)
    try:
        with wdm.Ensemble([a]) as e:
            e.step(2, 0.01)
            b.step(2, 0.01)
            a.set_hernquist(80.0, 1.5, (0.0, 0.3, 0.0))     # configuration change after capture
            b.set_hernquist(80.0, 1.5, (0.0, 0.3, 0.0))
            e.step(2, 0.01)
            b.step(2, 0.01)
            e.step(2, 0.02)                                # dt change
            b.step(2, 0.02)
        assert np.array_equal(a.get_psi(np.complex64), b.get_psi(np.complex64))
    finally:
        a.close()
        b.close()


def test_graph_refuses_time_dependent_potential(wdm, nativelib):
    p = _make(wdm, nativelib, (16, 16, 16), "f32", None, 2)
    try:
        p.set_quadrupole(np.diag([1.0, 2.0, -3.0]), (0, 0, 1), 1.0)
        with wdm.Ensemble([p]) as e:
            with pytest.raises(wdm.WaveDMError) as err:
                e.step(1, 0.01, graph=True)
            assert err.value.code == wdm._capi.WDM_ERR_UNSUPPORTED
            e.step(1, 0.01, graph=False)
    finally:
        p.close()


VARIANT_SHAPES = [(64, 8, 8), (8, 128, 8), (8, 8, 192), (256, 8, 8), (384, 8, 8), (32, 16, 8),
                  (8, 384, 8), (8, 8, 384), (8, 256, 8), (8, 8, 256), (8, 160, 8), (8, 8, 160)]


def _available(wdm, lib, shape, prec):
    """variants compiled for each axis of a plan of this shape"""
    out = []
    with wdm.Plan(shape, X3, prec, lib=lib) as p:
        for axis in range(3):
            ok = []
            for v in range(wdm._capi.WDM_MAX_VARIANTS):
                sel = [0, 0, 0]
                sel[axis] = v
                try:
                    p.select_variants(sel)
                    ok.append(v)
                except wdm.WaveDMError:
                    pass
            out.append(ok)
    return out


@pytest.mark.parametrize("shape", VARIANT_SHAPES, ids=["x".join(map(str, s)) for s in VARIANT_SHAPES])
def test_every_variant_matches_oracle(wdm, nativelib, shape):
    psi = blob(shape, 3)
    ref = O.run_ssfm(psi, X3, 2, 0.01)
    checked = 0
    for prec, tol in (("f64", 1e-12), ("f32", 2e-5)):
        avail = _available(wdm, nativelib, shape, prec)
        for v in range(wdm._capi.WDM_MAX_VARIANTS):
            sel = [v if v in avail[a] else 0 for a in range(3)]
            if v > 0 and sel == [0, 0, 0]:
                continue
            with wdm.Plan(shape, X3, prec, lib=nativelib) as p:
                p.select_variants(sel)
                p.set_psi(psi)
                p.step(2, 0.01)
                err = np.linalg.norm(p.get_psi() - ref) / np.linalg.norm(ref)
            assert err < tol, (prec, sel, err)
            checked += 1
    assert checked >= 4


def test_autotune_keeps_state_and_parity(wdm, nativelib):
    shape = (64, 64, 64)
    psi = blob(shape, 4)
    ref = O.run_ssfm(psi, X3, 3, 0.01)
    with wdm.Plan(shape, X3, "f64", lib=nativelib) as p:
        p.set_psi(psi)
        v, ms = p.autotune(2, 0.01)
        assert all(0 <= x < wdm._capi.WDM_MAX_VARIANTS for x in v) and np.all(ms >= 0)
        assert np.array_equal(p.get_psi(), psi)
        p.step(3, 0.01)
        assert np.linalg.norm(p.get_psi() - ref) / np.linalg.norm(ref) < 1e-12


def test_ensemble_autotune_shares_choices(wdm, nativelib):
    plans = [_make(wdm, nativelib, (64, 16, 8), "f32", None, i) for i in range(3)]
    try:
        with wdm.Ensemble(plans) as e:
            e.autotune(2, [0.01] * 3)
            e.step(2, 0.01)
        assert all(np.all(np.isfinite(p.get_psi())) for p in plans)
    finally:
        for p in plans:
            p.close()
