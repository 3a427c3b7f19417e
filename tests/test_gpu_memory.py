"""Device memory and the non-finite guard.

The reference frees consumed device buffers so that its peak stays near one
psi (src/solver/KDK.jl:28-35,103-113).  Here the plan allocates everything at
creation: psi (N complex), the Poisson half spectrum (hpitch*Ny*Nz complex)
and per-axis tables; stepping allocates nothing, and the reductions
(wdm_reduce, wdm_cfl) run as read-only passes with no N-sized scratch.  The
device's free memory (hipMemGetInfo) is read around each phase.

wdm_step reports a psi that became non-finite as WDM_ERR_NONFINITE (a
host-side status from the step's own mass partials, not a device fault)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _free_bytes():
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def test_step_and_reduce_allocate_nothing_384(wdm, nativelib):
    cfg = wdm.configs.build("c2")
    n = int(np.prod(cfg["shape"]))
    nh = ((cfg["shape"][0] // 2 + 1 + 7) // 8 * 8) * cfg["shape"][1] * cfg["shape"][2]
    f0 = _free_bytes()
    with wdm.Plan(cfg["shape"], cfg["xmax"], "f32", hbar_over_m=cfg["hbar_over_m"],
                  four_pi_G=cfg["four_pi_G"], lib=nativelib) as p:
        wdm.configs.install(p, cfg)
        dt = 7e-5
        p.step(2, dt)
        f1 = _free_bytes()
        p.step(100, dt)
        f2 = _free_bytes()
        p.reduce()
        p.cfl(0.0)
        f3 = _free_bytes()
    f4 = _free_bytes()
    plan_bytes = f0 - f1
    floor = 8 * (n + nh)  # psi + half spectrum, FP32 complex
    print(f"plan {plan_bytes / 2**20:.1f} MiB (psi + half spectrum {floor / 2**20:.1f} MiB); "
          f"100 steps {(f1 - f2) / 2**20:.1f} MiB; reduce+cfl {(f2 - f3) / 2**20:.1f} MiB; "
          f"released {(f4 - f3) / 2**20:.1f} MiB", flush=True)
    assert floor <= plan_bytes <= floor + 64 * 2**20
    assert f2 == f1, "stepping allocated device memory"
    assert f3 == f2, "the reductions allocated device memory"
    assert f4 - f0 >= -8 * 2**20, "plan memory not returned"


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_nonfinite_psi_reported_by_step(wdm, nativelib, prec):
    shape = (16, 12, 8)
    rng = np.random.default_rng(2)
    psi = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    psi[3, 4, 5] = np.nan
    with wdm.Plan(shape, 2.0, prec, lib=nativelib) as p:
        p.set_psi(psi)
        with pytest.raises(wdm.WaveDMError) as e:
            p.step(1, 1e-3)
        assert e.value.code == wdm._capi.WDM_ERR_NONFINITE
        # the plan stays usable: a finite psi steps again
        p.set_psi(np.nan_to_num(psi))
        p.step(1, 1e-3)
        assert np.all(np.isfinite(p.get_psi()))
