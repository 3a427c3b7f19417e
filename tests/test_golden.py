"""Golden fixtures (tests/golden/, written by tests/golden/make_golden.py from
the FP64 oracle): the oracle must keep reproducing them, and the device path
must match them (FP64 plan <= 1e-12 relative, FP32 plan <= 2e-5)."""
import os

import numpy as np
import pytest

from oracle import ssfm_oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
X3 = (3.0, 2.5, 3.5)
FOUR_PI_G = 4 * np.pi * 0.3
HQ = dict(M=20.0, a=1.3, c=(0.3, -0.2, 0.4))


def _load(name):
    with np.load(os.path.join(GOLD, name), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def _V(xxx, yyy, zzz, ps, t):
    G = FOUR_PI_G / (4 * np.pi)
    c = HQ["c"]
    return -G * HQ["M"] / (np.sqrt((xxx - c[0]) ** 2 + (yyy - c[1]) ** 2 + (zzz - c[2]) ** 2) + HQ["a"])


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


CASES = [("step_8x8x8.npz", False), ("step_16x12x8.npz", True)]


@pytest.mark.parametrize("name,host", CASES, ids=["8x8x8", "16x12x8-hernquist-border"])
def test_oracle_reproduces_golden(name, host):
    g = _load(name)
    kw = dict(V=_V, border=g["border"]) if host else {}
    psi = O.run_ssfm(g["psi0"], X3, 10, float(g["dt"]), four_pi_G=FOUR_PI_G, **kw)
    assert rel(psi, g["psi_10"]) < 1e-13


def test_oracle_reproduces_c0_summary(wdm):
    g = _load("c0_summary.npz")
    cfg = wdm.configs.build("c0")
    assert abs(np.linalg.norm(cfg["psi0"]) - g["psi0_l2"]) <= 1e-14 * g["psi0_l2"]
    psi = O.run_ssfm(cfg["psi0"], cfg["xmax"], cfg["steps"], cfg["dt"], hbar_over_m=cfg["hbar_over_m"],
                     four_pi_G=cfg["four_pi_G"], be=O.Backend("scipy", workers=os.cpu_count()))
    idx = g["idx"]
    assert np.allclose(psi[idx[:, 0], idx[:, 1], idx[:, 2]], g["vals"], rtol=1e-12, atol=0)
    assert abs(np.linalg.norm(psi) - g["l2"]) <= 1e-13 * g["l2"]


@pytest.mark.gpu
@pytest.mark.parametrize("name,host", CASES, ids=["8x8x8", "16x12x8-hernquist-border"])
def test_device_matches_golden(wdm, nativelib, name, host):
    g = _load(name)
    for prec, tol in (("f64", 1e-12), ("f32", 2e-5)):
        with wdm.Plan(g["psi0"].shape, X3, prec, four_pi_G=FOUR_PI_G, lib=nativelib) as p:
            if host:
                p.set_hernquist(HQ["M"], HQ["a"], HQ["c"])
                p.set_border(g["border"])
            p.set_psi(g["psi0"])
            done = 0
            for n in (1, 2, 10):
                p.step(n - done, float(g["dt"]), t0=done * float(g["dt"]))
                done = n
                if prec == "f64" or n <= 2:
                    assert rel(p.get_psi(), g[f"psi_{n}"]) < tol, (prec, n)


@pytest.mark.gpu
def test_device_matches_c0_summary(wdm, nativelib):
    g = _load("c0_summary.npz")
    cfg = wdm.configs.build("c0")
    with wdm.Plan(cfg["shape"], cfg["xmax"], "f64", hbar_over_m=cfg["hbar_over_m"], four_pi_G=cfg["four_pi_G"],
                  lib=nativelib) as p:
        p.set_psi(cfg["psi0"])
        p.step(cfg["steps"], cfg["dt"])
        psi = p.get_psi()
    idx = g["idx"]
    assert np.allclose(psi[idx[:, 0], idx[:, 1], idx[:, 2]], g["vals"], rtol=1e-12, atol=1e-12 * g["maxabs"])
    assert abs(np.linalg.norm(psi) - g["l2"]) <= 1e-12 * g["l2"]
