"""The 256^3 and 384^3 workloads at their full sizes against the oracle.

Three reference-split steps of each workload (c1: 256^3, 64 solitons,
Hernquist host, self-gravity; c2: 384^3, 128 solitons, rotating tidal
quadrupole, self-gravity; autoset dt) from the same initial state, on the
device and in the FP64 oracle (oracle/ssfm_oracle.run_ssfm, scipy FFTs): a
pass that mis-indexed lines the same way forward and inverse would keep every
size-independent property green but fails here.  Each workload is run with
the default pass tables, with the tables the autotuner picks on the
workload's own state, and, in FP32, with the y passes' 8-wave tables and
every axis on its variant 1 (the configurations the benchmark times).

Bars: relative L2 <= 1e-12 (FP64), <= 2e-5 (FP32, whose initial state is the
FP64 one rounded to complex64)."""
import os

import numpy as np
import pytest

from oracle import ssfm_oracle as O
from workload_v import oracle_V

pytestmark = pytest.mark.gpu

NSTEPS = 3
_REF = {}


def _reference(wdm, nativelib, name):
    if name not in _REF:
        cfg = wdm.configs.build(name)
        with wdm.Plan(cfg["shape"], cfg["xmax"], "f64", hbar_over_m=cfg["hbar_over_m"],
                      four_pi_G=cfg["four_pi_G"], lib=nativelib) as p:
            wdm.configs.install(p, cfg)
            dt, _ = wdm.configs.autoset_dt(p, cfg["steps"])
            psi0 = p.get_psi()
        print(f"{name} {cfg['shape']}: oracle, {NSTEPS} steps", flush=True)
        be = O.Backend("scipy", workers=min(16, os.cpu_count() or 1))
        ref = O.run_ssfm(psi0, cfg["xmax"], NSTEPS, dt, hbar_over_m=cfg["hbar_over_m"],
                         four_pi_G=cfg["four_pi_G"], V=oracle_V(cfg), be=be)
        _REF.clear()  # one workload's arrays at a time
        _REF[name] = (cfg, dt, psi0, ref)
    return _REF[name]


CASES = [(n, prec, v) for n in ("c1", "c2") for prec in ("f64", "f32") for v in ("default", "tuned")]
CASES += [("c2", "f32", (0, 1, 0)), ("c2", "f32", (1, 1, 1)), ("c1", "f32", (1, 1, 1))]
IDS = [f"{n}-{p}-{v if isinstance(v, str) else ''.join(map(str, v))}" for n, p, v in CASES]


@pytest.mark.parametrize("name,prec,variants", CASES, ids=IDS)
def test_fullsize_matches_oracle(wdm, nativelib, name, prec, variants):
    cfg, dt, psi0, ref = _reference(wdm, nativelib, name)
    tol = 1e-12 if prec == "f64" else 2e-5
    with wdm.Plan(cfg["shape"], cfg["xmax"], prec, hbar_over_m=cfg["hbar_over_m"],
                  four_pi_G=cfg["four_pi_G"], lib=nativelib) as p:
        wdm.configs.install(p, cfg)
        p.set_psi(psi0)
        if variants == "tuned":
            chosen, _ = p.autotune(2, dt)
            p.set_psi(psi0)
        elif variants != "default":
            p.select_variants(variants)
            chosen = list(variants)
        else:
            chosen = [0, 0, 0]
        p.step(NSTEPS, dt)
        out = p.get_psi()
    err = float(np.linalg.norm(out - ref) / np.linalg.norm(ref))
    print(f"  {name} {prec} variants {chosen}: rel L2 vs oracle {err:.3e}", flush=True)
    assert err < tol, (chosen, err)
