"""Writes the golden fixtures of the SSFM step (tests/golden/*.npz).

Inputs are seeded here, outputs come from the FP64 oracle (oracle/ssfm_oracle.py),
so the fixtures pin later changes of the oracle and of the device path to the
same numbers.  The reference itself (Julia) cannot run in this image; the
oracle is pinned against it by tests/test_oracle.py.

    python tests/golden/make_golden.py

Fixtures (numpy .npz, no pickles):
  step_8x8x8.npz      psi0 and psi after 1, 2 and 10 steps, self-gravity
  step_16x12x8.npz    same on an uneven grid, with a Hernquist host and an absorber
  c0_summary.npz      the 64^3 C0 run after 100 steps: norms and 64 sampled values
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import ssfm_oracle as O  # noqa: E402

X3 = (3.0, 2.5, 3.5)
HQ = dict(M=20.0, a=1.3, c=(0.3, -0.2, 0.4))
FOUR_PI_G = 4 * np.pi * 0.3


def seeded_psi(shape, seed):
    rng = np.random.default_rng(seed)
    x, y, z, _, _ = O.setup_grid(*X3, *shape)
    r2 = x[:, None, None] ** 2 + y[None, :, None] ** 2 + z[None, None, :] ** 2
    amp = 2.0 * np.exp(-r2 / 3.0) + 0.2
    return np.asfortranarray(amp * np.exp(1j * rng.uniform(-np.pi, np.pi, shape)))


def hernquist_V(xxx, yyy, zzz, ps, t):
    G = FOUR_PI_G / (4 * np.pi)
    c = HQ["c"]
    return -G * HQ["M"] / (np.sqrt((xxx - c[0]) ** 2 + (yyy - c[1]) ** 2 + (zzz - c[2]) ** 2) + HQ["a"])


def steps(psi0, nsteps, dt, **kw):
    out = {}
    psi = psi0
    done = 0
    for n in nsteps:
        psi = O.run_ssfm(psi, X3, n - done, dt, t0=done * dt, four_pi_G=FOUR_PI_G, **kw)
        out[f"psi_{n}"] = np.asarray(psi)
        done = n
    return out


def c0_summary():
    # 64^3, L = 20 kpc, seeded soliton x 0.3 rad random phase, dt = 1e-3 Gyr, 100 steps (FP64)
    ip = os.path.join(os.path.dirname(os.path.dirname(HERE)), "wavedm.jl_amd")
    sys.path.insert(0, os.path.dirname(ip))
    import importlib.util
    spec = importlib.util.spec_from_file_location("wavedm_amd", os.path.join(ip, "__init__.py"),
                                                  submodule_search_locations=[ip])
    W = importlib.util.module_from_spec(spec)
    sys.modules["wavedm_amd"] = W
    spec.loader.exec_module(W)
    cfg = W.configs.build("c0")
    psi = O.run_ssfm(cfg["psi0"], cfg["xmax"], cfg["steps"], cfg["dt"], hbar_over_m=cfg["hbar_over_m"],
                     four_pi_G=cfg["four_pi_G"], be=O.Backend("scipy", workers=os.cpu_count()))
    rng = np.random.default_rng(64)
    idx = np.stack([rng.integers(0, 64, 64) for _ in range(3)], axis=1)
    vals = psi[idx[:, 0], idx[:, 1], idx[:, 2]]
    _, _, _, d, dV = O.setup_grid(*cfg["xmax"], 64, 64, 64)
    return dict(l2=np.linalg.norm(psi), maxabs=np.abs(psi).max(), mass=O.mass(psi, dV),
                sum_re=psi.real.sum(), sum_im=psi.imag.sum(), idx=idx, vals=vals,
                psi0_l2=np.linalg.norm(cfg["psi0"]))


def main():
    a = seeded_psi((8, 8, 8), 11)
    np.savez(os.path.join(HERE, "step_8x8x8.npz"), psi0=a, dt=0.05, **steps(a, (1, 2, 10), 0.05))
    b = seeded_psi((16, 12, 8), 12)
    border = np.linspace(0.97, 1.0, 16)
    np.savez(os.path.join(HERE, "step_16x12x8.npz"), psi0=b, dt=0.04, border=border,
             **steps(b, (1, 2, 10), 0.04, V=hernquist_V, border=border))
    np.savez(os.path.join(HERE, "c0_summary.npz"), **c0_summary())
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")


if __name__ == "__main__":
    main()
