"""Benchmark: cell-steps/s of the fused FP32 SSFM step at 384^3 on one MI355X.

Workload (the c2 configuration): 384^3 grid, L = 20 kpc, 128 seeded solitons,
self-gravity plus a rotating tidal quadrupole, autoset dt, 1000 KDK steps.
A "step" is one reference-split iteration: kick (Poisson-consistent Phi of the
current psi + V(x,t)) and drift, i.e. the 8 pass kernels.

Printed JSON line: value = cell-steps/s of the FP32 run (inputs resident in
HBM, timed region = exactly --steps steps between synchronisations), plus
  roofline      dominant kernel, algorithmic bytes / average launch time vs 8 TB/s
  cpu_baseline  the oracle's restatement of the reference CPU step (numpy/scipy
                pocketfft, FP64, all host cores) on a bounded sample of the
                same workload
  fp64          the FP64 (plain form) run of the same steps, its speed and the
                FP32-vs-FP64 psi error, mass and energy drifts.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import math
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "wavedm.jl_amd")
HBM_PEAK_GBS = 8000.0


def load_pkg():
    if "wavedm_amd" in sys.modules:
        return sys.modules["wavedm_amd"]
    spec = importlib.util.spec_from_file_location(
        "wavedm_amd", os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["wavedm_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def pass_bytes(shape, elem_bytes):
    """Algorithmic HBM bytes per launch of each pass: every pass reads and
    writes its array once; the Poisson chain works on the (Nx/2+1) half
    spectrum."""
    nx, ny, nz = shape
    n = nx * ny * nz
    nh = (nx // 2 + 1) * ny * nz
    psi_rw = 2 * n * elem_bytes
    half_rw = 2 * nh * elem_bytes
    return {
        "x_kick_fwd": psi_rw + nh * elem_bytes,       # psi r/w + Phi_hat read
        "y_fwd": psi_rw,
        "z_drift": psi_rw,
        "y_inv": psi_rw,
        "x_inv_density": psi_rw + nh * elem_bytes,    # psi r/w + rho_hat write
        "poisson_y_fwd": half_rw,
        "poisson_z": half_rw,
        "poisson_y_inv": half_rw,
    }


def measured_traffic(shape, kernel):
    """HBM bytes per launch of one FP32 pass from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (profiles/traffic_384.json, scaled by the
    copy-kernel calibration), or None when there is none for this grid."""
    path = os.path.join(ROOT, "profiles", "traffic_384.json")
    if tuple(shape) != (384, 384, 384) or not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)["bytes_per_launch"].get("FP32 " + kernel)


def energy(W, red, dt):
    """The energy the reference split (one kick of dt/2 per drift of dt)
    conserves: K + (dt_kick/dt)(W_self + V_ext) (wavedm_amd.kdk.split_energy)."""
    return W.kdk.split_energy(red, dt, 0.5 * dt)


def run_gpu(W, cfg, precision, steps, warmup, ledger=False, diagnostics=True, round_ic=False):
    C = W.configs
    with W.Plan(cfg["shape"], cfg["xmax"], precision, hbar_over_m=cfg["hbar_over_m"],
                four_pi_G=cfg["four_pi_G"]) as p:
        C.install(p, cfg)
        dt, lims = (cfg["dt"], None) if cfg["dt"] else C.autoset_dt(p, cfg["steps"])
        if warmup:
            p.step(warmup, dt)
            C.install(p, cfg)
        p.set_mass_ledger(ledger)
        if round_ic:  # the same initial state rounded to complex64 (a perturbation of FP32 rounding size)
            p.set_psi(p.get_psi(np.complex64))
        variants, _ = p.autotune(3, dt)  # per-axis pass tables measured on this state (state kept)
        # per-pass breakdown with events around every kernel (untimed: each event
        # pair adds a few microseconds between kernels), then the timed region
        # with events on the dominant pass only
        npre = min(steps, 100)
        breakdown, _ = p.step_timed(npre, dt)
        dom = max(breakdown, key=breakdown.get)
        red0 = p.reduce()
        p.synchronize()
        t0 = time.perf_counter()
        timed, ev_ms = p.step_timed(steps, dt, t0=npre * dt, passes=[dom])
        p.synchronize()
        wall = time.perf_counter() - t0
        passes = {k: v / npre * steps for k, v in breakdown.items()}  # per-launch averages, scaled to `steps`
        passes[dom] = timed[dom]
        red1 = p.reduce()
        if not diagnostics:
            return dict(dt=dt, wall=wall, red0=red0, red1=red1, variants=variants,
                        psi=p.get_psi(np.complex128) if round_ic else None)
        # the reference loop's per-step diagnostics on the final state (device)
        diag_ms = {}
        for label, nslots in (("full", W._capi.WDM_DIAG_COUNT), ("without_potential", W._capi.WDM_DIAG_PE_ABS)):
            ts = []
            for _ in range(3):
                t1 = time.perf_counter()
                diag = p.diagnostics(nslots)
                ts.append(time.perf_counter() - t1)
            diag_ms[label] = 1e3 * sorted(ts)[1]
        diag = p.diagnostics()
        psi = p.get_psi(np.complex128)
    return dict(dt=dt, cfl=lims, wall=wall, event_ms=ev_ms, passes=passes, red0=red0, red1=red1, psi=psi,
                diag=diag, diag_ms=diag_ms, variants=variants)


def cpu_baseline(W, cfg, dt, nsteps):
    """The oracle's op-for-op restatement of the reference CPU step, timed on a
    bounded number of steps of the same workload (FP64, scipy pocketfft with
    one worker per host core)."""
    sys.path.insert(0, ROOT)
    from oracle import ssfm_oracle as O
    with W.Plan(cfg["shape"], cfg["xmax"], "f64", hbar_over_m=cfg["hbar_over_m"],
                four_pi_G=cfg["four_pi_G"]) as p:
        W.configs.install(p, cfg)
        psi0 = p.get_psi()
    V = None
    pot = cfg["potential"]
    if pot is not None and pot[0] == "quadrupole":
        kw = pot[1]
        T0, axis, period, cen = np.asarray(kw["T0"]), np.asarray(kw["axis"]), kw["period"], kw["center"]

        def V(xxx, yyy, zzz, ps, t):
            n = axis / np.linalg.norm(axis)
            K = np.array([[0, -n[2], n[1]], [n[2], 0, -n[0]], [-n[1], n[0], 0]])
            a = 2 * math.pi * t / period
            R = np.eye(3) + math.sin(a) * K + (1 - math.cos(a)) * K @ K
            T = R @ T0 @ R.T
            d = (xxx - cen[0], yyy - cen[1], zzz - cen[2])
            return 0.5 * sum(T[i, j] * d[i] * d[j] for i in range(3) for j in range(3))
    elif pot is not None and pot[0] == "hernquist":
        kw = pot[1]
        G = cfg["four_pi_G"] / (4 * math.pi)

        def V(xxx, yyy, zzz, ps, t):
            c = kw["center"]
            return -G * kw["M"] / (np.sqrt((xxx - c[0]) ** 2 + (yyy - c[1]) ** 2 + (zzz - c[2]) ** 2) + kw["a"])
    cores = os.cpu_count() or 1
    be = O.Backend("scipy", workers=cores)
    # one untimed step (planner / page-in), then the timed sample
    O.run_ssfm(psi0, cfg["xmax"], 1, dt, hbar_over_m=cfg["hbar_over_m"], four_pi_G=cfg["four_pi_G"], V=V, be=be)
    t0 = time.perf_counter()
    O.run_ssfm(psi0, cfg["xmax"], nsteps, dt, hbar_over_m=cfg["hbar_over_m"], four_pi_G=cfg["four_pi_G"],
               V=V, be=be)
    sec = time.perf_counter() - t0
    n = int(np.prod(cfg["shape"]))
    model = ""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        model = next((l.split(":", 1)[1].strip() for l in out.splitlines() if l.startswith("Model name")), "")
    except Exception:
        pass
    return {"value": n * nsteps / sec, "unit": "cell-steps/s", "cores": cores, "kind": "port",
            "sample": f"{nsteps} steps of the 384^3 c2 workload (after 1 untimed), FP64, "
                      f"oracle restatement of the reference CPU step with scipy pocketfft "
                      f"({cores} workers) standing in for FFTW; {sec / nsteps:.2f} s/step; CPU: {model}",
            "ms_per_step": 1e3 * sec / nsteps}


def copy_calibration(mib, reps):
    """A plain device-to-device copy of known size: the reference point for
    the HBM byte counters (FETCH_SIZE / WRITE_SIZE) and the achievable copy
    bandwidth.  Not part of the benchmark line."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(reps):
        b.copy_(a)
    t1.record()
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1) / reps
    print(json.dumps({"copy_bytes_each_way": 4 * n, "reps": reps, "avg_ms": ms,
                      "gbs_read_plus_write": 2 * 4 * n / (ms * 1e-3) / 1e9}), flush=True)


def adjoint_leg(W, steps=None):
    """Cost of the gradient on the adjoint workload (c4: 128^3, 16 solitons,
    Hernquist host, 200 FP32 steps, checkpoints every 20): wall time of the
    taped forward + backward sweep against one plain forward run."""
    cfg = W.configs.build("c4")
    nsteps = steps or cfg["steps"]
    with W.Plan(cfg["shape"], cfg["xmax"], "f32", hbar_over_m=cfg["hbar_over_m"],
                four_pi_G=cfg["four_pi_G"]) as p:
        W.configs.install(p, cfg)
        dt, _ = W.configs.autoset_dt(p, nsteps)
        p.autotune(3, dt)   # pass tables for this grid (kept state); forward and backward share them
        psi0 = p.get_psi(np.complex64)
        hq = cfg["potential"][1]
        loss = W.adjoint.ProfileLoss(cfg["shape"], cfg["xmax"], nbins=40)
        # target: the same run under seeded-perturbed host parameters
        p.set_hernquist(hq["M"] * 1.05, hq["a"] * 0.95, (hq["center"][0] + 0.5, 0.2, -0.3))
        p.set_psi(psi0)
        p.step(nsteps, dt)
        target = loss.profile(p.get_psi())
        p.clear_potential_models()
        p.set_hernquist(hq["M"], hq["a"], hq["center"])
        # warm up both paths once, then time
        p.set_psi(psi0)
        p.step(2, dt)
        W.adjoint.value_and_grad(p, psi0, 2, dt, loss, target, ckpt_every=20)
        p.set_psi(psi0)
        p.synchronize()
        t0 = time.perf_counter()
        p.step(nsteps, dt)
        p.synchronize()
        t_fwd = time.perf_counter() - t0
        p.set_psi(psi0)
        t0 = time.perf_counter()
        p.adjoint_forward(nsteps, dt, None, 0.0, 20)
        p.synchronize()
        t_tape = time.perf_counter() - t0
        L, gN = loss(p.get_psi(), target)
        gN = np.asfortranarray(gN.astype(np.complex64))  # the plan's layout: no conversion inside the timing
        t0 = time.perf_counter()
        g0, gp = p.adjoint_backward(gN, dt_scale=True)
        t_bwd = time.perf_counter() - t0
        # the same sweep without the dt-scale spectral term (fused conjugate drift)
        t0 = time.perf_counter()
        p.adjoint_backward(gN, dt_scale=False)
        t_bwd_nodt = time.perf_counter() - t0
    return {"workload": f"c4: FP32 128^3, 16 solitons + Hernquist host, {nsteps} steps, checkpoints every 20, "
                        "40-bin radial profile loss, gradients w.r.t. psi0 and (M, a, centre, dt-scale)",
            "forward_ms": 1e3 * t_fwd, "taped_forward_ms": 1e3 * t_tape, "backward_ms": 1e3 * t_bwd,
            "backward_over_forward": t_bwd / t_fwd, "gradient_over_forward": (t_tape + t_bwd) / t_fwd,
            "backward_ms_without_dt_scale": 1e3 * t_bwd_nodt,
            "backward_over_forward_without_dt_scale": t_bwd_nodt / t_fwd,
            "loss": L, "grad_params": dict(zip(W.adjoint.PARAM_NAMES, [float(v) for v in gp])),
            "grad_psi0_norm": float(np.linalg.norm(g0))}


def ensemble_leg(W, steps=None):
    """The batched workload (c3: 48 boxes, 16 each of 64^3, 128^3, 160^3, FP32,
    self-gravity): autotune per shape, then time the steps as one HIP-graph
    launch per step and as eager per-member launches."""
    cfg = W.configs.build("c3")
    nsteps = steps or cfg["steps"]
    plans = []
    try:
        dts = []
        for m in cfg["members"]:
            p = W.Plan(m["shape"], m["xmax"], "f32", hbar_over_m=m["hbar_over_m"], four_pi_G=m["four_pi_G"])
            plans.append(p)
            W.configs.install(p, m)
            dts.append(W.configs.autoset_dt(p, nsteps)[0])
        cells = sum(p.ncell for p in plans)
        psi0 = [p.get_psi(np.complex64) for p in plans]
        out = {"workload": f"c3: 48 FP32 boxes (16 x 64^3, 16 x 128^3, 16 x 160^3), self-gravity, "
                           f"{nsteps} steps, autoset dt per box", "cells": cells}
        with W.Ensemble(plans) as e:
            t0 = time.perf_counter()
            e.autotune(2, dts)
            out["autotune_s"] = time.perf_counter() - t0
            for mode in ("graph", "eager"):
                for p, s0 in zip(plans, psi0):
                    p.set_psi(s0)
                e.step(2, dts, graph=mode == "graph")   # capture / warm-up
                for p, s0 in zip(plans, psi0):
                    p.set_psi(s0)
                for p in plans:
                    p.synchronize()
                t0 = time.perf_counter()
                e.step(nsteps, dts, graph=mode == "graph")
                wall = time.perf_counter() - t0
                out[mode] = {"ms_per_step": 1e3 * wall / nsteps, "cell_steps_per_s": cells * nsteps / wall,
                             "hbm_frac_112B": 112.0 * cells * nsteps / wall / (HBM_PEAK_GBS * 1e9)}
        out["mass_finite"] = bool(all(np.isfinite(p.reduce(1)["mass"]) for p in plans))
        return out
    finally:
        for p in plans:
            p.close()


def c1_leg(W, steps=None):
    """The c1 workload (256^3, 64 solitons, static Hernquist host,
    self-gravity, autoset dt, 500 steps): FP32 (with and without the mass
    projection) against FP64 from the same complex64-rounded initial state,
    and FP64 from the unrounded state (how much the dynamics amplify a
    rounding-sized perturbation).  Reports psi relative L2, relative mass drift
    and relative drift of the split's energy K + (dt_kick/dt)(W + V_ext),
    which the static potential conserves up to the splitting error."""
    cfg = W.configs.build("c1")
    nsteps = steps or cfg["steps"]
    res = {}
    with W.Plan(cfg["shape"], cfg["xmax"], "f64", hbar_over_m=cfg["hbar_over_m"],
                four_pi_G=cfg["four_pi_G"]) as p:
        W.configs.install(p, cfg)
        dt, _ = W.configs.autoset_dt(p, cfg["steps"])
        psi_exact = p.get_psi()
    psi0 = psi_exact.astype(np.complex64)
    for label, prec, ledger, start in (("fp64", "f64", False, psi0), ("fp64_unrounded_ic", "f64", False, psi_exact),
                                       ("fp32", "f32", False, psi0), ("fp32_mass_projection", "f32", True, psi0)):
        with W.Plan(cfg["shape"], cfg["xmax"], prec, hbar_over_m=cfg["hbar_over_m"],
                    four_pi_G=cfg["four_pi_G"]) as p:
            W.configs.install(p, cfg)
            p.set_psi(start)
            p.set_mass_ledger(ledger)
            r0 = p.reduce()
            p.synchronize()
            t0 = time.perf_counter()
            p.step(nsteps, dt)
            p.synchronize()
            wall = time.perf_counter() - t0
            r1 = p.reduce()
            e0, e1 = energy(W, r0, dt), energy(W, r1, dt)
            res[label] = {"ms_per_step": 1e3 * wall / nsteps, "mass_drift": abs(r1["mass"] - r0["mass"]) / r0["mass"],
                          "energy_drift": abs(e1 - e0) / abs(e0), "energy0": e0, "energy1": e1,
                          "psi": p.get_psi()}
    ref = res["fp64"].pop("psi")
    for k in ("fp64_unrounded_ic", "fp32", "fp32_mass_projection"):
        res[k]["psi_rel_l2_vs_fp64"] = float(np.linalg.norm(res[k].pop("psi") - ref) / np.linalg.norm(ref))
    res["workload"] = (f"c1: 256^3, L=40 kpc, 64 solitons + static Hernquist host, self-gravity, {nsteps} "
                       f"steps, dt={dt:.6g} Gyr (autoset)")
    res["speedup_fp32_vs_fp64"] = res["fp64"]["ms_per_step"] / res["fp32"]["ms_per_step"]
    res["speedup_fp32_projection_vs_fp64"] = res["fp64"]["ms_per_step"] / res["fp32_mass_projection"]["ms_per_step"]
    return res


def error_curve(W, name="c2", checkpoints=(10, 20, 50, 100, 200, 500, 1000)):
    """psi relative L2 against FP64 along a run: FP32 (with the mass
    projection and without) and FP64 started from the complex64-rounded state
    (the dynamics' own amplification of a rounding-sized perturbation), all
    against FP64 from the exact device initial state, at each checkpoint."""
    cfg = W.configs.build(name)
    mk = lambda prec: W.Plan(cfg["shape"], cfg["xmax"], prec, hbar_over_m=cfg["hbar_over_m"],  # noqa: E731
                             four_pi_G=cfg["four_pi_G"])
    plans = {"fp64": mk("f64"), "fp64_rounded_ic": mk("f64"), "fp32": mk("f32"), "fp32_mass_projection": mk("f32")}
    try:
        for p in plans.values():
            W.configs.install(p, cfg)
        dt, _ = W.configs.autoset_dt(plans["fp64"], cfg["steps"])
        psi0 = plans["fp64"].get_psi().astype(np.complex64)
        for k in ("fp64_rounded_ic", "fp32", "fp32_mass_projection"):
            plans[k].set_psi(psi0)
        plans["fp32_mass_projection"].set_mass_ledger(True)
        rows, done = [], 0
        m0 = {k: p.reduce(1)["mass"] for k, p in plans.items()}
        for c in checkpoints:
            for p in plans.values():
                p.step(c - done, dt, t0=done * dt)
            done = c
            ref = plans["fp64"].get_psi()
            nref = np.linalg.norm(ref)
            row = {"step": c}
            for k in ("fp64_rounded_ic", "fp32", "fp32_mass_projection"):
                row[k] = float(np.linalg.norm(plans[k].get_psi() - ref) / nref)
                row[k + "_mass_drift"] = abs(plans[k].reduce(1)["mass"] - m0[k]) / m0[k]
            del ref
            rows.append(row)
            print(json.dumps(row), file=sys.stderr, flush=True)
        return {"workload": f"{name} {cfg['shape']}, dt={dt:.6g} Gyr", "rows": rows}
    finally:
        for p in plans.values():
            p.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-adjoint", action="store_true")
    ap.add_argument("--skip-ensemble", action="store_true")
    ap.add_argument("--skip-c1", action="store_true")
    ap.add_argument("--error-curve", action="store_true",
                    help="only run the psi-error-versus-step curve of the workload (FP32 and rounded-IC FP64 "
                         "against FP64) and print it as JSON")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this library runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU slab decomposition is a design note, not implemented", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration, max(args.steps, 1))
        return
    W = load_pkg()
    if args.error_curve:
        print(json.dumps(error_curve(W, args.config)), flush=True)
        return
    cfg = W.configs.build(args.config)
    n = int(np.prod(cfg["shape"]))
    print(f"[bench] {cfg['name']} {cfg['shape']} FP32 warmup={args.warmup} steps={args.steps}", file=sys.stderr,
          flush=True)
    r32 = run_gpu(W, cfg, "f32", args.steps, args.warmup)
    ms32 = 1e3 * r32["wall"] / args.steps
    value = n * args.steps / r32["wall"]
    bytes32 = pass_bytes(cfg["shape"], 8)
    step_bytes32 = sum(bytes32.values())
    dom = max(r32["passes"], key=r32["passes"].get)
    dom_avg_ms = r32["passes"][dom] / args.steps
    achieved = bytes32[dom] / (dom_avg_ms * 1e-3) / 1e9
    passes = {k: {"avg_us": 1e3 * v / args.steps, "bytes": bytes32[k],
                  "gbs": bytes32[k] / (v / args.steps * 1e-3) / 1e9 if v > 0 else None}
              for k, v in r32["passes"].items()}
    out = {
        "metric": "cell-steps/s + HBM GB/s at 384^3 (FP32 vs FP64 SSFM); adjoint/forward cost",
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms32,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded soliton superposition, device-generated)",
        "config": {"workload": f"{cfg['name']}: FP32 SSFM {cfg['shape'][0]}^3, L={2 * cfg['xmax'][0]:g} kpc, "
                               f"128 solitons, self-gravity + rotating tidal quadrupole, autoset dt, "
                               f"{args.steps} reference-split KDK steps",
                   "grid": list(cfg["shape"]), "dt_gyr": r32["dt"], "cfl_limits": r32["cfl"],
                   "pass_variants_xyz": list(r32["variants"])},
        "hbm_gbs_step": step_bytes32 / (ms32 * 1e-3) / 1e9,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(cfg["shape"], dom),
                     "algorithmic_bytes": bytes32[dom],
                     "step_frac": step_bytes32 / (ms32 * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "passes": passes,
        "passes_note": "timed region: HIP events on the dominant pass only; the other passes from a "
                       "preceding 100-step run with events on every kernel",
    }
    m0, m1 = r32["red0"]["mass"], r32["red1"]["mass"]
    out["fp32"] = {"mass_drift": abs(m1 - m0) / m0, "energy0": energy(W, r32["red0"], r32["dt"]),
                   "energy1": energy(W, r32["red1"], r32["dt"]),
                   "energy_note": "K + (dt_kick/dt)(W_self + V_ext), the split's energy; the c2 tidal field "
                                  "rotates, so it is not conserved here (see the c1 leg)"}
    print("[bench] FP32 leg with the mass projection", file=sys.stderr, flush=True)
    rl = run_gpu(W, cfg, "f32", args.steps, min(args.warmup, 5), ledger=True, diagnostics=False)
    out["fp32"]["mass_projection"] = {
        "ms_per_step": 1e3 * rl["wall"] / args.steps,
        "mass_drift": abs(rl["red1"]["mass"] - rl["red0"]["mass"]) / rl["red0"]["mass"],
        "note": "wdm_set_mass_ledger(1): scale from the x pass's mass partials applied in the next z pass"}
    out["diagnostics"] = {"ms_per_call": r32["diag_ms"], "values": r32["diag"],
                          "note": "wdm_diagnostics on the final FP32 state: mass, peak, R, R1..R9, momenta, "
                                  "PE_abs/KE/QE/Virial (full includes the Phi_all solve)"}
    if not args.skip_fp64:
        print("[bench] FP64 leg", file=sys.stderr, flush=True)
        r64 = run_gpu(W, cfg, "f64", args.steps, min(args.warmup, 5))
        ms64 = 1e3 * r64["wall"] / args.steps
        err = float(np.linalg.norm(r32["psi"] - r64["psi"]) / np.linalg.norm(r64["psi"]))
        e32, e64 = energy(W, r32["red1"], r32["dt"]), energy(W, r64["red1"], r64["dt"])
        out["fp64"] = {"ms_per_step": ms64, "value": n * args.steps / r64["wall"],
                       "speedup_fp32_vs_fp64": ms64 / ms32,
                       "speedup_fp32_projection_vs_fp64": ms64 / out["fp32"]["mass_projection"]["ms_per_step"],
                       "psi_rel_l2_fp32_vs_fp64": err,
                       "mass_drift_fp64": abs(r64["red1"]["mass"] - r64["red0"]["mass"]) / r64["red0"]["mass"],
                       "energy_rel_diff_fp32_vs_fp64": abs(e32 - e64) / abs(e64),
                       "hbm_gbs_step": 2 * step_bytes32 / (ms64 * 1e-3) / 1e9}
        # sensitivity of the workload: FP64 from the complex64-rounded initial state
        print("[bench] FP64 leg from the FP32-rounded initial state", file=sys.stderr, flush=True)
        rr = run_gpu(W, cfg, "f64", args.steps, min(args.warmup, 5), diagnostics=False, round_ic=True)
        out["fp64"]["psi_rel_l2_fp64_from_rounded_ic"] = float(
            np.linalg.norm(rr["psi"] - r64["psi"]) / np.linalg.norm(r64["psi"]))
        del rr
        del r64
    del r32
    if not args.skip_c1:
        print("[bench] c1 leg (256^3, 500 steps, FP32 vs FP64)", file=sys.stderr, flush=True)
        out["c1"] = c1_leg(W)
    if not args.skip_adjoint:
        print("[bench] adjoint leg (c4)", file=sys.stderr, flush=True)
        out["adjoint"] = adjoint_leg(W)
    if not args.skip_ensemble:
        print("[bench] ensemble leg (c3)", file=sys.stderr, flush=True)
        out["ensemble"] = ensemble_leg(W)
    if not args.skip_cpu:
        print("[bench] CPU baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(W, cfg, out["config"]["dt_gyr"], args.cpu_steps)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
