"""Build driver for the wavedm native libraries.

    python build.py [hip|emu|all] [--jobs N] [--incremental]

* ``hip``: lib/libwavedm_hip.so, the product, HIP for gfx950 (hipcc).
* ``emu``: lib/libwavedm_cpuemu.so, a CPU fiber-emulation build of the same
  kernel source compiled with g++ and AddressSanitizer.  Debug/test only: it
  exists so the host-side tests can bounds-check every global and LDS access
  of the kernels; nothing in the package loads it.

One translation unit per supported axis length (the WDM_FOR_EACH_SIZE table in
wdm_ops.hpp) so the kernel instantiations compile in parallel.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import re
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB = os.path.join(PKG, "lib")
BUILD = os.path.join(PKG, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("WDM_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: a multiply is fused into an add only where the source
# writes fmaT.  Left to the compiler, contraction is decided per kernel, so two
# kernels that evaluate the same expression (the density of the fused inverse
# pass and of the standalone one, a batched pass and its single-plan twin)
# could round it differently; with it off they agree bit for bit.
HIP_FLAGS = ["-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]
EMU_FLAGS = ["-O1", "-g", "-std=c++17", "-fPIC", "-fsanitize=address", "-fno-omit-frame-pointer",
             "-Wno-unknown-pragmas", "-Wno-ignored-attributes"]


def sizes():
    src = open(os.path.join(HERE, "wdm_ops.hpp")).read()
    block = src.split("#define WDM_FOR_EACH_SIZE(X)")[1].split("\n\n")[0]
    return [tuple(int(v) for v in m) for m in re.findall(r"X\((\d+),\s*(\d+),\s*(\d+)\)", block)]


def _deps_mtime():
    t = 0.0
    for d in (HERE, os.path.join(HERE, "emu"), os.path.join(os.path.dirname(PKG), "include")):
        for f in os.listdir(d):
            if f.endswith((".hpp", ".h", ".cpp", ".inl")):
                t = max(t, os.path.getmtime(os.path.join(d, f)))
    return t


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return cmd[-1]


def _jobs(kind, incremental):
    out = os.path.join(BUILD, kind)
    os.makedirs(out, exist_ok=True)
    dep = _deps_mtime()
    jobs, objs = [], []
    if kind == "hip":
        base = [HIPCC] + HIP_FLAGS + ["-x", "hip", "-c"]
    else:
        base = ["g++"] + EMU_FLAGS + ["-c"]
    units = [("wdm_plan", os.path.join(HERE, "wdm_plan.cpp"), [])]
    for (n, a, b) in sizes():
        units.append((f"wdm_inst_{n}", os.path.join(HERE, "wdm_inst.cpp"),
                      [f"-DWDM_INST_N={n}", f"-DWDM_INST_A={a}", f"-DWDM_INST_B={b}"]))
    if kind == "emu":
        units.append(("emu_runtime", os.path.join(HERE, "emu", "emu_runtime.cpp"), []))
    for name, src, defs in units:
        obj = os.path.join(out, name + ".o")
        objs.append(obj)
        if incremental and os.path.exists(obj) and os.path.getmtime(obj) > dep:
            continue
        jobs.append(base + defs + [src, "-o", obj])
    return jobs, objs


def build(kind="hip", jobs=None, incremental=False, verbose=True, defines=(), out=None):
    """defines/out: extra -D flags and an alternate output .so (tuning variants)."""
    global BUILD
    jobs = jobs or min(16, os.cpu_count() or 4)
    saved = BUILD
    if defines:
        BUILD = os.path.join(saved, "var_" + "_".join(d.replace("-D", "").replace("=", "") for d in defines))
    cmds, objs = _jobs(kind, incremental)
    BUILD = saved
    cmds = [c[:1] + list(defines) + c[1:] for c in cmds]
    if verbose:
        print(f"[build:{kind}] compiling {len(cmds)} units with {jobs} jobs", flush=True)
    with cf.ThreadPoolExecutor(jobs) as ex:
        for done in ex.map(_run, cmds):
            if verbose:
                print(f"[build:{kind}] {os.path.basename(done)}", flush=True)
    os.makedirs(LIB, exist_ok=True)
    if out:
        target = out
        os.makedirs(os.path.dirname(out), exist_ok=True)
        link = ([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}"] if kind == "hip"
                else ["g++", "-shared", "-fPIC", "-fsanitize=address"]) + ["-o", target + ".tmp"] + objs
    elif kind == "hip":
        target = os.path.join(LIB, "libwavedm_hip.so")
        link = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", target + ".tmp"] + objs
    if not out and kind == "emu":
        target = os.path.join(LIB, "libwavedm_cpuemu.so")
        link = ["g++", "-shared", "-fPIC", "-fsanitize=address", "-o", target + ".tmp"] + objs
    _run(link)
    shutil.move(target + ".tmp", target)
    if verbose:
        print(f"[build:{kind}] -> {target}", flush=True)
    return target


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("what", nargs="?", default="hip", choices=["hip", "emu", "all"])
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--incremental", action="store_true")
    a = ap.parse_args(argv)
    kinds = ["hip", "emu"] if a.what == "all" else [a.what]
    for k in kinds:
        build(k, a.jobs, a.incremental)


if __name__ == "__main__":
    sys.exit(main())
