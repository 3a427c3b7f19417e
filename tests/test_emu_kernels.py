"""Host-side bounds check of the kernels: the GPU parity suite run against the
CPU fiber-emulation build of the same kernel source, compiled with
AddressSanitizer (every global and LDS access is checked).  Runs in a
subprocess because the ASan runtime must be preloaded.  Debug build only: the
package never loads it; the GPU run of the same suite is the parity claim."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SUBSET = ("roundtrip or (poisson and (8x8x8 or 32x16x8 or 160x16x8 or 12x8x384)) or "
          "(step_fp64 and (16x16x16 or 128x8x16 or 8x384x16)) or kick_and_drift or strang or "
          "quadrupole or static_potential or nls or reductions or error_paths or "
          "(adjoint_matches_oracle or checkpoint_interval or adjoint_error_paths or adjoint_tape or (quadrupole_gradients and 16x12x8)) or "
          "ensemble_equals or recaptures or time_dependent or (every_variant and (64x8x8 or 32x16x8 or 8x8x384 or 8x8x192)) or "
          "device_matches_golden or (diagnostics_match_oracle and (32x24x16 or 16x12x8 or 96x48x12)) or "
          "diagnostics_after or diagnostics_error or restart_continues or restart_with_mass or mass_reference or "
          "density_average or "
          "projection_error_paths or projection_reference_follows or projection_enabled or ic_velocity or "
          "ic_acceleration or velocity_ic_chain or baryon_mesh or (tidal_kicks and 32x24x16) or tidal_refusals")


def test_gpu_suite_on_emulated_kernels(wdm):
    emu = wdm._capi.EMU_LIB
    if not os.path.exists(emu):
        pytest.fail(f"{emu} missing: run __graft_entry__.build()")
    gxx = lambda f: subprocess.run(["g++", f"-print-file-name={f}"], capture_output=True,  # noqa: E731
                                   text=True).stdout.strip()
    env = dict(os.environ)
    env["LD_PRELOAD"] = f"{gxx('libasan.so')} {gxx('libstdc++.so')}"
    env["ASAN_OPTIONS"] = "detect_leaks=0"
    r = subprocess.run([sys.executable, "-m", "pytest", os.path.join(ROOT, "tests", "test_gpu_parity.py"),
                        os.path.join(ROOT, "tests", "test_gpu_adjoint.py"),
                        os.path.join(ROOT, "tests", "test_gpu_ensemble.py"),
                        os.path.join(ROOT, "tests", "test_golden.py"),
                        os.path.join(ROOT, "tests", "test_gpu_diag.py"),
                        os.path.join(ROOT, "tests", "test_gpu_snapshot.py"),
                        os.path.join(ROOT, "tests", "test_gpu_ledger.py"),
                        os.path.join(ROOT, "tests", "test_gpu_ic.py"),
                        os.path.join(ROOT, "tests", "test_gpu_tidal.py"),
                        "-m", "gpu", "--emu", "-q", "-x", "-p", "no:cacheprovider", "-k", SUBSET],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    tail = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "WARNING: ASan" not in l)[-3000:]
    assert r.returncode == 0, tail
    assert " passed" in r.stdout, tail
