"""The C-ABI library loads on a host without a GPU and exports every symbol
include/wavedm.h declares; error paths that need no device."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    src = open(os.path.join(ROOT, "include", "wavedm.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(wdm_\w+)\s*\(", src)))


def test_header_and_binding_agree(wdm):
    assert header_symbols() == sorted(wdm._capi.SIGNATURES)


@pytest.mark.parametrize("which", ["hip", "emu"])
def test_library_exports_every_symbol(wdm, which):
    path = wdm._capi.HIP_LIB if which == "hip" else wdm._capi.EMU_LIB
    if not os.path.exists(path):
        pytest.fail(f"{path} missing: run __graft_entry__.build()")
    if which == "emu":
        # do not dlopen the ASan build into this process: read its dynamic symbols
        import subprocess
        out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True).stdout
        names = set(line.split()[-1] for line in out.splitlines() if line.strip())
        missing = [s for s in header_symbols() if s not in names]
    else:
        lib = ctypes.CDLL(path)
        missing = [s for s in header_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_version_and_argument_errors_without_device(wdm):
    A = wdm._capi
    lib = A.load()
    assert lib.wdm_version() == 1
    h = ctypes.c_void_p()
    assert lib.wdm_plan_create(ctypes.byref(h), 0, 8, 8, 1.0, 1.0, 1.0, A.WDM_F64, 0) == A.WDM_ERR_ARG
    assert b"bad grid size" in lib.wdm_last_error(None)
    assert lib.wdm_plan_create(ctypes.byref(h), 8, 8, 8, -1.0, 1.0, 1.0, A.WDM_F64, 0) == A.WDM_ERR_ARG
    assert lib.wdm_plan_create(ctypes.byref(h), 8, 8, 8, 1.0, 1.0, 1.0, 7, 0) == A.WDM_ERR_ARG
    assert lib.wdm_plan_create(None, 8, 8, 8, 1.0, 1.0, 1.0, 0, 0) == A.WDM_ERR_ARG
    assert lib.wdm_step(None, 1, 1.0, 0.5, 0.0, 0) == A.WDM_ERR_ARG


def test_plan_create_rejects_arrays_past_32bit_offsets(wdm):
    """The pass kernels form byte offsets in 32 bits: a plan whose psi (or half
    spectrum) would reach 4 GiB is refused before any device call.  1024^3
    has 2^30 cells (below the 2^31-cell limit) but 8 GiB of FP32 psi; 646^3
    FP64 and 813^3 FP32 are the first cubes past 2^32 bytes."""
    A = wdm._capi
    lib = A.load()
    h = ctypes.c_void_p()
    for n, prec in ((1024, A.WDM_F32), (812, A.WDM_F64), (646, A.WDM_F64), (813, A.WDM_F32)):
        rc = lib.wdm_plan_create(ctypes.byref(h), n, n, n, 1.0, 1.0, 1.0, prec, 0)
        assert rc == A.WDM_ERR_UNSUPPORTED, (n, prec, rc)
        assert b"4 GiB" in lib.wdm_last_error(None)
        assert not h.value
