"""ctypes binding of the C ABI in include/wavedm.h.

``load()`` opens the in-tree HIP library ``lib/libwavedm_hip.so`` and raises if
it is missing: there is no CPU fallback in the package.  ``load_emu()`` opens
the CPU fiber-emulation debug build, which only the host-side tests use to
bounds-check the kernels (it must be loaded in a process started with the
AddressSanitizer runtime preloaded).
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")
HIP_LIB = os.path.join(LIB_DIR, "libwavedm_hip.so")
EMU_LIB = os.path.join(LIB_DIR, "libwavedm_cpuemu.so")

# constants mirrored from include/wavedm.h
WDM_OK = 0
WDM_ERR_ARG = -1
WDM_ERR_OOM = -2
WDM_ERR_HIP = -3
WDM_ERR_NONFINITE = -4
WDM_ERR_UNSUPPORTED = -5
WDM_ERR_STATE = -6
WDM_F64 = 0
WDM_F32 = 1
WDM_C128 = 0
WDM_C64 = 1
WDM_GRID_CONSISTENT = 1
WDM_SPLIT_REFERENCE = 0
WDM_SPLIT_STRANG = 1
WDM_V_HERNQUIST = 2
WDM_V_QUADRUPOLE = 4
WDM_V_KAPPA = 8
WDM_V_TIDAL = 16
WDM_RED_MASS = 0
WDM_RED_KINETIC = 1
WDM_RED_WSELF = 2
WDM_RED_VEXT = 3
WDM_RED_PSIMAX = 4
WDM_RED_PHIMAX = 5
WDM_RED_COUNT = 6
WDM_ADJ_DT_SCALE = 1
WDM_ADJ_NPARAM = 12
WDM_DIAG_MASS = 0
WDM_DIAG_RHOMAX = 1
WDM_DIAG_PEAK_IX = 2
WDM_DIAG_R = 5
WDM_DIAG_R1 = 6
WDM_DIAG_MOMX = 15
WDM_DIAG_PE_ABS = 18
WDM_DIAG_KE = 19
WDM_DIAG_QE = 20
WDM_DIAG_VIRIAL = 21
WDM_DIAG_COUNT = 22
WDM_DIAG_BINS = 4096
WDM_MAX_VARIANTS = 3
WDM_NOALLOC_SCRATCH_BYTES = 64 << 20

# every symbol include/wavedm.h declares: name -> (restype, argtypes)
_P = C.c_void_p
_D = C.c_double
_I = C.c_int
_DP = C.POINTER(C.c_double)
SIGNATURES = {
    "wdm_version": (_I, []),
    "wdm_last_error": (C.c_char_p, [_P]),
    "wdm_plan_create": (_I, [C.POINTER(_P), _I, _I, _I, _D, _D, _D, _I, C.c_uint]),
    "wdm_plan_destroy": (_I, [_P]),
    "wdm_set_coeffs": (_I, [_P, _D, _D]),
    "wdm_set_self_gravity": (_I, [_P, _I]),
    "wdm_set_psi": (_I, [_P, _P, _I]),
    "wdm_get_psi": (_I, [_P, _P, _I]),
    "wdm_set_psi_device": (_I, [_P, _P, _I]),
    "wdm_get_psi_device": (_I, [_P, _P, _I]),
    "wdm_set_border": (_I, [_P, _DP, _I]),
    "wdm_set_static_potential": (_I, [_P, _DP]),
    "wdm_set_potential_model": (_I, [_P, _I, _DP, _I]),
    "wdm_clear_potential_models": (_I, [_P]),
    "wdm_set_tidal_model": (_I, [_P, _DP, _I, _I, _I, _DP, _DP, _DP, _DP, _I]),
    "wdm_poisson": (_I, [_P, _DP, _DP, _D]),
    "wdm_get_phi": (_I, [_P, _DP]),
    "wdm_kick": (_I, [_P, _D, _D]),
    "wdm_drift": (_I, [_P, _D]),
    "wdm_step": (_I, [_P, _I, _D, _D, _D, _I]),
    "wdm_reduce": (_I, [_P, _DP, _I]),
    "wdm_diagnostics": (_I, [_P, _DP, _I]),
    "wdm_acceleration": (_I, [_P, _DP, _DP, _DP]),
    "wdm_density_accumulate": (_I, [_P]),
    "wdm_density_average": (_I, [_P, _DP, C.POINTER(C.c_longlong)]),
    "wdm_density_reset": (_I, [_P]),
    "wdm_density_sum": (_I, [_P, _DP, C.POINTER(C.c_longlong)]),
    "wdm_density_restore": (_I, [_P, _DP, C.c_longlong]),
    "wdm_set_mass_ledger": (_I, [_P, _I]),
    "wdm_get_mass_ledger": (_I, [_P, C.POINTER(_I), _DP]),
    "wdm_set_mass_reference": (_I, [_P, _D]),
    "wdm_poisson_noalloc": (_I, [_I, _I, _I, _D, _D, _D, _I, _P, _P, _D, _P]),
    "wdm_ic_acceleration": (_I, [_P, _DP, _DP, _DP, _DP, _DP, _DP]),
    "wdm_ic_velocity": (_I, [_P, _DP, _DP, _DP, _DP, _I]),
    "wdm_step_timed_mask": (_I, [_P, _I, C.c_double, C.c_double, C.c_double, C.c_uint, _DP, _I, _DP]),
    "wdm_cfl": (_I, [_P, _D, _DP]),
    "wdm_synchronize": (_I, [_P]),
    "wdm_step_timed": (_I, [_P, _I, _D, _D, _D, _DP, _I, _DP]),
    "wdm_ic_solitons": (_I, [_P, _DP, _I]),
    "wdm_adjoint_forward": (_I, [_P, _I, _D, _D, _D, _I]),
    "wdm_adjoint_backward": (_I, [_P, _P, _I, _P, _DP, _I, C.c_uint]),
    "wdm_autotune": (_I, [_P, _I, _D, C.POINTER(C.c_int), _DP]),
    "wdm_select_variants": (_I, [_P, C.POINTER(C.c_int)]),
    "wdm_ensemble_create": (_I, [C.POINTER(_P), C.POINTER(_P), _I]),
    "wdm_ensemble_destroy": (_I, [_P]),
    "wdm_ensemble_step": (_I, [_P, _I, _DP, _DP, _D, _I]),
    "wdm_ensemble_autotune": (_I, [_P, _I, _DP]),
    "wdm_ensemble_layout": (_I, [_P, C.POINTER(_I), C.POINTER(_I)]),
}
PASS_NAMES = ["x_kick_fwd", "y_fwd", "z_drift", "y_inv", "x_inv_density", "poisson_y_fwd", "poisson_z",
              "poisson_y_inv"]

_cache: dict[str, C.CDLL] = {}


class WaveDMError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def _bind(path: str) -> C.CDLL:
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def load(path: str | None = None) -> C.CDLL:
    """The HIP library; raises if it has not been built."""
    path = path or HIP_LIB
    if path not in _cache:
        if not os.path.exists(path):
            raise ImportError(
                f"{path} not found: build it with `python wavedm.jl_amd/csrc/build.py hip` "
                "(or __graft_entry__.build()); the package has no CPU fallback")
        _cache[path] = _bind(path)
    return _cache[path]


def load_emu() -> C.CDLL:
    """CPU emulation debug build (tests only)."""
    if EMU_LIB not in _cache:
        if not os.path.exists(EMU_LIB):
            raise ImportError(f"{EMU_LIB} not found: `python wavedm.jl_amd/csrc/build.py emu`")
        _cache[EMU_LIB] = _bind(EMU_LIB)
    return _cache[EMU_LIB]


def check(lib: C.CDLL, plan, rc: int):
    if rc != WDM_OK:
        msg = lib.wdm_last_error(plan)
        raise WaveDMError(rc, msg.decode() if msg else "unknown error")
