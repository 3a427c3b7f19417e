"""wavedm-hip: MI355X backend for WaveDM.jl's SSFM kick-drift step and FFT
Poisson gravity.

The native library (lib/libwavedm_hip.so, C ABI in include/wavedm.h) does all
the arithmetic on the GPU; this package is the host-side mirror of the
reference's solver-step API (src/solver/KDK.jl, src/solver/gravity.jl,
src/parallel_poisson.jl, src/physics/poisson.jl, src/solver/fft_operators.jl)
over that ABI, used by the tests, the benchmark and Python callers.  The
directory name is not an identifier, so it is imported as ``wavedm_amd``.
"""
from . import _capi, adjoint, configs, ensemble, ic, kdk, snapshot, units
from ._capi import WaveDMError
from .ensemble import Ensemble
from .plan import Plan

__all__ = ["Ensemble", "Plan", "WaveDMError", "_capi", "adjoint", "configs", "ensemble", "ic", "kdk", "snapshot", "units"]
