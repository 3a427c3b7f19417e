import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_addoption(parser):
    parser.addoption("--emu", action="store_true", default=False,
                     help="run GPU-marked tests against the CPU emulation debug build "
                          "(libwavedm_cpuemu.so, needs the ASan runtime preloaded)")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (HIP library calls)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


def load_wavedm():
    """Import the host package that lives in the directory ``wavedm.jl_amd``
    (not a valid Python identifier) under the module name ``wavedm_amd``."""
    if "wavedm_amd" in sys.modules:
        return sys.modules["wavedm_amd"]
    pkg_dir = os.path.join(ROOT, "wavedm.jl_amd")
    spec = importlib.util.spec_from_file_location(
        "wavedm_amd", os.path.join(pkg_dir, "__init__.py"),
        submodule_search_locations=[pkg_dir])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["wavedm_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def wdm():
    return load_wavedm()


@pytest.fixture(scope="session")
def nativelib(wdm, request):
    """The native library the GPU tests drive: the HIP build, or the CPU
    emulation debug build when pytest runs with --emu."""
    if request.config.getoption("--emu"):
        return wdm._capi.load_emu()
    return wdm._capi.load()
