"""wdm_poisson_noalloc, the allocation-free variant of wdm_poisson: device
buffers in and out, its work memory in the library's fixed device array.
Checked through the C ABI against wdm_poisson on a plan of the same grid
(same kernels and bit-identical tables: equal to the last bit), on two calls
in a row, the second on a smaller grid, in FP64 and FP32; a grid above the
array's size is refused; the device's free memory does not move across the
calls (the sequence runs twice; the reading is taken after the first, once
the device has loaded the library's code, which holds the array)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hip():
    return ctypes.CDLL("libamdhip64.so")


def _free_bytes(hip):
    free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


class DevBuf:
    """A device buffer from hipMalloc (made before the calls under test)."""

    def __init__(self, hip, nbytes):
        self.hip, self.nbytes = hip, nbytes
        self.ptr = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(nbytes)) == 0

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), 1) == 0

    def download(self, n):
        out = np.empty(n, dtype=np.float64)
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def _rho(shape, seed):
    rng = np.random.default_rng(seed)
    x = [np.linspace(-1, 1, n) for n in shape]
    X, Y, Z = np.meshgrid(*x, indexing="ij")
    return np.asfortranarray(np.exp(-4 * (X ** 2 + Y ** 2 + Z ** 2)) + 0.1 * rng.uniform(size=shape))


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_noalloc_poisson_equals_plan_poisson(wdm, prec):
    A = wdm._capi
    lib = A.load()
    hip = _hip()
    cases = [((128, 96, 64), (6.0, 5.0, 4.0), 3.5, 1), ((48, 32, 24), (3.0, 2.5, 2.0), 0.7, 2)]
    nmax = max(int(np.prod(s)) for s, _, _, _ in cases)
    rho_d, phi_d = DevBuf(hip, 8 * nmax), DevBuf(hip, 8 * nmax)
    try:
        refs = []
        for shape, X, scale, seed in cases:
            rho = _rho(shape, seed)
            with wdm.Plan(shape, X, prec) as p:
                refs.append((rho, p.poisson(rho, scale=scale).ravel(order="F")))
        # the sequence twice: the first pass lets the runtime load the
        # library's device code (which holds the fixed array) and size its
        # own copy staging; the second must leave the free memory unchanged
        free0 = None
        for rep in range(2):
            for (shape, X, scale, _), (rho, ref) in zip(cases, refs):
                n = rho.size
                rho_d.upload(rho.ravel(order="F"))
                rc = lib.wdm_poisson_noalloc(*shape, *X, A.WDM_F64 if prec == "f64" else A.WDM_F32,
                                             rho_d.ptr, phi_d.ptr, scale, None)
                assert rc == A.WDM_OK, lib.wdm_last_error(None)
                assert hip.hipDeviceSynchronize() == 0
                got = phi_d.download(n)
                print(f"{prec} {shape} pass {rep}: max |diff| {np.abs(got - ref).max():.3e} "
                      f"of max |phi| {np.abs(ref).max():.3e}", flush=True)
                assert np.array_equal(got, ref)
            if rep == 0:
                free0 = _free_bytes(hip)
        assert _free_bytes(hip) == free0, "the calls allocated device memory"
    finally:
        rho_d.free()
        phi_d.free()


def test_noalloc_refuses_oversized_grid_and_bad_arguments(wdm):
    A = wdm._capi
    lib = A.load()
    hip = _hip()
    buf = DevBuf(hip, 8 * 1024)
    try:
        # 384^3 FP64 needs a 227 MB half spectrum, above the 64 MiB array
        rc = lib.wdm_poisson_noalloc(384, 384, 384, 1.0, 1.0, 1.0, A.WDM_F64, buf.ptr, buf.ptr, 1.0, None)
        assert rc == A.WDM_ERR_UNSUPPORTED and b"scratch" in lib.wdm_last_error(None)
        assert lib.wdm_poisson_noalloc(8, 8, 8, 1.0, 1.0, 1.0, A.WDM_F64, buf.ptr, buf.ptr, 0.0, None) == A.WDM_ERR_ARG
        assert lib.wdm_poisson_noalloc(8, 8, 8, 1.0, 1.0, 1.0, A.WDM_F64, None, buf.ptr, 1.0, None) == A.WDM_ERR_ARG
        assert lib.wdm_poisson_noalloc(7, 8, 8, 1.0, 1.0, 1.0, A.WDM_F64, buf.ptr, buf.ptr, 1.0,
                                       None) == A.WDM_ERR_UNSUPPORTED
    finally:
        buf.free()
