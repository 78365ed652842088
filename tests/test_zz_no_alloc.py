"""tbd_dft_forward_noalloc (device buffers in, device buffers out, the
library's fixed scratch array instead of device allocations) against
tbd_dft_forward on the same inputs, through the C ABI.  The test's own device
buffers for the inputs and outputs come from the HIP runtime the library
links (plumbing only)."""
import ctypes

import numpy as np
import pytest

from turbid_amd import runtime as RT

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    RT.lib()   # loads libamdhip64 as the library's dependency
    h = ctypes.CDLL("libamdhip64.so", mode=ctypes.RTLD_GLOBAL)
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


class _Dev:
    """A device buffer of the test (hipMalloc / hipFree)."""

    def __init__(self, nbytes):
        self.h = _hip()
        self.p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(self.p), nbytes) == 0
        self.n = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.n
        assert self.h.hipMemcpy(self.p, a.ctypes.data, self.n, H2D) == 0

    def get(self, dtype):
        out = np.empty(self.n // np.dtype(dtype).itemsize, dtype=dtype)
        assert self.h.hipMemcpy(out.ctypes.data, self.p, self.n, D2H) == 0
        return out

    def free(self):
        if self.p:
            self.h.hipFree(self.p)
            self.p = ctypes.c_void_p()


def _noalloc(phys, fp32):
    nv, nz, ny, nx = phys.shape
    _, nyc, mx = RT.dims(nx, ny, nz)
    src = np.ascontiguousarray(phys, dtype=np.float64)
    dev_in, dev_out = _Dev(src.nbytes), _Dev(16 * nv * mx * nyc * nz)
    try:
        dev_in.put(src)
        # the output starts as NaN: every element must be written
        dev_out.put(np.full(2 * nv * mx * nyc * nz, np.nan))
        RT.dft_forward_noalloc(nx, ny, nz, nv, dev_in.p.value, dev_out.p.value, fp32=fp32)
        return dev_out.get(np.float64).view(np.complex128).reshape(nv, mx, nyc, nz)
    finally:
        dev_in.free()
        dev_out.free()


def _field(shape, seed):
    return np.random.default_rng(seed).standard_normal(shape)


def _check(phys, fp32):
    ref = RT.dft_forward(phys, fp32=fp32)
    got = _noalloc(phys, fp32)
    assert np.all(np.isfinite(got))
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    # FP64: the twiddles are computed on the device instead of the host, so
    # the two may differ in the last bit of a twiddle; FP32: the same FP32
    # arithmetic on the same rounded inputs
    tol = 1e-6 if fp32 else 1e-13
    print("noalloc %s fp32=%d rel max err %.3e" % (phys.shape, fp32, err), flush=True)
    assert err < tol, err
    return err


@pytest.mark.parametrize("fp32", [0, 1])
def test_two_calls_second_smaller(fp32):
    _check(_field((3, 33, 36, 48), 1), fp32)
    _check(_field((1, 17, 18, 24), 2), fp32)


@pytest.mark.parametrize("fp32", [0, 1])
def test_several_plane_chunks(fp32):
    # one z-plane of three components at 256 x 256 takes 1.1 MB (FP64) or
    # 1.7 MB (FP32) of the 64 MiB array: 65 planes need two chunks
    _check(_field((3, 65, 256, 256), 3), fp32)
    _check(_field((2, 9, 48, 32), 4), fp32)


def test_size_beyond_the_array_is_rejected():
    small = _Dev(256)
    try:
        # one plane of 3 components at 3072 x 3072 needs 151 MB of scratch:
        # rejected before anything is launched
        with pytest.raises(RT.TurbidError, match="limit"):
            RT.dft_forward_noalloc(3072, 3072, 4, 3, small.p.value, small.p.value)
    finally:
        small.free()
    # the entry point still works after the rejection
    _check(_field((1, 5, 12, 12), 5), 0)
