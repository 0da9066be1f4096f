"""fhd_fft_r2c_noalloc, the entry point that allocates nothing, called through the C ABI with
device buffers the test allocates itself, against fhd_fft_r2c (which allocates its own context):
two calls in a row, the largest grid it accepts and then a smaller one, and a grid beyond its
work space, which it must reject."""
import ctypes as C

import numpy as np
import pytest

import fluam_amd as F

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    """The HIP runtime libfhd.so is linked against (already loaded with it), for the test's own
    device buffers."""
    F.lib()
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


def _noalloc():
    L = F.lib()
    L.fhd_fft_r2c_noalloc.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    L.fhd_fft_r2c_noalloc.restype = C.c_int
    return L


def _run_noalloc(h, L, x):
    mz, my, mx = x.shape
    P = mx // 2 + 1
    x32 = np.ascontiguousarray(x, dtype=np.float32)
    out = np.zeros((mz, my, P), dtype=np.complex64)
    d_in, d_out = C.c_void_p(), C.c_void_p()
    assert h.hipMalloc(C.byref(d_in), x32.nbytes) == 0
    assert h.hipMalloc(C.byref(d_out), out.nbytes) == 0
    try:
        assert h.hipMemcpy(d_in, x32.ctypes.data, x32.nbytes, H2D) == 0
        ok = L.fhd_fft_r2c_noalloc(mx, my, mz, d_in, d_out, None)
        assert ok == 1, L.fhd_last_error().decode()
        assert h.hipDeviceSynchronize() == 0
        assert h.hipMemcpy(out.ctypes.data, d_out, out.nbytes, D2H) == 0
    finally:
        h.hipFree(d_in)
        h.hipFree(d_out)
    return out


def test_noalloc_fft_equals_fft_r2c_twice():
    h, L = _hip(), _noalloc()
    rng = np.random.default_rng(7)
    for shape in ((128, 128, 128), (24, 32, 48)):  # (mz, my, mx): the largest accepted, then smaller
        print(f"[gpu-case] fhd_fft_r2c_noalloc shape={shape}", flush=True)
        x = rng.standard_normal(shape).astype(np.float32)
        got = _run_noalloc(h, L, x)
        ref = F.fft_r2c(x.astype(np.float64))  # the allocating entry point, same FP32 inputs
        scale = np.abs(ref).max()
        err = np.abs(got.astype(np.complex128) - ref).max() / scale
        print(f"  max |noalloc - fft_r2c| / max |X| = {err:.2e}, bit-equal {np.array_equal(got.astype(np.complex128), ref)}",
              flush=True)
        assert err < 1e-6
        # and the transform itself, against numpy in FP64
        npy = np.fft.rfftn(x.astype(np.float64))
        assert np.abs(got - npy).max() / np.abs(npy).max() < 1e-5


def test_noalloc_fft_rejects_a_grid_beyond_its_work_space():
    h, L = _hip(), _noalloc()
    d = C.c_void_p()
    assert h.hipMalloc(C.byref(d), 16) == 0
    try:
        # 192 * 128 * 65 complex values: a valid grid, larger than the work space
        assert L.fhd_fft_r2c_noalloc(128, 128, 192, d, d, None) == 0
        assert b"exceeds" in L.fhd_last_error()
    finally:
        h.hipFree(d)
