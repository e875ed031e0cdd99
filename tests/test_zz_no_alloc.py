"""The allocation-free r2c (fans_gpu_fft_r2c_noalloc) against fans_gpu_fft_r2c.

The no-allocation entry point takes device buffers and a stream from the caller
and keeps its twiddle tables in a file-scope device array; its spectrum must
equal the allocating entry point's (FP64, 1e-13 relative to the spectrum's
largest entry: the two differ only in how the twiddles are rounded, by an ulp).
Two calls in a row, the second on a smaller grid, check that the shared array
is rewritten per call; an over-sized grid is refused before any launch."""
import ctypes as C

import numpy as np
import pytest

from fans_amd import gpu

pytestmark = pytest.mark.gpu


def _fn():
    f = gpu.lib().fans_gpu_fft_r2c_noalloc
    f.argtypes = [C.c_int32] * 5 + [C.c_void_p, C.c_void_p, C.c_void_p]
    f.restype = C.c_int
    return f


def _hip():
    """The HIP runtime the library itself links (torch bundles another one, and
    two runtimes in one process do not share devices): plain device buffers and
    a stream for the test, which the entry point then must not allocate."""
    gpu.lib()
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    h.hipStreamSynchronize.argtypes = [C.c_void_p]
    h.hipStreamDestroy.argtypes = [C.c_void_p]
    return h


def _noalloc_r2c(x):
    hip = _hip()
    nx, ny, nz, h = x.shape
    out = np.zeros((nx, ny, nz // 2 + 1, h), dtype=np.complex128)
    din, dout, st = C.c_void_p(), C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(din), x.nbytes) == 0 and hip.hipMalloc(C.byref(dout), out.nbytes) == 0
    assert hip.hipStreamCreate(C.byref(st)) == 0
    try:
        assert hip.hipMemcpy(din, x.ctypes.data, x.nbytes, 1) == 0  # host to device
        rc = _fn()(0, nx, ny, nz, h, din, dout, st)
        assert rc == 0, gpu.lib().fans_gpu_last_error(None)
        assert hip.hipStreamSynchronize(st) == 0
        assert hip.hipMemcpy(out.ctypes.data, dout, out.nbytes, 2) == 0  # device to host
    finally:
        hip.hipStreamDestroy(st)
        hip.hipFree(din)
        hip.hipFree(dout)
    return out


def test_noalloc_r2c_matches_allocating_entry_point():
    rng = np.random.default_rng(21)
    for shape in [(16, 32, 64, 3), (8, 12, 20, 1)]:  # second call: smaller grid, radix 3 / 5
        print(f"[gpu-case] fft_r2c_noalloc shape={shape}", flush=True)
        x = rng.standard_normal(shape)
        ref = gpu.fft_r2c(x)
        got = _noalloc_r2c(x)
        err = float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))
        assert err < 1e-13, (shape, err)
        assert float(np.max(np.abs(got - np.fft.rfftn(x, axes=(0, 1, 2)))) / np.max(np.abs(ref))) < 1e-13


def test_noalloc_rejects_grids_beyond_the_scratch_array():
    rc = _fn()(0, 8192, 8192, 8192, 3, C.c_void_p(16), C.c_void_p(16), None)
    assert rc == -1
    assert b"scratch array holds" in gpu.lib().fans_gpu_last_error(None)
