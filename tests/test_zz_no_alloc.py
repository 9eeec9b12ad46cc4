"""ConvolveDrfYpadNoAlloc (the allocation-free variant of ConvolveDrfYpad,
staged in the library's file-scope device array) against ConvolveDrfYpad:
bit-identical results on two calls in a row, the second smaller, sizes past
its limits refused with a message, and a valid call working after that."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(plane, psf, fft=0, pad=0, avg=0):
    import libirl_amd as L
    lib = L.load()
    lib.ConvolveDrfYpadNoAlloc.restype = ctypes.c_int
    lib.ConvolveDrfYpadNoAlloc.argtypes = [L._FP, ctypes.c_int, ctypes.c_int,
                                           ctypes.POINTER(L.Drf)] + [ctypes.c_int] * 3 + [L._FP]
    psf = np.ascontiguousarray(psf, np.float32)
    d = L.Drf((psf.shape[1] + 1) // 2, (psf.shape[0] + 1) // 2, 0,
              psf.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    plane = np.ascontiguousarray(plane, np.float32)
    out = np.full_like(plane, np.nan)
    rc = lib.ConvolveDrfYpadNoAlloc(L.fptr(plane), plane.shape[1], plane.shape[0],
                                    ctypes.byref(d), fft, pad, avg, L.fptr(out))
    return rc, out


def test_no_alloc_convolution_equals_the_allocating_entry_point():
    import libirl_amd as L
    from oracle import oracle as O
    rng = np.random.default_rng(12)
    cases = [((96, 128), (31, 31), 3, 2), ((24, 40), (15, 9), 0, 0)]   # second call smaller
    for shape, pshape, pad, avg in cases:
        print("no-alloc convolution plane %s psf %s pad %d avg %d" % (shape, pshape, pad, avg),
              flush=True)
        a = rng.random(shape).astype(np.float32)
        p = rng.random(pshape).astype(np.float32)
        p /= p.sum()
        rc, got = _noalloc(a, p, 0, pad, avg)
        assert rc == 0, L.error_string()
        want = L.convolve_drf_ypad(a, p, 0, pad, avg)
        assert L.last_error() == 0
        assert np.array_equal(got, want)
        ref = O.conv2d_ypad(a, p, pad, avg, real="f64")
        assert np.linalg.norm(got - ref) <= 1e-5 * np.linalg.norm(ref)


def test_no_alloc_convolution_refuses_sizes_past_its_limits():
    import libirl_amd as L
    rc, _ = _noalloc(np.ones((300, 300), np.float32), np.ones((3, 3), np.float32) / 9)
    assert rc != 0
    assert "at most 65536" in L.error_string()
    rc, _ = _noalloc(np.ones((16, 16), np.float32), np.ones((65, 3), np.float32))
    assert rc != 0 and "63 x 63" in L.error_string()
    # a valid call after the refusals
    a = np.random.default_rng(1).random((20, 30)).astype(np.float32)
    p = np.full((5, 7), 1 / 35.0, np.float32)
    rc, got = _noalloc(a, p, 1)
    assert rc == 0 and np.array_equal(got, L.convolve_drf_ypad(a, p, 1))
