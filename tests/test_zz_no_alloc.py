"""The FFT entry point that allocates nothing (``sp_fft_noalloc``).

It runs the library's own transform plan on device buffers the caller owns, with its
twiddle tables and the four-step intermediate in a static device array, so its results
must equal the allocating testing entry's (``sp_testing_fft``) to the bit: the same
passes, the same tables, the same order of arithmetic. Two calls in a row check that the
array's leftovers from a larger call do not leak into a smaller one, and a third returns
to the first length (the tables rebuilt again).
"""

import ctypes as C

import numpy as np
import pytest

from solprop_amd import _lib


def _ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class _Hip:
    """The caller's device buffers, from the HIP runtime the library itself links (torch
    carries a runtime of its own, which is not mixed in here)."""

    def __init__(self):
        _lib.lib()
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]

    def check(self, rc):
        assert rc == 0, f"HIP error {rc}"

    def alloc(self, nbytes):
        p = C.c_void_p()
        self.check(self.rt.hipMalloc(C.byref(p), nbytes))
        return p

    def upload(self, dst, a):
        self.check(self.rt.hipMemcpy(dst, a.ctypes.data_as(C.c_void_p), a.nbytes, 1))

    def download(self, a, src):
        self.check(self.rt.hipMemcpy(a.ctypes.data_as(C.c_void_p), src, a.nbytes, 2))


def _reference(x, sign):
    lib = _lib.lib()
    out = np.empty_like(x)
    msg = C.create_string_buffer(256)
    batch, n = x.shape
    rc = lib.sp_testing_fft(_ptr(x), _ptr(out), n, batch, sign, msg, 256)
    assert rc == _lib.SP_OK, msg.value
    return out


def test_no_alloc_refuses_sizes_beyond_its_array():
    """Refused before any device call: more than 2^20 points in all, or rows above 2^16."""
    lib = _lib.lib()
    msg = C.create_string_buffer(256)
    fake = C.c_void_p(16)  # never dereferenced: the size check comes first
    rc = lib.sp_fft_noalloc(fake, fake, 1 << 17, 1, -1, None, msg, 256)
    assert rc == _lib.SP_VALUE and b"at most" in msg.value
    rc = lib.sp_fft_noalloc(fake, fake, 1 << 12, 257, -1, None, msg, 256)
    assert rc == _lib.SP_VALUE and b"at most" in msg.value
    rc = lib.sp_fft_noalloc(fake, fake, 1000, 1, -1, None, msg, 256)
    assert rc == _lib.SP_VALUE


@pytest.mark.gpu
def test_no_alloc_fft_matches_the_allocating_entry():
    lib = _lib.lib()
    hip = _Hip()
    msg = C.create_string_buffer(256)
    rng = np.random.default_rng(11)
    # Four-step (8192 points: the intermediate in the static array), then a smaller
    # one-pass transform, then the first length again in place with the other sign.
    cases = ((8192, 16, -1, False), (2048, 4, -1, False), (8192, 3, +1, True))
    for n, batch, sign, in_place in cases:
        print(f"sp_fft_noalloc n={n} batch={batch} sign={sign} in_place={in_place}",
              flush=True)
        x = rng.standard_normal((batch, n)) + 1j * rng.standard_normal((batch, n))
        expected = _reference(x, sign)
        source = hip.alloc(x.nbytes)
        target = source if in_place else hip.alloc(x.nbytes)
        try:
            hip.upload(source, x)
            if not in_place:
                hip.check(hip.rt.hipMemset(target, 0xFF, x.nbytes))  # NaN everywhere
            rc = lib.sp_fft_noalloc(source, target, n, batch, sign, None, msg, 256)
            assert rc == _lib.SP_OK, msg.value
            got = np.empty_like(x)
            hip.download(got, target)  # on the default stream, after the transform
        finally:
            hip.check(hip.rt.hipFree(source))
            if not in_place:
                hip.check(hip.rt.hipFree(target))
        assert np.array_equal(got, expected)
        ref = np.fft.fft(x, axis=1) if sign < 0 else np.fft.ifft(x, axis=1) * n
        assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
