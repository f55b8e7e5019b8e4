"""gf_site_frequency_spectrum_noalloc: the sample SFS of a device row into
device memory, with no allocation, against gf_site_frequency_spectrum (the
host-buffer entry point) bit for bit, on two calls in a row, the second with
a smaller row and a smaller sample; plus its argument checks.  Device
memory for the test's own inputs and outputs comes straight from the HIP
runtime (the library is not asked to allocate)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class Hip:
    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so")
        for name in ("hipMalloc", "hipFree", "hipMemcpy", "hipDeviceSynchronize", "hipStreamCreate",
                     "hipStreamDestroy", "hipStreamSynchronize"):
            getattr(self.lib, name).restype = C.c_int
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.lib.hipStreamDestroy.argtypes = [C.c_void_p]
        self.lib.hipStreamSynchronize.argtypes = [C.c_void_p]

    def ok(self, rc):
        assert rc == 0, "HIP error %d" % rc

    def malloc(self, nbytes):
        p = C.c_void_p()
        self.ok(self.lib.hipMalloc(C.byref(p), max(int(nbytes), 4)))
        return p

    def upload(self, arr):
        p = self.malloc(arr.nbytes)
        self.ok(self.lib.hipMemcpy(p, arr.ctypes.data_as(C.c_void_p), arr.nbytes, 1))   # host to device
        return p

    def download(self, p, out):
        self.ok(self.lib.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, 2))    # device to host
        return out


def _row(rng, M, N):
    row = np.concatenate([rng.integers(0, N + 1, M // 2) / N,
                          np.exp(rng.uniform(np.log(1.0 / N), 0.0, M - M // 2))]).astype(np.float32)
    row[:4] = [0.0, 1.0, 0.5, 1.0 / N]
    return row


def test_noalloc_sfs_equals_host_entry_point(gpu):
    lib = gpu.load()
    lib.gf_site_frequency_spectrum_noalloc.restype = C.c_int32
    lib.gf_site_frequency_spectrum_noalloc.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_float, C.c_void_p,
                                                       C.c_void_p]
    hip = Hip()
    stream = C.c_void_p()
    hip.ok(hip.lib.hipStreamCreate(C.byref(stream)))
    rng = np.random.default_rng(71)
    L = 5e8
    try:
        # two calls in a row, the second smaller (its scratch still holds the
        # first call's sums), then an empty row
        for M, n, N in [(400_000, 1000, 200_000), (3_000, 50, 20_000), (0, 7, 100)]:
            print("noalloc sfs M=%d n=%d" % (M, n), flush=True)
            row = _row(rng, M, N) if M else np.zeros(0, np.float32)
            d_row = hip.upload(row) if M else hip.malloc(4)
            d_out = hip.malloc(4 * n)
            try:
                rc = lib.gf_site_frequency_spectrum_noalloc(d_row, M, n, L, d_out, stream)
                assert rc == 0, lib.gf_last_error().decode()
                hip.ok(hip.lib.hipStreamSynchronize(stream))
                got = hip.download(d_out, np.empty(n, np.float32))
            finally:
                hip.lib.hipFree(d_row)
                hip.lib.hipFree(d_out)
            want = gpu.site_frequency_spectrum(row, n, L)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (M, n)
        # sizes beyond the maximum and bad arguments are refused before any work
        d_out = hip.malloc(4 * 8)
        try:
            assert lib.gf_site_frequency_spectrum_noalloc(d_out, 8, 65537, L, d_out, stream) != 0
            assert b"maximum" in lib.gf_last_error()
            assert lib.gf_site_frequency_spectrum_noalloc(d_out, 8, 0, L, d_out, stream) != 0
            assert lib.gf_site_frequency_spectrum_noalloc(None, 8, 5, L, d_out, stream) != 0
        finally:
            hip.lib.hipFree(d_out)
    finally:
        hip.lib.hipStreamDestroy(stream)
