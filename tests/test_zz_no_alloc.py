"""The allocation-free transform entry point (erp_spectral_transform_noalloc)
against the host-buffer entry points it shares kernels with
(erp_dct2_fft2_forward, erp_idct2_fft2_forward, erp_idct_idxst_forward,
erp_idxst_idct_forward): identical results, bit for bit, on two calls in a
row, the second on a smaller grid (the scratch keeps the first call's
contents), and sizes beyond the scratch refused.  Device buffers are made with
the HIP runtime here; the entry point itself allocates nothing."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class _Hip:
    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so")
        self.bufs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), C.c_size_t(nbytes)) == 0
        self.bufs.append(p)
        return p

    def put(self, p, a):
        assert self.lib.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), H2D) == 0

    def get(self, p, a):
        assert self.lib.hipMemcpy(a.ctypes.data_as(C.c_void_p), p, C.c_size_t(a.nbytes), D2H) == 0

    def sync(self):
        assert self.lib.hipDeviceSynchronize() == 0

    def free(self):
        for p in self.bufs:
            self.lib.hipFree(p)
        self.bufs = []


LEAF = ["erp_dct2_fft2_forward", "erp_idct2_fft2_forward", "erp_idct_idxst_forward",
        "erp_idxst_idct_forward"]


def _host_leaf(lib, which, x):
    m, n = x.shape
    out = np.zeros_like(x)
    rc = getattr(lib, LEAF[which])(x.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p),
                                   C.c_int32(m), C.c_int32(n), C.c_int32(0))
    assert rc == 0
    return out


def test_noalloc_matches_host_entry_points(gpu_lib):
    lib = gpu_lib.load()
    hip = _Hip()
    try:
        rng = np.random.default_rng(17)
        for m, n in ((512, 1024), (128, 256)):            # second call smaller
            print(f"noalloc transforms {m}x{n}", flush=True)
            x = rng.standard_normal((m, n)).astype(np.float32)
            dx, dout = hip.alloc(x.nbytes), hip.alloc(x.nbytes)
            hip.put(dx, x)
            for which in range(4):
                rc = lib.erp_spectral_transform_noalloc(dx, dout, C.c_int32(m), C.c_int32(n),
                                                        C.c_int32(which), None)
                assert rc == 0, lib.erp_last_error()
                hip.sync()
                got = np.zeros_like(x)
                hip.get(dout, got)
                ref = _host_leaf(lib, which, x)
                assert np.array_equal(got, ref), (m, n, which)
        # beyond the scratch: refused before any launch
        rc = lib.erp_spectral_transform_noalloc(dx, dout, C.c_int32(4096), C.c_int32(64),
                                                C.c_int32(0), None)
        assert rc == gpu_lib.ERP_ERR_UNSUPPORTED
        assert b"2048" in lib.erp_last_error()
        rc = lib.erp_spectral_transform_noalloc(dx, dx, C.c_int32(128), C.c_int32(256),
                                                C.c_int32(0), None)
        assert rc == gpu_lib.ERP_ERR_INVALID
    finally:
        hip.sync()
        hip.free()
