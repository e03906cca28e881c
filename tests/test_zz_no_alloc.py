"""pngf_mlp_predict_dev, the entry point that allocates nothing, against pngf_mlp_predict on the
same network and inputs: bitwise the same outputs on two calls in a row, the second with fewer
rows (the scratch keeps the first call's contents), on the network's own stream and on a
caller's stream; sizes beyond the scratch are refused.  Device buffers come from the HIP runtime
directly (hipMalloc in the test, not in the library)."""
import ctypes as ct

import numpy as np
import pytest

import pngf_amd as P

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip():
    L = ct.CDLL("libamdhip64.so")
    L.hipMalloc.argtypes = [ct.POINTER(ct.c_void_p), ct.c_size_t]
    L.hipFree.argtypes = [ct.c_void_p]
    L.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
    L.hipDeviceSynchronize.argtypes = []
    L.hipStreamCreate.argtypes = [ct.POINTER(ct.c_void_p)]
    L.hipStreamSynchronize.argtypes = [ct.c_void_p]
    L.hipStreamDestroy.argtypes = [ct.c_void_p]
    return L


class DevBuf:
    def __init__(self, L, nbytes):
        self.L = L
        self.p = ct.c_void_p()
        assert L.hipMalloc(ct.byref(self.p), nbytes) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        assert a.nbytes <= self.nbytes
        assert self.L.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, shape):
        out = np.zeros(shape, dtype=np.float32)
        assert self.L.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.L.hipFree(self.p)


def test_predict_dev_matches_predict():
    L = hip()
    din, dh, dout = 123, 256, 9
    mlp = P.MLP(din, dh, dout, seed=4)
    rng = np.random.default_rng(3)
    mlp.set_params((rng.standard_normal(mlp.nparam) * 0.05).astype(np.float32))
    n1, n2 = 2048, 333  # 2 n1 dh = 2^20: the largest batch the scratch takes
    X1 = rng.standard_normal((n1, din)).astype(np.float32)
    X2 = rng.standard_normal((n2, din)).astype(np.float32)
    xb, ob = DevBuf(L, n1 * din * 4), DevBuf(L, n1 * dout * 4)
    stream = ct.c_void_p()
    assert L.hipStreamCreate(ct.byref(stream)) == 0
    try:
        lib = P.lib()
        for X, st in ((X1, None), (X2, stream)):
            n = len(X)
            want = mlp.predict(X)
            xb.put(X)
            assert L.hipDeviceSynchronize() == 0
            rc = lib.pngf_mlp_predict_dev(mlp.h, xb.p, n, ob.p, st)
            assert rc == 0, lib.pngf_mlp_last_error(mlp.h)
            assert L.hipDeviceSynchronize() == 0
            got = ob.get((n, dout))
            assert np.array_equal(got, want), n
        rc = lib.pngf_mlp_predict_dev(mlp.h, xb.p, n1 + 1, ob.p, None)
        assert rc == -1 and b"scratch" in lib.pngf_mlp_last_error(mlp.h)
        assert lib.pngf_mlp_predict_dev(mlp.h, xb.p, 0, ob.p, None) == 0
    finally:
        L.hipDeviceSynchronize()
        L.hipStreamDestroy(stream)
        xb.free()
        ob.free()
        mlp.close()
