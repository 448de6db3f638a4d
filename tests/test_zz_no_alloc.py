"""emt6ro_mlp_predict_noalloc (the allocation-free forward pass) against
emt6ro_mlp_predict, through the C ABI: the same outputs bit for bit on two
calls in a row, the second with a smaller input, and the size limit.

The test's own device buffers come from the HIP runtime the library is
linked against (libamdhip64, through ctypes); the entry point under test
allocates nothing itself."""
import ctypes as C

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs an MI355X GPU")]


def _hip():
    from emt6ro import _native  # noqa: F401  (loads the library and its HIP runtime)
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


H2D, D2H = 1, 2


class DeviceArray:
    def __init__(self, hip, host):
        self.hip, self.nbytes = hip, host.nbytes
        self.ptr = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), max(self.nbytes, 4)) == 0
        if host.nbytes:
            assert hip.hipMemcpy(self.ptr, host.ctypes.data, host.nbytes, H2D) == 0

    def get(self, like):
        out = np.empty_like(like)
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def _net(in_features, hidden, seed):
    from emt6ro.surrogate import MLP
    net = MLP(in_features, hidden, seed=seed)
    rng = np.random.default_rng(seed)
    p = net.get_params()
    net.set_params(p + rng.normal(0, 0.02, p.size).astype(np.float32))
    return net


def _noalloc(hip, net, X):
    from emt6ro._native import lib, check_mlp
    dX = DeviceArray(hip, X)
    y = np.full(len(X), np.nan, np.float32)
    dy = DeviceArray(hip, y)
    try:
        check_mlp(lib.emt6ro_mlp_predict_noalloc(net._h, dX.ptr, len(X), dy.ptr))
        return dy.get(y)
    finally:
        dX.free()
        dy.free()


@pytest.mark.parametrize("in_features,n1,n2", [(71, 9000, 1000), (152, 4096, 37)])
def test_noalloc_predict_matches_predict(in_features, n1, n2):
    hip = _hip()
    net = _net(in_features, 512, seed=in_features)
    rng = np.random.default_rng(1)
    for n in (n1, n2):
        X = rng.normal(0, 1, (n, in_features)).astype(np.float32)
        want = net.predict(X)
        got = _noalloc(hip, net, X)
        assert np.isfinite(got).all()
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), n


def test_noalloc_rejects_networks_over_the_limit():
    from emt6ro._native import lib
    hip = _hip()
    net = _net(600, 64, seed=3)
    X = np.zeros((8, 600), np.float32)
    dX = DeviceArray(hip, X)
    dy = DeviceArray(hip, np.zeros(8, np.float32))
    try:
        rc = lib.emt6ro_mlp_predict_noalloc(net._h, dX.ptr, 8, dy.ptr)
        assert rc != 0
        assert b"at most 512" in lib.emt6ro_mlp_last_error()
    finally:
        dX.free()
        dy.free()


def test_noalloc_zero_rows_is_a_no_op():
    from emt6ro._native import lib
    net = _net(71, 64, seed=4)
    assert lib.emt6ro_mlp_predict_noalloc(net._h, None, 0, None) == 0
