"""bw_mlp_forward_noalloc (no allocation: activations in the library's fixed
device array) against bw_mlp_forward on the same network and features, bit
for bit, on two calls in a row (the second smaller), through the C-ABI.  The
caller's device buffers come from the HIP runtime the library itself links
(one runtime in the process)."""
import ctypes as C
import os

import numpy as np
import pytest

import brownian_amd as B

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime libbrownian links (already loaded with it)."""
    path = next((l.split()[-1] for l in open("/proc/self/maps") if "libamdhip64.so" in l and "/torch/" not in l),
                "/opt/rocm/lib/libamdhip64.so")
    h = C.CDLL(path)
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2


def test_forward_noalloc_equals_forward():
    lib = B.lib()
    net = B.MLP(seed=3)
    hip = _hip()
    rng = np.random.default_rng(4)
    x_mean = rng.normal(0, 0.2, 144).astype(np.float32)
    x_std = rng.uniform(0.5, 2.0, 144).astype(np.float32)
    net.set_normalisation(x_mean, x_std, np.zeros(6, np.float32), np.ones(6, np.float32))
    fn = lib.bw_mlp_forward_noalloc
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
    dx, dy = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(dx), 1025 * 144 * 4) == 0
    assert hip.hipMalloc(C.byref(dy), 1025 * 12 * 4) == 0
    try:
        for n in (700, 45):
            x = rng.normal(0, 1, (n, 144)).astype(np.float32)
            ref = net.forward(x)
            assert hip.hipMemcpy(dx, x.ctypes.data, x.nbytes, H2D) == 0
            y = np.full((n, 12), np.nan, np.float32)
            assert hip.hipMemcpy(dy, y.ctypes.data, y.nbytes, H2D) == 0
            rc = fn(net.h, dx, n, dy)
            assert rc == 0, lib.bw_mlp_last_error(net.h)
            assert hip.hipMemcpy(y.ctypes.data, dy, y.nbytes, D2H) == 0
            np.testing.assert_array_equal(y, ref, err_msg=f"n={n}")
        # beyond the scratch array: a clear error, nothing launched
        assert fn(net.h, dx, 1025, dy) == -1  # BW_ERR_ARG
        assert b"1024" in lib.bw_mlp_last_error(net.h)
    finally:
        hip.hipDeviceSynchronize()
        hip.hipFree(dx)
        hip.hipFree(dy)
