"""bw_mlp_forward_noalloc (no allocation: activations in the library's fixed
device array) against bw_mlp_forward on the same network and features, bit
for bit, on two calls in a row (the second smaller), through the C-ABI;
device buffers for the inputs and outputs come from torch."""
import ctypes as C

import numpy as np
import pytest

import brownian_amd as B

pytestmark = pytest.mark.gpu


def test_forward_noalloc_equals_forward():
    import torch
    lib = B.lib()
    net = B.MLP(seed=3)
    rng = np.random.default_rng(4)
    x_mean = rng.normal(0, 0.2, 144).astype(np.float32)
    x_std = rng.uniform(0.5, 2.0, 144).astype(np.float32)
    net.set_normalisation(x_mean, x_std, np.zeros(6, np.float32), np.ones(6, np.float32))
    fn = lib.bw_mlp_forward_noalloc
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
    for n in (700, 45):
        x = rng.normal(0, 1, (n, 144)).astype(np.float32)
        ref = net.forward(x)
        xd = torch.from_numpy(x).to("cuda")
        yd = torch.full((n, 12), float("nan"), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        rc = fn(net.h, xd.data_ptr(), n, yd.data_ptr())
        assert rc == 0, lib.bw_mlp_last_error(net.h)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(yd.cpu().numpy(), ref, err_msg=f"n={n}")
    # beyond the scratch array: a clear error, nothing launched
    xd = torch.zeros((1025, 144), dtype=torch.float32, device="cuda")
    yd = torch.zeros((1025, 12), dtype=torch.float32, device="cuda")
    assert fn(net.h, xd.data_ptr(), 1025, yd.data_ptr()) == -1  # BW_ERR_ARG
    assert b"1024" in lib.bw_mlp_last_error(net.h)
