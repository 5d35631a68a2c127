"""pcc_forward_noalloc: the differentiable form's forward with its work space
in the library's file-scope device array instead of allocations.  Through the
C ABI: two calls in a row (the second with a smaller input) must equal
pcc_forward's output bitwise, on a partition with small and four-step levels;
an input beyond the array's 64 MiB is rejected with a clear error."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_noalloc_forward_matches_forward_twice():
    import torch
    from partconv_amd._lib import last_error, lib
    from partconv_amd.workloads import decaying_noise_ir
    L = lib()
    N = (ctypes.c_int * 3)(256, 2048, 16384)
    P = (ctypes.c_int * 3)(14, 14, 0)
    taps = 65536
    c = L.pcc_create(3, N, P, taps)
    assert c, last_error()
    try:
        h = torch.from_numpy(decaying_noise_ir(taps, 0.8, 48000, 3)).cuda()
        s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        assert L.pcc_set_filter(c, ctypes.c_void_p(h.data_ptr()), s) == 0, last_error()
        g = torch.Generator(device="cuda").manual_seed(7)
        for C, T in ((16, 131072), (3, 40000)):        # the second call smaller
            x = torch.randn((C, T), device="cuda", generator=g)
            y_ref = torch.empty_like(x)
            y = torch.full_like(x, float("nan"))
            assert L.pcc_forward(c, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y_ref.data_ptr()), C, T, s) == 0, \
                last_error()
            assert L.pcc_forward_noalloc(c, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()), C, T,
                                         s) == 0, last_error()
            torch.cuda.synchronize()
            print(f"C={C} T={T}: max |diff| {float((y - y_ref).abs().max()):.3e}", flush=True)
            assert torch.equal(y, y_ref), (C, T)
        # beyond the array: rejected before any GPU work
        xb = torch.zeros((64, 131072), device="cuda")
        rc = L.pcc_forward_noalloc(c, ctypes.c_void_p(xb.data_ptr()), ctypes.c_void_p(xb.data_ptr()), 64, 131072, s)
        assert rc != 0 and "maximum" in last_error()
    finally:
        L.pcc_destroy(c)
