"""les_reduce_sum_exact_noalloc (no handle, no allocation: partials in the library's fixed
device array) against les_reduce_sum_exact on a handle: the same bits, on two calls in a row,
the second over fewer values (fewer workgroups than the first left partials for)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_noalloc_sum_equals_handle_sum():
    import torch
    from les_halo_amd import LES
    from les_halo_amd._lib import LES_E_INVALID, lib
    L = LES(16, 12, 10)
    rng = np.random.default_rng(5)
    out = torch.full((1,), 7.0, dtype=torch.float32, device="cuda:0")
    stream = torch.cuda.current_stream().cuda_stream
    for n in (3_000_001, 4099):
        # magnitudes over twelve decades and both signs: a sum that rounding order would change
        x = (rng.standard_normal(n) * 10.0 ** rng.uniform(-6, 6, n)).astype(np.float32)
        xd = torch.from_numpy(x).to("cuda:0")
        torch.cuda.synchronize()
        want = L.reduce_sum_exact(xd.data_ptr(), n)
        r = lib().les_reduce_sum_exact_noalloc(ctypes.c_void_p(xd.data_ptr()), n,
                                               ctypes.c_void_p(out.data_ptr()),
                                               ctypes.c_void_p(stream))
        assert r == 0
        torch.cuda.synchronize()
        got = out.cpu().numpy()[0]
        assert np.float32(got).view(np.uint32) == np.float32(want).view(np.uint32), (n, got, want)
    # sizes beyond the maximum and null pointers are rejected before any launch
    big = (1 << 31)
    assert lib().les_reduce_sum_exact_noalloc(ctypes.c_void_p(xd.data_ptr()), big,
                                              ctypes.c_void_p(out.data_ptr()), None) == LES_E_INVALID
    assert lib().les_reduce_sum_exact_noalloc(None, 10, ctypes.c_void_p(out.data_ptr()),
                                              None) == LES_E_INVALID
    L.close()
