"""The C ABI library on the CPU: it loads, exports every entry point include/les_hip.h
declares, and its host-only parts (field bounds, halo sizes, input generators) behave."""
import ctypes
import os
import re

import numpy as np
import pytest

from les_halo_amd import fields as fl
from les_halo_amd import inputs
from les_halo_amd._lib import LIB_PATH, PROTOS, lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "les_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(les_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(LIB_PATH)
    names = declared_functions()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the Python binding covers all of them
    assert not [n for n in names if n not in PROTOS]


def test_field_bounds_match_fortran_declarations():
    ip, jp, kp = 17, 13, 9
    lo = (ctypes.c_int * 3)()
    hi = (ctypes.c_int * 3)()
    for name, (fid, _) in fl.FIELDS.items():
        n = lib().les_field_bounds(fid, ip, jp, kp, lo, hi)
        b = fl.bounds(name, ip, jp, kp)
        assert n == fl.size(name, ip, jp, kp), name
        for q, (l, h) in enumerate(b):
            assert (lo[q], hi[q]) == (l, h), name
    assert lib().les_field_bounds(fl.NFIELDS, ip, jp, kp, lo, hi) == -1


def test_halo_buffer_sizes_match_reference_module():
    """s_p, s_uvw, ... of module_LES_combined_ocl (Wrappers/...:80-87) for h_w = 1: the byte
    counts there are for float2/float4/float16 vectors, i.e. v_dim values per point."""
    ip, jp, kp, hw = 300, 300, 90, 1
    L = lib()
    assert L.les_halo_size(2, ip + 3, jp + 3, kp + 2, hw, 0) == 4 * (ip + jp + 6 - 2 * hw) * hw * (kp + 2)
    assert L.les_halo_size(4, ip + 2, jp + 3, kp + 3, hw, 0) == 8 * (ip + jp + 5 - 2 * hw) * hw * (kp + 3)
    assert L.les_halo_size(4, ip + 1, jp + 1, kp + 1, hw, 0) == 8 * (ip + jp + 2 - 2 * hw) * hw * (kp + 1)
    assert L.les_halo_size(4, ip, jp, kp, hw, 0) == 8 * (ip + jp - 2 * hw) * hw * kp
    assert L.les_halo_size(16, ip + 4, jp + 3, kp + 3, hw, 0) == 32 * (ip + jp + 7 - 2 * hw) * hw * (kp + 3)
    assert L.les_halo_size(1, ip + 2, jp + 2, kp + 2, hw, 0) == 2 * (ip + jp + 4 - 2 * hw) * hw * (kp + 2)
    # halo_exchange_correct.cl:164: v_sz = 2*km*(im+jm-4*h_w)*h_w
    assert L.les_halo_size(1, 10, 8, 5, 1, 1) == 2 * 5 * (10 + 8 - 4)


def test_grid_generator():
    g = inputs.gen_grid(20, 16, 40)
    dzn = fl.FArray("dzn", 20, 16, 40, data=g["dzn"])
    z2 = g["z2"]
    # 2 m cells while the top is <= 40 m, then x1.05 per level (FP32, sequential)
    assert all(dzn[k] == np.float32(2.0) for k in range(1, 21))
    for k in range(21, 41):
        assert dzn[k] == np.float32(dzn[k - 1] * np.float32(1.05))
    assert z2[19] == np.float32(40.0)
    # halos of grid.f95:60-65 and the staggered spacings of grid.f95:67-82
    assert dzn[0] == dzn[1] and dzn[-1] == dzn[0] and dzn[41] == dzn[40] and dzn[42] == dzn[41]
    dzs = fl.FArray("dzs", 20, 16, 40, data=g["dzs"])
    for k in range(0, 41):
        assert dzs[k] == np.float32(dzn[k + 1] / np.float32(2) + dzn[k] / np.float32(2))
    assert np.all(g["dx1"] == 4.0) and np.all(g["dxs"] == 4.0) and np.all(g["dys"] == 4.0)


def test_terrain_generator_coverage_and_determinism():
    ip, jp = 120, 100
    a = inputs.gen_terrain(ip, jp, seed=42)
    b = inputs.gen_terrain(ip, jp, seed=42)
    c = inputs.gen_terrain(ip, jp, seed=43)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    z = a.reshape(jp + 3, ip + 3)[2:jp + 2, 2:ip + 2]  # zbm(-1:ip+1,-1:jp+1) as (j, i)
    cover = np.count_nonzero(z) / (ip * jp)
    assert 0.30 <= cover < 0.40
    assert z.max() <= 60.0 and z.min() >= 0.0
    # halo ring of zbm(-1:ip+1,-1:jp+1) stays zero
    full = a.reshape(jp + 3, ip + 3)
    assert not full[:2].any() and not full[:, :2].any() and not full[-1].any() and not full[:, -1].any()


def test_mask_rule():
    """feedbfm.f95:42-63: solid where zbm > z2(k) + 0.5*dzn(k); b = c = d = 1 - a inside."""
    ip, jp, kp = 24, 20, 30
    case = inputs.make_case(ip, jp, kp, seed=5)
    z2 = case["z2"]
    dzn = fl.FArray("dzn", ip, jp, kp, data=case["dzn"])
    zbm = case["zbm"].reshape(jp + 3, ip + 3)
    a = fl.FArray("amask1", ip, jp, kp, data=case["amask1"])
    b = fl.FArray("bmask1", ip, jp, kp, data=case["bmask1"])
    c = fl.FArray("cmask1", ip, jp, kp, data=case["cmask1"])
    d = fl.FArray("dmask1", ip, jp, kp, data=case["dmask1"])
    for k in range(1, kp + 1):
        zk = np.float32(z2[k - 1] + np.float32(0.5) * dzn[k])
        solid = zbm[2:jp + 2, 2:ip + 2].T > zk
        assert np.array_equal(a[1:ip + 1, 1:jp + 1, k] == 0, solid)
        for m in (b, c, d):
            assert np.array_equal(m[1:ip + 1, 1:jp + 1, k] == 1, solid)
    assert case["amask1"].sum() < ip * jp * kp


def test_inflow_generator():
    ip, jp, kp = 10, 50, 40
    g = inputs.gen_grid(ip, jp, kp)
    w = inputs.gen_inflow(jp, kp, g, ustar=0.2, noise=0.01, seed=9).reshape(kp, jp)
    dzn = fl.FArray("dzn", ip, jp, kp, data=g["dzn"])
    z = np.array([np.float32(g["z2"][k - 1] + np.float32(0.5) * dzn[k]) for k in range(1, kp + 1)])
    mean = (0.2 / 0.4) * np.log(z.astype(np.float64) / 0.1)
    rel = w / mean[:, None] - 1.0
    assert abs(rel.mean()) < 0.003 and 0.007 < rel.std() < 0.013
    w0 = inputs.gen_inflow(jp, kp, g, ustar=0.2, noise=0.0, seed=9).reshape(kp, jp)
    assert np.allclose(w0, mean[:, None], rtol=1e-6)


def test_reference_power_law_profile():
    """initial_wind_profile.f95:22-32 for kp = 90."""
    ip, jp, kp = 8, 6, 90
    g = inputs.gen_grid(ip, jp, kp)
    w = np.zeros(jp * kp, np.float32)
    from les_halo_amd._lib import ptr
    assert lib().les_initial_wind_profile(jp, kp, ptr(g["z2"]), ptr(g["dzn"]), ptr(w)) == 0
    w = w.reshape(kp, jp)
    dzn = fl.FArray("dzn", ip, jp, kp, data=g["dzn"])
    for k in (1, 40, 78):
        zk = np.float32(g["z2"][k - 1] + np.float32(0.5) * dzn[k])
        assert abs(w[k - 1, 0] - 5.0 * (zk / 600.0) ** 0.2) < 1e-5
    assert np.all(w[78:] == w[76])


def test_create_without_gpu_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from les_halo_amd import LES, LESError
    with pytest.raises(LESError) as e:
        LES(8, 8, 8)
    assert e.value.code == 5  # LES_E_NODEV


def test_adjoint_create_without_gpu_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from les_halo_amd import LESError
    from les_halo_amd.adjoint import AdjointLES
    with pytest.raises(LESError) as e:
        AdjointLES(8, 8, 8)
    assert e.value.code == 5  # LES_E_NODEV


def test_invalid_sizes_rejected():
    from les_halo_amd._lib import LesConfig
    cfg = LesConfig()
    lib().les_default_config(ctypes.byref(cfg), 1, 8, 8)
    h = ctypes.c_void_p()
    assert lib().les_create(ctypes.byref(cfg), ctypes.byref(h)) == 1


def test_size_limit_of_32_bit_indexing():
    """les_create admits a grid only if its largest array, diu1(-1:ip+2, 0:jp+2, 0:kp+2), has
    fewer than 2^30 elements (device indices and buffer byte offsets are 32-bit): 2048 x 2048 x
    256 has (2052)(2051)(259) >= 2^30 and is rejected as invalid before any device work."""
    from les_halo_amd._lib import LesConfig
    assert 2052 * 2051 * 259 >= 1 << 30
    cfg = LesConfig()
    lib().les_default_config(ctypes.byref(cfg), 2048, 2048, 256)
    h = ctypes.c_void_p()
    assert lib().les_create(ctypes.byref(cfg), ctypes.byref(h)) == 1  # LES_E_INVALID
