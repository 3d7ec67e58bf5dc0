"""Halo-exchange pack/unpack (halo_exchange.cl, halo_exchange_correct.cl).

The oracle (oracle/halo.py) is pinned to the output of the reference's own C test programs
(src/OpenCL/test_halo_{read,write}{,_correct}.c, run by tests/golden/make_halo_vectors.py into
tests/golden/halo_vectors.json) and to the identity read(write(b)) == b.  The HIP kernels are
compared with the same fixtures and with the oracle on arrays of distinct values, bit for bit.
"""
import json
import os

import numpy as np
import pytest

from oracle import halo as oh

HERE = os.path.dirname(os.path.abspath(__file__))
VEC = json.load(open(os.path.join(HERE, "golden", "halo_vectors.json")))


def _fixture_inputs(name):
    """The inputs each reference test program's main() builds (see make_halo_vectors.py)."""
    e = VEC[name]
    im, jm, km, hw, vd = e["im"], e["jm"], e["km"], e["h_w"], e["v_dim"]
    n = im * jm * km
    if name == "test_halo_read_correct":
        arrays = [np.arange(v * n, (v + 1) * n, dtype=np.float32) for v in range(vd)]
        buf = np.full(oh.halo_size(vd, im, jm, km, hw, 1), -1.0, np.float32)
    else:
        arrays = [np.full(n, 1.0 + v, np.float32) for v in range(vd)]  # first half 1, second 2
        if name == "test_halo_read":
            v_sz = oh.halo_size(1, im, jm, km, hw)
            ns, we = im * hw * km, (jm - 2 * hw) * hw * km
            seg = np.concatenate([np.full(ns, 6.0), np.full(ns, 7.0), np.full(we, 8.0),
                                  np.full(we, 9.0)]).astype(np.float32)
            assert seg.size == v_sz
            buf = np.tile(seg, vd)
        else:
            buf = np.full(oh.halo_size(vd, im, jm, km, hw), 5.0, np.float32)
    return e, arrays, buf


def _oracle_program(name):
    """What the reference program computes, through the oracle."""
    e, arrays, buf = _fixture_inputs(name)
    im, jm, km, hw = e["im"], e["jm"], e["km"], e["h_w"]
    if name.startswith("test_halo_write"):
        return {"array": np.concatenate(oh.halo_write(arrays, buf, im, jm, km, hw))}
    if name == "test_halo_read":
        # test_halo_read.c's own unpack indexes the west block with (i-h_w)+w instead of
        # (i-h_w)*h_w+w; the block is uniform (8.0) there, so the result is the same
        arrays = oh.halo_write(arrays, buf, im, jm, km, hw)
        buf = np.full_like(buf, 4.0)
        return {"buffer": oh.halo_read(arrays, im, jm, km, hw, 0, buf)}
    return {"array": np.concatenate(arrays), "buffer": oh.halo_read(arrays, im, jm, km, hw, 1, buf)}


@pytest.mark.parametrize("name", sorted(VEC))
def test_oracle_matches_reference_programs(name):
    got = _oracle_program(name)
    for what, ref in got.items():
        assert np.array_equal(ref, np.float32(VEC[name][what])), (name, what)


@pytest.mark.parametrize("im,jm,km,hw", [(7, 5, 3, 1), (20, 20, 2, 3), (33, 17, 4, 2)])
def test_oracle_read_inverts_write(im, jm, km, hw):
    rng = np.random.default_rng(im * jm)
    vd = 3
    buf = rng.standard_normal(oh.halo_size(vd, im, jm, km, hw)).astype(np.float32)
    arrays = [rng.standard_normal(im * jm * km).astype(np.float32) for _ in range(vd)]
    out = oh.halo_write(arrays, buf, im, jm, km, hw)
    assert np.array_equal(oh.halo_read(out, im, jm, km, hw), buf)
    # interior untouched, ring fully overwritten
    ring = np.zeros((km, jm, im), bool)
    ring[:, :hw, :] = ring[:, -hw:, :] = ring[:, :, :hw] = ring[:, :, -hw:] = True
    for v in range(vd):
        assert np.array_equal(out[v][~ring.reshape(-1)], arrays[v][~ring.reshape(-1)])


def test_oracle_correct_read_is_inner_ring_without_corners():
    im, jm, km, hw = 9, 7, 2, 1
    a = np.arange(im * jm * km, dtype=np.float32)
    buf = oh.halo_read([a], im, jm, km, hw, 1)
    A = a.reshape(km, jm, im)
    north = A[:, hw, hw:im - hw].reshape(-1)
    south = A[:, jm - 2 * hw, hw:im - hw].reshape(-1)
    west = A[:, hw:jm - hw, hw].reshape(-1)
    east = A[:, hw:jm - hw, im - 2 * hw].reshape(-1)
    assert np.array_equal(buf, np.concatenate([north, south, west, east]))


# ---------------------------------------------------------------------------------------------
# HIP kernels
# ---------------------------------------------------------------------------------------------
def _dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x, np.float32)).to("cuda:0")


def _gpu_read(L, arrays, im, jm, km, hw, correct, buf):
    import torch
    da = [_dev(a) for a in arrays]
    db = _dev(buf)
    torch.cuda.synchronize()
    L.halo_read_arrays([t.data_ptr() for t in da], im, jm, km, db.data_ptr(), hw, correct)
    L.sync()
    return db.cpu().numpy()


def _gpu_write(L, arrays, im, jm, km, hw, buf):
    import torch
    da = [_dev(a) for a in arrays]
    db = _dev(buf)
    torch.cuda.synchronize()
    L.halo_write_arrays([t.data_ptr() for t in da], im, jm, km, db.data_ptr(), hw)
    L.sync()
    return [t.cpu().numpy() for t in da]


@pytest.fixture(scope="module")
def handle():
    from les_halo_amd import LES
    L = LES(8, 6, 4)
    yield L
    L.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(VEC))
def test_gpu_matches_reference_programs(handle, name):
    e, arrays, buf = _fixture_inputs(name)
    im, jm, km, hw = e["im"], e["jm"], e["km"], e["h_w"]
    if name.startswith("test_halo_write"):
        got = {"array": np.concatenate(_gpu_write(handle, arrays, im, jm, km, hw, buf))}
    elif name == "test_halo_read":
        arrays = _gpu_write(handle, arrays, im, jm, km, hw, buf)
        got = {"buffer": _gpu_read(handle, arrays, im, jm, km, hw, 0, np.full_like(buf, 4.0))}
    else:
        got = {"buffer": _gpu_read(handle, arrays, im, jm, km, hw, 1, buf)}
    for what, val in got.items():
        assert np.array_equal(val, np.float32(VEC[name][what])), (name, what)


@pytest.mark.gpu
@pytest.mark.parametrize("vd", [1, 2, 4, 16])
@pytest.mark.parametrize("im,jm,km,hw", [(303, 303, 92, 1), (302, 304, 93, 1), (37, 29, 5, 3)])
def test_gpu_pack_unpack_vs_oracle(handle, vd, im, jm, km, hw):
    if vd == 16 and im > 100:
        km = 7  # keep the 16-component case small
    rng = np.random.default_rng(vd * 1000 + im)
    n = im * jm * km
    arrays = [(np.arange(n, dtype=np.float32) + np.float32(v * 0.5)) for v in range(vd)]
    for correct in (0, 1):
        buf0 = np.full(oh.halo_size(vd, im, jm, km, hw, correct), -7.0, np.float32)
        got = _gpu_read(handle, arrays, im, jm, km, hw, correct, buf0)
        want = oh.halo_read(arrays, im, jm, km, hw, correct, buf0)
        assert np.array_equal(got, want), ("read", correct)
    buf = rng.standard_normal(oh.halo_size(vd, im, jm, km, hw)).astype(np.float32)
    got = _gpu_write(handle, arrays, im, jm, km, hw, buf)
    want = oh.halo_write(arrays, buf, im, jm, km, hw)
    for v in range(vd):
        assert np.array_equal(got[v], want[v]), ("write", v)


@pytest.mark.gpu
def test_gpu_exchange_on_handle_fields():
    """les_exchange_halo_read/write on the handle's own p and u, v, w (extents of
    main_ocl.f95:40-167, h_w = 1), against the oracle on the downloaded arrays."""
    import torch
    from les_halo_amd import LES
    from les_halo_amd import fields as fl
    ip, jp, kp = 20, 14, 9
    with LES(ip, jp, kp) as L:
        rng = np.random.default_rng(3)
        for nm in ("p", "u", "v", "w"):
            L.upload(nm, rng.standard_normal(fl.size(nm, ip, jp, kp)).astype(np.float32))
        for names, ext in ((["p"], (ip + 3, jp + 3, kp + 2)), (["u", "v"], (ip + 2, jp + 3, kp + 2))):
            host = [L.download(n) for n in names]
            for correct in (0, 1):
                nb = L.halo_size(names, 1, correct)
                assert nb == oh.halo_size(len(names), *ext, 1, correct)
                db = torch.zeros(nb, dtype=torch.float32, device="cuda:0")
                torch.cuda.synchronize()
                L.exchange_halo_read(names, db.data_ptr(), 1, correct)
                L.sync()
                assert np.array_equal(db.cpu().numpy(), oh.halo_read(host, *ext, 1, correct))
            nb = L.halo_size(names, 1, 0)
            src = rng.standard_normal(nb).astype(np.float32)
            db = _dev(src)
            torch.cuda.synchronize()
            L.exchange_halo_write(names, db.data_ptr(), 1)
            L.sync()
            want = oh.halo_write(host, src, *ext, 1)
            for q, n in enumerate(names):
                assert np.array_equal(L.download(n), want[q]), n
