"""Restart files in the reference's Fortran unformatted layout (press.f95:192-199,
adam.f95:24-29, ifdata.f95:227-247), pinned to files written by the Fortran runtime itself
(tests/golden/restart_records.f95 -> restart_data30.bin / restart_data31.bin)."""
import os

import numpy as np
import pytest

from les_halo_amd import fields as fl

HERE = os.path.dirname(os.path.abspath(__file__))
D30 = os.path.join(HERE, "golden", "restart_data30.bin")
D31 = os.path.join(HERE, "golden", "restart_data31.bin")
IM, JM, KM = 6, 4, 3
NAMES30 = ["u", "v", "w", "p", "usum", "vsum", "wsum"]
NAMES31 = ["fold", "gold", "hold"]


def _records(path):
    raw = open(path, "rb").read()
    out, pos = [], 0
    while pos < len(raw):
        n = int(np.frombuffer(raw, "<u4", 1, pos)[0])
        out.append(raw[pos + 4:pos + 4 + n])
        assert int(np.frombuffer(raw, "<u4", 1, pos + 4 + n)[0]) == n
        pos += n + 8
    return out


def _expected(q):
    i, j, k = np.meshgrid(np.arange(1, IM + 1), np.arange(1, JM + 1), np.arange(1, KM + 1),
                          indexing="ij")
    return (q * 1000 + i + 10 * j + 100 * k).astype(np.float32)  # (i, j, k)


def test_fixture_layout():
    """4-byte length markers around each record; (n, time) header; interior arrays i fastest."""
    r30, r31 = _records(D30), _records(D31)
    assert len(r30) == 8 and len(r31) == 4
    assert np.frombuffer(r30[0][:4], "<i4")[0] == 1234
    assert np.frombuffer(r30[0][4:], "<f4")[0] == np.float32(61.5)
    for q in range(7):
        a = np.frombuffer(r30[1 + q], "<f4").reshape(KM, JM, IM).transpose(2, 1, 0)
        assert np.array_equal(a, _expected(q + 1))
    assert not np.frombuffer(r31[3], "<f4").any()  # fghold


def _interior_put(name, full, vals):
    a = fl.FArray(name, IM, JM, KM, data=full)
    a[1:IM + 1, 1:JM + 1, 1:KM + 1] = vals
    return full


@pytest.mark.gpu
def test_write_matches_fortran_and_read_round_trips(tmp_path):
    from les_halo_amd import LES
    rng = np.random.default_rng(4)
    L = LES(IM, JM, KM)
    for q, nm in enumerate(NAMES30 + NAMES31):
        full = rng.standard_normal(fl.size(nm, IM, JM, KM)).astype(np.float32)  # halos: noise
        L.upload(nm, _interior_put(nm, full, _expected(q + 1)))
    p20, p21 = str(tmp_path / "data20"), str(tmp_path / "data21")
    L.write_restart(p20, p21, n=1234, time=61.5, with_header=True)
    assert open(p20, "rb").read() == open(D30, "rb").read()
    assert open(p21, "rb").read() == open(D31, "rb").read()
    L.write_restart(p20, p21)  # the data20 form: no header record
    assert open(p20, "rb").read() == open(D30, "rb").read()[16:]
    # read the Fortran-written pair into a fresh handle: interiors replaced, halos kept
    M = LES(IM, JM, KM)
    before = {nm: rng.standard_normal(fl.size(nm, IM, JM, KM)).astype(np.float32)
              for nm in NAMES30 + NAMES31}
    for nm, v in before.items():
        M.upload(nm, v)
    n, t = M.read_restart(D30, D31)
    assert (n, t) == (1234, np.float32(61.5))
    for q, nm in enumerate(NAMES30 + NAMES31):
        got = fl.FArray(nm, IM, JM, KM, data=M.download(nm))
        assert np.array_equal(got[1:IM + 1, 1:JM + 1, 1:KM + 1], _expected(q + 1)), nm
        want = _interior_put(nm, before[nm].copy(), _expected(q + 1))
        assert np.array_equal(M.download(nm), want), nm
