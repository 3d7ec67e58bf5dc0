"""GIS terrain loader and on-device mask setup (feedbfm.f95:31-63), against the reference's own
GIS file and the masks its Fortran produced from it (tests/golden/tokyo_150x150x90.npz)."""
import os

import numpy as np
import pytest

import common
from les_halo_amd import fields as fl
from les_halo_amd import inputs

HERE = os.path.dirname(os.path.abspath(__file__))
TOKYO = os.path.join(HERE, "golden", "tokyo_150x150x90.npz")
GIS = "/root/reference/src/GIS/Tokyo_20mgrid.txt"
IP = JP = 150
KP = 90
MASKS = ["amask1", "bmask1", "cmask1", "dmask1"]


def _tokyo():
    d = np.load(TOKYO)
    masks = {n: np.unpackbits(d[n])[:fl.size(n, IP, JP, KP)].astype(np.float32) for n in MASKS}
    return d, masks


def test_gis_reader_layout(tmp_path):
    """Rows j = 100..1, i = 1..100 per row, first value of each line, at (i+25, j+25)."""
    path = tmp_path / "gis.txt"
    vals = [(i, j) for j in range(100, 0, -1) for i in range(1, 101)]
    path.write_text("".join("  %d.5 7 junk\n" % (1000 * j + i) for i, j in vals))
    zbm = inputs.gis_zbm(str(path), 130, 127).reshape(127 + 3, 130 + 3)  # [j+1, i+1]
    for i, j in [(1, 1), (100, 1), (1, 100), (37, 64), (100, 100)]:
        assert zbm[j + 25 + 1, i + 25 + 1] == np.float32(1000 * j + i + 0.5)
    assert zbm[1 + 25 + 1, :26 + 1].sum() == 0 and zbm[:26 + 1, :].sum() == 0
    with pytest.raises(ValueError):
        inputs.gis_zbm(str(path), 124, 150)


@pytest.mark.skipif(not os.path.exists(GIS), reason="reference GIS file not present")
def test_gis_reader_matches_reference_read():
    d, _ = _tokyo()
    assert common.bits_equal(inputs.gis_zbm(GIS, IP, JP), d["zbm"])


@pytest.mark.gpu
def test_device_masks_match_reference_feedbfm():
    from les_halo_amd import LES
    d, ref = _tokyo()
    case = inputs.make_case(IP, JP, KP)
    case["z2"], case["dzn"] = d["z2"], d["dzn"]
    A = LES(IP, JP, KP, **common.REF_PHYS, **common.CFG_SOR)
    A.load_case(case)
    A.set_terrain(d["zbm"])
    for n in MASKS:
        assert common.bits_equal(A.download(n), ref[n]), n
    # the stencils' bit mask equals the one built from uploaded masks: one step agrees
    B = LES(IP, JP, KP, **common.REF_PHYS, **common.CFG_SOR)
    B.load_case(case)
    for n in MASKS:
        B.upload(n, ref[n])
    for L in (A, B):
        L.ifdata()
        L.step(1)
    for n in ("u", "v", "w", "p", "usum"):
        assert common.bits_equal(A.download(n), B.download(n)), n
