"""One domain as 2 x 2 (and 2 x 1, 1 x 2) tiles on one GPU (les_tiled) against one handle on
the whole domain: after the init sequence and three steps every interior value, and every
halo value on the domain's faces that the stages set, is bitwise the single-domain value.  The
tiles exchange their shared faces and the periodic j faces between stages, keep the domain's
inflow/outflow and boundary rules on the domain's faces, and combine the exact sums."""
import numpy as np
import pytest

import common

pytestmark = pytest.mark.gpu

SOR = dict(omega=1.5, nmaxp=20)
FIELDS = ["u", "v", "w", "p", "f", "g", "h", "usum", "vsum", "wsum", "fold", "gold", "hold",
          "sm", "rhs"] + ["diu%d" % q for q in range(1, 10)]


def _interior(name, a, dims):
    return common.interior(name, a, *dims)


# even tile sizes sweep the tiles' colour-split copies with split-layout face exchanges; an
# odd jp (32 x 21: the colour flips across the periodic seam) sweeps in place, as the single
# domain does there.  18 columns per tile give colour rows of 9 entries (a float4 group of the
# split rows holds the tile's last entry and padding); 3 x 2 tiles have middle tiles with
# neighbours on both i faces.
@pytest.mark.parametrize("tiles,grid", [((2, 2), (32, 24)), ((2, 1), (32, 24)),
                                        ((1, 2), (32, 24)), ((2, 1), (32, 21)),
                                        ((2, 2), (36, 24)), ((3, 2), (54, 24))])
def test_tiled_equals_single_domain(tiles, grid):
    from les_halo_amd import LES, TiledLES
    case = common.tiny_case(grid[0], grid[1], 16, seed=21)
    dims = case["dims"]
    L = LES(*dims, **common.REF_PHYS, pjuge=0.0, **SOR)
    T = TiledLES(*dims, sub_i=tiles[0], sub_j=tiles[1], **common.REF_PHYS, **SOR)
    for X in (L, T):
        X.load_case(case)
        X.ifdata()
    u0 = T.download("u")
    L.run(1, 3)
    T.step(1)  # eager, then graph replays of steps 2 and 3
    T.run(2, 2)
    for nm in FIELDS:
        a, b = _interior(nm, L.download(nm), dims), _interior(nm, T.download(nm), dims)
        assert common.bits_equal(b, a), common.mismatch_report(b, a, nm)
    # the domain's faces of u, v, w and p (inflow, outflow, periodic j, top/bottom)
    for nm in ("u", "v", "w", "p"):
        b, a = T.download(nm), L.download(nm)
        assert common.bits_equal(b, a), common.mismatch_report(b, a, nm + " (full)")
    # the steps moved the state (a vacuous comparison would pass too)
    assert not common.bits_equal(T.download("u"), u0)
    assert np.abs(_interior("p", T.download("p"), dims)).max() > 0
    T.close()
    L.close()


def test_tiled_rejects_odd_tile_sizes():
    from les_halo_amd import LESError, TiledLES
    with pytest.raises(LESError):
        TiledLES(30, 24, 8, sub_i=2, sub_j=2)  # 15 columns per tile: colours would flip
