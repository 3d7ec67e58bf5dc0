"""The thread-to-work-item maps of the colour-split kernels (SOR sweeps, conjugate-gradient
kernels), checked with exact integers on the CPU.

A thread index t becomes (group, row) and a row becomes (level chunk, j) by multiply-high
quotients: q = (t * m) >> 32 with m = floor((2^32 - 1) / n) + 1 (les_sor.hip: magic_div),
admitted by sor_split_divisors when ng * (jp * kp) * ng < 2^32 and jp * kp * jp < 2^32.  For
every grid the kernels run on here, and for the level chunks they use (one level, the
residual-summing sweep's four, the CG kernels' ten), each quotient must equal the true one for
every thread index of the launch: a single wrong quotient would skip or repeat a row of cells.
"""
import numpy as np
import pytest


def magic(n):
    return (0xFFFFFFFF // n) + 1


def admitted(ip, jp, kp):
    nr = ((ip + 1) // 2 + 3) // 4 * 4
    ng, rows = nr // 4, jp * kp
    return ng * rows * ng < 2 ** 32 and rows * jp < 2 ** 32


GRIDS = [(300, 300, 90), (600, 600, 90), (150, 150, 90), (150, 150, 60), (320, 320, 90),
         (96, 96, 60), (16, 12, 9), (37, 18, 23), (40, 32, 20), (14, 10, 8)]


@pytest.mark.parametrize("dims", GRIDS)
@pytest.mark.parametrize("kc", [1, 4, 10])
def test_multiply_high_quotients_are_exact(dims, kc):
    ip, jp, kp = dims
    assert admitted(ip, jp, kp)
    nr = ((ip + 1) // 2 + 3) // 4 * 4
    ng = nr // 4
    nch = (kp + kc - 1) // kc
    t = np.arange(ng * jp * nch, dtype=np.uint64)
    row = (t * np.uint64(magic(ng))) >> np.uint64(32)
    assert np.array_equal(row, t // np.uint64(ng))
    grp = t - row * np.uint64(ng)
    ch = (row * np.uint64(magic(jp))) >> np.uint64(32)
    assert np.array_equal(ch, row // np.uint64(jp))
    j = 1 + row - ch * np.uint64(jp)
    # every (group, j, chunk) once, so every (group, j, k) once over the chunk's levels
    key = (ch * np.uint64(jp) + (j - np.uint64(1))) * np.uint64(ng) + grp
    assert np.array_equal(np.sort(key), t)


def test_large_sizes_are_refused():
    # 2048 x 2048 x 256: ng^2 jp kp exceeds 2^32, so the split kernels are not used there
    assert not admitted(2048, 2048, 256)
