"""Ensemble of uneven members advanced together (les_ensemble_run): every member ends exactly
where its own les_run would leave it, and where the oracle does."""
import numpy as np
import pytest

import common
from les_halo_amd import inputs
from oracle.oracle import SUM_EXACT

pytestmark = pytest.mark.gpu

DIMS = [(24, 16, 10), (40, 18, 12), (16, 20, 8), (33, 12, 9)]


def _member(dims, seed, ustar):
    from les_halo_amd import LES
    case = inputs.make_case(*dims, seed=seed, ustar=ustar, hmax=12.0)
    L = LES(*dims, **common.REF_PHYS, **common.CFG_SOR)
    L.load_case(case)
    L.ifdata()
    return L, case


def test_ensemble_equals_members_and_oracle():
    from les_halo_amd import Ensemble
    members, alone, cases = [], [], []
    for q, dims in enumerate(DIMS):
        members.append(_member(dims, 100 + q, 0.1 + 0.1 * q)[0])
        L, case = _member(dims, 100 + q, 0.1 + 0.1 * q)
        alone.append(L)
        cases.append(case)
    E = Ensemble(members)
    E.run(1, 5)
    E.sync()
    for L in alone:
        L.run(1, 5)
        L.sync()
    for q, (M, L) in enumerate(zip(members, alone)):
        for nm in ("u", "v", "w", "p", "sm"):
            assert common.bits_equal(M.download(nm), L.download(nm)), (DIMS[q], nm)
    # one member against the oracle
    q = 1
    o = common.oracle_from_case(cases[q], **common.REF_PHYS, **common.CFG_SOR, sum_mode=SUM_EXACT)
    o.ifdata()
    for n in range(1, 6):
        o.step(n)
    for nm in ("u", "v", "w", "p"):
        assert common.bits_equal(members[q].download(nm), o.a[nm]), nm
    # continuing from a later step number
    E.run(6, 2)
    E.sync()
    for L in alone:
        L.run(6, 2)
        L.sync()
    for M, L in zip(members, alone):
        assert np.array_equal(M.download("u"), L.download("u"))
    E.close()


BATCH_DIMS = [(24, 16, 10), (40, 18, 12), (32, 20, 8), (34, 12, 9), (64, 32, 14)]


def _tuned_member(dims, seed, ustar, tile):
    L, case = _member(dims, seed, ustar)
    L.set_sor_tile(*tile)
    return L, case


def test_batched_ensemble_equals_members_and_oracle():
    """The batched form (velfg, calc_sm, visc+adam and every fused SOR iteration as one launch
    over all members, in one captured graph per step) leaves every member exactly where its own
    les_run leaves it, and where the oracle does; later runs continue it, reads of rhs and p
    between runs see the members' own values."""
    from les_halo_amd import Ensemble
    tiles = [(4, 5), (8, 10), (4, 3), (16, 4), (4, 7)]
    members, alone, cases = [], [], []
    for q, dims in enumerate(BATCH_DIMS):
        members.append(_tuned_member(dims, 200 + q, 0.15 + 0.07 * q, tiles[q])[0])
        L, case = _tuned_member(dims, 200 + q, 0.15 + 0.07 * q, tiles[q])
        alone.append(L)
        cases.append(case)
    E = Ensemble(members, batched=True)
    E.run(1, 4)
    E.sync()
    assert E.batched
    for L in alone:
        L.run(1, 4)
        L.sync()
    names = ("u", "v", "w", "p", "sm", "rhs", "f", "usum", "fold", "diu3")
    for q, (M, L) in enumerate(zip(members, alone)):
        for nm in names:
            assert common.bits_equal(M.download(nm), L.download(nm)), (BATCH_DIMS[q], nm)
    q = 4
    o = common.oracle_from_case(cases[q], **common.REF_PHYS, **common.CFG_SOR, sum_mode=SUM_EXACT)
    o.ifdata()
    for n in range(1, 5):
        o.step(n)
    for nm in ("u", "v", "w", "p", "rhs"):
        assert common.bits_equal(members[q].download(nm), o.a[nm]), nm
    # continuing, after the reads above; then the concurrent form on the same members
    E.run(5, 3)
    E.sync()
    for L in alone:
        L.run(5, 3)
    for M, L in zip(members, alone):
        for nm in ("u", "p", "rhs"):
            assert common.bits_equal(M.download(nm), L.download(nm)), nm
    E.close()
    E2 = Ensemble(members, batched=False)
    E2.run(8, 2)
    E2.sync()
    assert not E2.batched
    for L in alone:
        L.run(8, 2)
    for M, L in zip(members, alone):
        assert common.bits_equal(M.download("p"), L.download("p"))
    E2.close()
