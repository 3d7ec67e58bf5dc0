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
