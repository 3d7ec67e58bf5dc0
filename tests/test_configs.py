"""The HIP path on the inputs of the benchmark configurations, against the oracle bit for bit
(exact-sum mode, the GPU's reduction semantics; the oracle's sequential-sum mode is pinned to
the reference build at the same settings by test_oracle_reference.py):

* 150 x 150 x 90, the CPU configuration's grid, terrain and inflow, omega 1.5, exactly 50 SOR
  iterations: the init sequence and five steps, every state array;
* 300 x 300 x 90 as 2 x 2 tiles on one GPU (les_tiled, face exchanges between the tiles):
  the init sequence and two steps, compared with the oracle on the whole domain;
* members of the 64-member ensemble (sizes, terrain seeds and u* from inputs.ensemble_specs,
  kp 60 and 90, ip and jp up to 320) advanced together: each one after two steps.
"""
import pytest

import common
from les_halo_amd import inputs
from oracle.oracle import SUM_EXACT

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

STATE = list(common.DUMP_ORDER)
TILE_FIELDS = ["u", "v", "w", "p", "f", "g", "h", "usum", "vsum", "wsum", "fold", "gold", "hold",
               "sm", "rhs"] + ["diu%d" % q for q in range(1, 10)]


def _bad(get, o, names, tag, interior_of=None):
    bad = []
    for n in names:
        g, want = get(n), o.a[n]
        if interior_of:
            g, want = interior_of(n, g), interior_of(n, want)
        if not common.bits_equal(g, want):
            bad.append(common.mismatch_report(g, want, n))
    assert not bad, "%s:\n  %s" % (tag, "\n  ".join(bad))


def test_cpu_configuration_grid_150():
    from les_halo_amd import LES
    ip, jp, kp = 150, 150, 90
    case = inputs.make_case(ip, jp, kp)
    L = LES(ip, jp, kp, **inputs.CONFIG_PHYSICS)
    L.load_case(case)
    L.autotune(2)
    o = common.oracle_from_case(case, **inputs.CONFIG_PHYSICS, sum_mode=SUM_EXACT)
    L.ifdata()
    o.ifdata()
    L.run(1, 5)
    for n in range(1, 6):
        o.step(n)
    _bad(L.download, o, STATE, "150x150x90, 5 steps")
    s = L.scalars()
    assert s["sor_iters"] == 50
    L.close()


def test_tiles_2x2_grid_300_against_oracle():
    from les_halo_amd import TiledLES
    ip, jp, kp = 300, 300, 90
    case = inputs.make_case(ip, jp, kp)
    ph = inputs.CONFIG_PHYSICS
    T = TiledLES(ip, jp, kp, sub_i=2, sub_j=2, **{k: ph[k] for k in
                                                   ("dt", "ro", "vn", "alpha", "beta", "cs0",
                                                    "omega", "nmaxp")})
    T.load_case(case)
    o = common.oracle_from_case(case, **ph, sum_mode=SUM_EXACT)
    T.ifdata()
    o.ifdata()
    T.step(1)
    T.run(2, 1)
    for n in (1, 2):
        o.step(n)
    _bad(T.download, o, TILE_FIELDS, "2x2 tiles of 300x300x90, 2 steps (interiors)",
         interior_of=lambda n, a: common.interior(n, a, ip, jp, kp))
    _bad(T.download, o, ("u", "v", "w", "p"), "2x2 tiles of 300x300x90, 2 steps (full arrays)")
    T.close()


# the largest member (320 x 320 x 90), the smallest (96 x 96 x 60), large kp = 60 members and a
# narrow one
ENSEMBLE_MEMBERS = (40, 20, 63, 15, 4)


def test_ensemble_members_against_oracle():
    from les_halo_amd import LES, Ensemble
    specs = inputs.ensemble_specs(64)
    ph = inputs.CONFIG_PHYSICS
    members, cases = [], []
    for q in ENSEMBLE_MEMBERS:
        sp = specs[q]
        case = inputs.make_case(*sp["dims"], seed=sp["seed"], ustar=sp["ustar"])
        L = LES(*sp["dims"], **ph)
        L.load_case(case)
        L.autotune(2)
        L.ifdata()
        members.append(L)
        cases.append(case)
    assert {c["dims"][2] for c in cases} == {60, 90}
    E = Ensemble(members)
    E.run(1, 2)
    E.sync()
    for q, L, case in zip(ENSEMBLE_MEMBERS, members, cases):
        o = common.oracle_from_case(case, **ph, sum_mode=SUM_EXACT)
        o.ifdata()
        for n in (1, 2):
            o.step(n)
        _bad(L.download, o, ("u", "v", "w", "p", "sm", "f", "usum"),
             "ensemble member %d %s, 2 steps" % (q, case["dims"]))
    E.close()
    for L in members:
        L.close()
