"""The reference-order sums (les_config::sum_order = LES_SUM_REFERENCE): rhsav, pav and the SOR
residual accumulated sequentially in FP32 in press.f95's loop order, on the GPU.  With them the
whole step is the reference's arithmetic, so the GPU must reproduce

* the reference's own outputs: the digests of the golden vectors written by the reference
  build (tests/golden/ref_vectors.json), with press.f95's SOR settings (early exit at 1e-4,
  omega 1) and with the benchmark's (omega 1.5, no early exit);
* the oracle in sequential-sum mode, bit for bit, on other grids (odd jp: the in-place sweeps;
  even jp: the colour-split copy), eagerly and from the captured graph, and stage by stage.
"""
import json
import os

import numpy as np
import pytest

import common
from les_halo_amd import inputs
from oracle.oracle import SUM_SEQ

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
VEC = json.load(open(os.path.join(HERE, "golden", "ref_vectors.json")))
STATE = list(common.DUMP_ORDER)


def _les(case, **sor):
    from les_halo_amd import LES, LES_SUM_REFERENCE
    L = LES(*case["dims"], **common.REF_PHYS, **sor, sum_order=LES_SUM_REFERENCE)
    L.load_case(case)
    return L


def _digests_equal(L, dump, tag):
    bad = [n for n, (digest, _) in dump.items() if common.sha(L.download(n)) != digest]
    assert not bad, "%s: arrays differ from the reference: %s" % (tag, bad)


@pytest.mark.parametrize("key,sor", [("tiny_trajectory", "ref"), ("odd_trajectory", "ref"),
                                     ("tiny_trajectory_cfg", "cfg"),
                                     ("odd_trajectory_cfg", "cfg")])
def test_gpu_reproduces_the_reference_outputs(key, sor):
    entry = VEC[key]
    case = common.tiny_case(*entry["dims"], seed=entry["seed"])
    L = _les(case, **(common.REF_SOR if sor == "ref" else common.CFG_SOR))
    L.ifdata()
    dumps = {int(k): v for k, v in entry["dumps"].items()}
    if 0 in dumps:
        _digests_equal(L, dumps[0], "%s after init" % key)
    for n in range(1, max(dumps) + 1):
        L.step(n)
        if n in dumps:
            _digests_equal(L, dumps[n], "%s after step %d" % (key, n))
    L.close()


@pytest.mark.parametrize("dims,sor", [((40, 32, 20), "ref"), ((40, 32, 20), "cfg"),
                                      ((33, 17, 12), "ref"), ((34, 20, 11), "cfg")])
def test_trajectory_matches_sequential_oracle(dims, sor):
    sorp = common.REF_SOR if sor == "ref" else common.CFG_SOR
    case = common.tiny_case(*dims, seed=31)
    L = _les(case, **sorp)
    o = common.oracle_from_case(case, **common.REF_PHYS, **sorp, sum_mode=SUM_SEQ)
    L.ifdata()
    o.ifdata()
    for n in range(1, 4):
        L.step(n)
        o.step(n)
    L.run(4, 3)  # graph replays
    for n in range(4, 7):
        o.step(n)
    bad = [common.mismatch_report(L.download(n), o.a[n], n) for n in STATE
           if not common.bits_equal(L.download(n), o.a[n])]
    assert not bad, "\n".join(bad)
    s = L.scalars()
    assert np.float32(s["pav"]) == np.float32(o.scalars.pav)
    assert np.float32(s["rhsav"]) == np.float32(o.scalars.rhsav)
    assert s["sor_iters"] == o.scalars.sor_iters
    if sor == "ref":
        assert np.float32(s["sor"]) == np.float32(o.scalars.sor)
    L.close()


def test_stages_match_sequential_oracle():
    """les_press_rhsav, les_press_pav and the whole press on a random state."""
    case = common.tiny_case()
    ip, jp, kp = case["dims"]
    for stage, oname in (("press_rhsav", "press_rhsav"), ("press_pav", "press_pav_adj"),
                         ("press", "press")):
        L = _les(case, **common.REF_SOR)
        o = common.oracle_from_case(case, **common.REF_PHYS, **common.REF_SOR, sum_mode=SUM_SEQ)
        for name, arr in common.random_state(ip, jp, kp, 55).items():
            L.upload(name, arr)
            o.set(name, arr)
        getattr(L, stage)()
        o.call(oname)
        s = L.scalars()
        if stage == "press_pav":
            assert np.float32(s["pav"]) == np.float32(o.scalars.pav), stage
            continue
        assert np.float32(s["rhsav"]) == np.float32(o.scalars.rhsav), stage
        for n in ("rhs", "p"):
            assert common.bits_equal(L.download(n), o.a[n]), common.mismatch_report(
                L.download(n), o.a[n], "%s: %s" % (stage, n))
        L.close()


def test_rejected_where_unsupported():
    from les_halo_amd import LES_PREC_MIXED, LES_SUM_REFERENCE, LESError, TiledLES
    from les_halo_amd._lib import LesConfig, lib
    import ctypes
    with pytest.raises(LESError):
        _les(inputs.make_case(16, 12, 10, hmax=12.0), precision=LES_PREC_MIXED)
    cfg = LesConfig()
    lib().les_default_config(ctypes.byref(cfg), 32, 24, 16)
    cfg.sub_i, cfg.sub_j, cfg.pjuge, cfg.sum_order = 2, 2, 0.0, LES_SUM_REFERENCE
    h = ctypes.c_void_p()
    assert lib().les_tiled_create(ctypes.byref(cfg), ctypes.byref(h)) != 0
    assert TiledLES  # the exact-sum tiles are covered by test_tiled.py
