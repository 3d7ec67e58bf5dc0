"""The oracle against the reference's own implementation: golden vectors written by
tests/golden/make_ref_vectors.py from the reference Fortran (built with flang from its own
sources).  Bit-exact equality of every array, in the oracle's sequential-sum mode."""
import json
import os

import numpy as np
import pytest

import common
from les_halo_amd import inputs
from oracle.oracle import SUM_SEQ

VEC = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.json")))


def _check(o, dump, tag):
    bad = []
    for name, (digest, norm) in dump.items():
        if common.sha(o.a[name]) != digest:
            bad.append("%s (norm %.9g vs ref %.9g)" % (name, np.linalg.norm(o.a[name].astype(np.float64)), norm))
    assert not bad, "%s: arrays differ from the reference: %s" % (tag, ", ".join(bad))


def _case_of(entry):
    ip, jp, kp = entry["dims"]
    return common.tiny_case(ip, jp, kp, seed=entry["seed"])


@pytest.mark.parametrize("key", ["tiny_trajectory", "odd_trajectory"])
def test_trajectory_bitwise(key):
    entry = VEC[key]
    case = _case_of(entry)
    o = common.oracle_from_case(case, **common.REF_PHYS, **common.REF_SOR, sum_mode=SUM_SEQ)
    o.ifdata()
    dumps = {int(k): v for k, v in entry["dumps"].items()}
    if 0 in dumps:
        _check(o, dumps[0], "%s after init" % key)
    for n in range(1, max(dumps) + 1):
        o.step(n)
        if n in dumps:
            _check(o, dumps[n], "%s after step %d" % (key, n))


@pytest.mark.parametrize("key", ["tiny_trajectory_cfg", "odd_trajectory_cfg"])
def test_trajectory_bitwise_benchmark_sor(key):
    """The benchmark's SOR settings (omega 1.5, no early exit, 50 iterations): the oracle in
    sequential-sum mode against the reference built with press.f95:44-46 set to those values."""
    entry = VEC[key]
    case = _case_of(entry)
    o = common.oracle_from_case(case, **common.REF_PHYS, **common.CFG_SOR, sum_mode=SUM_SEQ)
    o.ifdata()
    dumps = {int(k): v for k, v in entry["dumps"].items()}
    if 0 in dumps:
        _check(o, dumps[0], "%s after init" % key)
    for n in range(1, max(dumps) + 1):
        o.step(n)
        if n in dumps:
            _check(o, dumps[n], "%s after step %d" % (key, n))


@pytest.mark.slow
def test_reference_grid_150_trajectory_benchmark_sor():
    """150x150x90 (the CPU configuration's grid, terrain and inflow) with the benchmark's SOR
    settings: the sequential-sum oracle equals the reference bit for bit after 1, 5, 10 steps."""
    entry = VEC["ref150_trajectory_cfg"]
    ip, jp, kp = entry["dims"]
    case = inputs.make_case(ip, jp, kp, seed=entry["seed"])
    o = common.oracle_from_case(case, **common.REF_PHYS, **common.CFG_SOR, sum_mode=SUM_SEQ)
    o.ifdata()
    dumps = {int(k): v for k, v in entry["dumps"].items()}
    for n in range(1, max(dumps) + 1):
        o.step(n)
        if n in dumps:
            _check(o, dumps[n], "150x150x90 (omega 1.5) after step %d" % n)


STAGE_ORACLE = {"velnw": "velnw", "bondv1": "bondv1", "velfg": "velfg", "feedbf": "feedbf",
                "les": "les", "adam": "adam", "press": "press", "boundp1": "boundp1",
                "boundp2": "boundp2", "bondfg": "bondfg", "boundsm": "boundsm", "step": "step"}


@pytest.mark.parametrize("key", sorted(VEC["tiny_stages"]["stages"]))
def test_stage_bitwise(key):
    ent = VEC["tiny_stages"]["stages"][key]
    ip, jp, kp = VEC["tiny_stages"]["dims"]
    case = common.tiny_case(ip, jp, kp, seed=VEC["tiny_stages"]["seed"])
    o = common.oracle_from_case(case, **common.REF_PHYS, **common.REF_SOR, sum_mode=SUM_SEQ)
    for name, arr in common.random_state(ip, jp, kp, ent["seed"]).items():
        o.set(name, arr)
    o.call(STAGE_ORACLE[ent["stage"]], ent["n"])
    _check(o, ent["out"], key)


@pytest.mark.slow
def test_reference_grid_150_trajectory():
    if "ref150_trajectory" not in VEC:
        pytest.skip("150x150x90 vectors not generated")
    entry = VEC["ref150_trajectory"]
    ip, jp, kp = entry["dims"]
    case = inputs.make_case(ip, jp, kp, seed=entry["seed"])
    o = common.oracle_from_case(case, **common.REF_PHYS, **common.REF_SOR, sum_mode=SUM_SEQ)
    o.ifdata()
    dumps = {int(k): v for k, v in entry["dumps"].items()}
    for n in range(1, max(dumps) + 1):
        o.step(n)
        if n in dumps:
            _check(o, {k: v for k, v in dumps[n].items() if k in ("u", "v", "w", "p", "sm", "rhs", "usum")},
                   "150x150x90 after step %d" % n)


def test_tokyo_mask_rule():
    """feedbfm.f95:42-63 with the reference's own GIS terrain: the oracle's mask rule and the
    library's les_gen_masks reproduce the reference's four masks bit for bit."""
    path = os.path.join(os.path.dirname(__file__), "golden", "tokyo_150x150x90.npz")
    if not os.path.exists(path):
        pytest.skip("tokyo vectors not generated")
    d = np.load(path)
    ip, jp, kp = 150, 150, 90
    from les_halo_amd import fields as fl
    ref = {n: np.unpackbits(d[n])[:fl.size(n, ip, jp, kp)].astype(np.float32)
           for n in ["amask1", "bmask1", "cmask1", "dmask1"]}
    for n in ref:
        assert common.sha(ref[n]) == VEC["tokyo_masks"][n]
    assert common.sha(d["zbm"]) == VEC["tokyo_masks"]["zbm"]
    grid = {"z2": d["z2"], "dzn": d["dzn"]}
    masks = inputs.gen_masks(ip, jp, kp, d["zbm"], grid)
    for n in ref:
        assert common.bits_equal(masks[n], ref[n]), n
    from oracle.oracle import OracleState
    o = OracleState(ip, jp, kp)
    o.set("z2", d["z2"])
    o.set("dzn", d["dzn"])
    o.feedbfm(d["zbm"])
    for n in ref:
        assert common.bits_equal(o.a[n], ref[n]), n


def test_aveflow_statistics_bitwise():
    """aveflow.f95:42-157 after every step of a short run: every statistics array of the oracle
    equals the reference's (run with aveu/avew preceded by zeros, see ref_harness.f95 mode 3)."""
    ent = VEC["aveflow"]
    ip, jp, kp = ent["dims"]
    case = common.tiny_case(ip, jp, kp, seed=ent["seed"])
    o = common.oracle_from_case(case, **common.REF_PHYS, **common.REF_SOR, sum_mode=SUM_SEQ)
    o.ifdata()
    st = o.new_stats()
    for n in range(1, ent["nsteps"] + 1):
        o.step(n)
        o.aveflow(st, n, ent["n1"], ent["nsteps"])
    bad = [n for n, (h, _) in ent["stats"].items() if common.sha(st[n]) != h]
    assert not bad, ["%s: |%g| vs |%g|" % (n, np.linalg.norm(st[n].astype(np.float64)),
                                          ent["stats"][n][1]) for n in bad]
