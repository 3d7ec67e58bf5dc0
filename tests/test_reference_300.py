"""The HIP path against the reference's own implementation at 300x300x90 (the benchmark grid and
inputs), after 10 and 100 steps, as one domain and as 2 x 2 tiles.  The reference is the
Fortran built from the reference sources (oracle/_ref/ref300*), run on the same seeded inputs
through the same init sequence (ifdata.f95:258-273); its runs start in the background when
these tests are collected (conftest.py).

Benchmark settings (omega 1.5, exactly 50 SOR iterations, no early exit): the reference build
ref300cfg differs from the reference only in press.f95:44-46's two values.

* Reference-order sums (LES_SUM_REFERENCE: rhsav, pav and the residual accumulated in FP32 in
  press.f95's loop order): every step is the reference's arithmetic, and u, v, w, p equal the
  reference's bit for bit after 10 and 100 steps.
* Exact sums (the default and the benchmark path: each sum rounded once from its exact value):
  the only difference left is the rounding of rhsav and pav.  The reference's sequential FP32
  pav over 8.1 M cells is off by about 8e-4 (p itself is about 31 rms), which shifts its p
  uniformly: p differs by 2.7e-5 relative L2, almost all of it that constant (press.f95's
  Neumann-periodic problem fixes p only up to a constant; only its gradient enters the
  velocities).  Bars: u, v, w within 1e-5 after 10 steps; p within 1e-5 once each field's
  volume-weighted mean is removed (the raw p is an expected failure at 1e-5); every field within
  1e-3 after 100 steps.

The reference's own settings (omega 1, early exit once the FP32 residual sum is below 1e-4,
press.f95:44-46,114): with reference-order sums the GPU equals the reference bit for bit after
10 steps; with exact sums the residual's sum order can also move the exit by one iteration,
and after 100 steps every field is within 1e-3.
"""
import json
import os

import numpy as np
import pytest

import common

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

HERE = os.path.dirname(os.path.abspath(__file__))
VEC = json.load(open(os.path.join(HERE, "golden", "ref_vectors.json")))
FIELDS = ("u", "v", "w", "p")
GRID = (300, 300, 90)
_GPU = {}


def _gpu_run(tiled, sor, seq=False, steps=(10, 100)):
    """u, v, w, p after the given step counts on the GPU (cached per configuration)."""
    key = (tiled, tuple(sorted(sor.items())), seq, steps)
    if key in _GPU:
        return _GPU[key]
    from les_halo_amd import LES, LES_SUM_EXACT, LES_SUM_REFERENCE, TiledLES, inputs
    case = inputs.make_case(*GRID)
    if tiled:
        X = TiledLES(*GRID, sub_i=2, sub_j=2, **common.REF_PHYS, omega=sor["omega"],
                     nmaxp=sor["nmaxp"])
    else:
        X = LES(*GRID, **common.REF_PHYS, **sor,
                sum_order=LES_SUM_REFERENCE if seq else LES_SUM_EXACT)
    X.load_case(case)
    X.ifdata()
    out, done = {}, 0
    for n in steps:
        X.run(done + 1, n - done)
        done = n
        out[n] = {f: X.download(f) for f in FIELDS}
    X.close()
    _GPU[key] = out
    return out


def _p_gauge_free_rel_l2(a, b):
    """Relative L2 of interior p after removing each field's volume-weighted mean (FP64)."""
    from les_halo_amd import inputs
    ip, jp, kp = GRID
    case = inputs.make_case(*GRID)
    vol = (case["dx1"][2:ip + 2].astype(np.float64)[:, None, None]
           * case["dy1"][1:jp + 1].astype(np.float64)[None, :, None]
           * case["dzn"][2:kp + 2].astype(np.float64)[None, None, :])
    x = common.interior("p", a, ip, jp, kp).astype(np.float64)
    y = common.interior("p", b, ip, jp, kp).astype(np.float64)
    x = x - (vol * x).sum() / vol.sum()
    y = y - (vol * y).sum() / vol.sum()
    return float(np.linalg.norm(x - y) / np.linalg.norm(y))


def _errors(gpu, ref):
    return {n: common.rel_l2(gpu[n], ref[n]) for n in FIELDS}


def _reference_is_pinned(ref):
    """The reference run on this machine reproduced the committed digests of the same build
    (tests/golden/ref_vectors.json, written where the reference sources are)."""
    ent = VEC["ref300_trajectory_cfg"]["dumps"]
    for idx, n in enumerate((10, 100)):
        for f in FIELDS:
            assert common.sha(ref[idx][f]) == ent[str(n)][f][0], (n, f)


@pytest.mark.ref300(cfg=True)
@pytest.mark.timeout(1200)
def test_reference_order_sums_bitwise_benchmark_sor(ref300):
    ref = ref300(True)
    _reference_is_pinned(ref)
    gpu = _gpu_run(False, common.CFG_SOR, seq=True)
    for idx, n in enumerate((10, 100)):
        bad = [f for f in FIELDS if not common.bits_equal(gpu[n][f], ref[idx][f])]
        assert not bad, (n, bad, _errors(gpu[n], ref[idx]))


@pytest.mark.ref300(cfg=True)
@pytest.mark.timeout(1200)
@pytest.mark.parametrize("tiled", [False, True], ids=["single", "tiles2x2"])
def test_exact_sums_vs_reference_benchmark_sor(ref300, tiled):
    ref = ref300(True)
    gpu = _gpu_run(tiled, common.CFG_SOR)
    err10 = _errors(gpu[10], ref[0])
    pg = _p_gauge_free_rel_l2(gpu[10]["p"], ref[0]["p"])
    err100 = _errors(gpu[100], ref[1])
    print("%s: rel L2 vs reference after 10 steps %s, p without its mean %.3e; after 100 "
          "steps %s" % ("2x2 tiles" if tiled else "single domain", err10, pg, err100))
    assert all(err10[f] <= 1e-5 for f in ("u", "v", "w")), err10
    assert pg <= 1e-5, pg
    assert all(e <= 1e-3 for e in err100.values()), err100


@pytest.mark.ref300(cfg=True)
@pytest.mark.timeout(1200)
@pytest.mark.xfail(strict=True, reason="the reference's sequential FP32 pav is off by about 8e-4, "
                   "a uniform shift of its p: raw p differs by 2.7e-5 after 10 steps")
def test_exact_sums_raw_p_vs_reference_benchmark_sor(ref300):
    ref = ref300(True)
    err = _errors(_gpu_run(False, common.CFG_SOR)[10], ref[0])
    print("exact sums, 10 steps: rel L2 vs reference", err)
    assert err["p"] <= 1e-5, err


@pytest.mark.ref300(cfg=False)
@pytest.mark.timeout(1200)
def test_reference_order_sums_bitwise_early_exit(ref300):
    """press.f95's own settings, 10 steps, every sum in the reference's order."""
    ref = ref300(False)
    gpu = _gpu_run(False, common.REF_SOR, seq=True, steps=(10,))
    bad = [f for f in FIELDS if not common.bits_equal(gpu[10][f], ref[0][f])]
    assert not bad, (bad, _errors(gpu[10], ref[0]))


@pytest.mark.ref300(cfg=False)
@pytest.mark.timeout(1200)
def test_exact_sums_vs_reference_early_exit(ref300):
    """press.f95's own settings with exact sums: u, v within 1e-5 after 10 steps (w and p move by
    1.6e-5 and 2.3e-5 when the residual's order moves the exit), every field within 1e-3 after
    100 steps."""
    ref = ref300(False)
    gpu = _gpu_run(False, common.REF_SOR)
    err10, err100 = _errors(gpu[10], ref[0]), _errors(gpu[100], ref[1])
    print("reference settings, exact sums: rel L2 after 10 steps %s, after 100 steps %s"
          % (err10, err100))
    assert err10["u"] <= 1e-6 and err10["v"] <= 1e-5, err10
    assert all(e <= 1e-3 for e in err100.values()), err100
