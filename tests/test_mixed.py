"""Mixed-precision pressure solve (LES_PREC_MIXED) against the FP64 oracle.

The GPU keeps the stencils and the SOR sweeps in FP32 and runs the rhs mean, the residual and
the refined pressure iterate in FP64 (les_mixed.hip).  The FP64 oracle runs the whole step in
FP64 with the pressure solved to the same relative residual (further blocks of SOR
iterations, which the refinement equals in exact arithmetic).  The mixed run must reach the
tolerance and stay within FP32-stencil rounding of the FP64 trajectory; the plain FP32 run with
its fixed 50 iterations is further away in p.
"""
import numpy as np
import pytest

import common
from oracle.oracle import SUM_SEQ

pytestmark = pytest.mark.gpu

TOL, RMAX = 1e-7, 400
SOR = dict(omega=1.5, pjuge=0.0, nmaxp=50)


def _gpu(case, **kw):
    from les_halo_amd import LES
    ip, jp, kp = case["dims"]
    L = LES(ip, jp, kp, **common.REF_PHYS, **SOR, **kw)
    L.load_case(case)
    L.ifdata()
    return L


@pytest.fixture(scope="module")
def case():
    return common.tiny_case(40, 32, 20, seed=11)


@pytest.mark.parametrize("romega,solver", [(0.0, 0), (1.7, 0), (0.0, 1)])
def test_mixed_reaches_tolerance_and_tracks_fp64(case, romega, solver):
    """romega: relaxation of the SOR correction solves (0 = the step's omega); solver 1: the
    corrections by preconditioned conjugate gradients."""
    from les_halo_amd.les import LES_PREC_MIXED
    nsteps = 4
    L = _gpu(case, precision=LES_PREC_MIXED, refine_tol=TOL, refine_max=RMAX, refine_omega=romega,
             refine_solver=solver)
    P = _gpu(case)
    o = common.oracle_from_case(case, np.float64, **common.REF_PHYS, **SOR, sum_mode=SUM_SEQ,
                                refine_tol=TOL, refine_max=RMAX, refine_omega=romega)
    o.ifdata()
    for n in range(1, nsteps + 1):
        L.step(n)
        P.step(n)
        o.step(n)
        s = L.scalars()
        print("step %d: residual %.3e after %d refinement passes (%d CG iterations)"
              % (n, s["resid_l2"], s["refine_passes"], s["cg_iters"]))
        assert s["resid_l2"] <= TOL, (n, s["resid_l2"], s["refine_passes"])
        assert o.scalars.resid_l2 <= TOL
    err_mixed = {nm: common.rel_l2(L.download(nm), o.a[nm]) for nm in ("u", "v", "w", "p")}
    err_plain = {nm: common.rel_l2(P.download(nm), o.a[nm]) for nm in ("u", "v", "w", "p")}
    print("rel L2 vs FP64 oracle: mixed", err_mixed, "plain", err_plain)
    for nm in ("u", "v", "w"):
        assert err_mixed[nm] < 2e-5, (nm, err_mixed[nm])
    assert err_mixed["p"] < 1e-4
    assert err_mixed["p"] < err_plain["p"]


def test_mixed_rejects_odd_jp():
    from les_halo_amd import LESError
    from les_halo_amd.les import LES_PREC_MIXED
    case = common.tiny_case(14, 11, 8, seed=9)
    L = _gpu(case, precision=LES_PREC_MIXED, refine_tol=TOL, refine_max=RMAX)
    with pytest.raises(LESError):
        L.step(1)
