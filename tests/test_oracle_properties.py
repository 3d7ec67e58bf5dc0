"""Pins of the oracle that do not come from the reference run: closed forms, invariants and
exact arithmetic.  They cover what the golden vectors leave open (omega != 1, which press.f95
hard-codes to 1; the exact summation mode; FP64 instantiation)."""
from fractions import Fraction

import numpy as np
import pytest

import common
from les_halo_amd import fields as fl
from les_halo_amd import inputs
from oracle.oracle import SUM_EXACT, SUM_SEQ, OracleState, exact_sum_f32


# ---------------------------------------------------------------- exact summation
def _round_f32(q):
    """Correctly rounded (ties to even) float32 of a Fraction, from first principles."""
    if q == 0:
        return np.float32(0.0)
    s = -1 if q < 0 else 1
    a = abs(q)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    while Fraction(2) ** e > a:
        e -= 1
    while Fraction(2) ** (e + 1) <= a:
        e += 1
    e = max(e, -126)  # subnormal range keeps the 2^-149 quantum
    quantum = Fraction(2) ** (e - 23)
    m = a / quantum
    fl_m = m.numerator // m.denominator
    rem = m - fl_m
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and fl_m % 2 == 1):
        fl_m += 1
    return np.float32(s * float(fl_m * quantum))


@pytest.mark.parametrize("kind", ["normal", "wide", "subnormal", "cancel", "ties"])
def test_exact_sum_is_correctly_rounded(kind):
    rng = np.random.default_rng(hash(kind) % 2**32)
    n = 3001
    if kind == "normal":
        x = rng.standard_normal(n).astype(np.float32)
    elif kind == "wide":
        x = (rng.standard_normal(n) * np.exp(rng.uniform(-60, 60, n))).astype(np.float32)
    elif kind == "subnormal":
        x = (rng.integers(-2**20, 2**20, n) * 2.0 ** -149).astype(np.float32)
    elif kind == "cancel":
        h = rng.standard_normal(n // 2).astype(np.float32)
        x = np.concatenate([h, -h[::-1], np.float32([3e-39])]).astype(np.float32)
    else:  # sums that land exactly between two floats
        x = np.float32([1.0, 2.0 ** -24, 2.0 ** -24 * 0.0, 1.0, 2.0 ** -23, 2.0 ** -24])
    exact = sum((Fraction(float(v)) for v in x), Fraction(0))
    want = _round_f32(exact)
    got = exact_sum_f32(x)
    assert got.view(np.uint32) == want.view(np.uint32), (kind, got, want)


def test_exact_sum_order_independent():
    x = (np.random.default_rng(3).standard_normal(20000) * 1e3).astype(np.float32)
    a = exact_sum_f32(x)
    b = exact_sum_f32(x[::-1].copy())
    c = exact_sum_f32(np.random.default_rng(4).permutation(x))
    assert a == b == c


# ---------------------------------------------------------------- init.f95 closed forms
def test_poisson_coefficients_closed_form():
    """Uniform 4 m horizontal spacing and a 2 m vertical cell: cn2* = cn3* = 1/16,
    cn4 from the two dzs, cn1 = 1/(sum of the six)."""
    case = inputs.make_case(12, 10, 30, seed=1)
    o = common.oracle_from_case(case)
    assert np.all(o.a["cn2l"] == np.float32(1 / 16)) and np.all(o.a["cn2s"] == np.float32(1 / 16))
    assert np.all(o.a["cn3l"] == np.float32(1 / 16)) and np.all(o.a["cn3s"] == np.float32(1 / 16))
    # k = 5: dzs(4) = dzs(5) = 2 -> cn4l = cn4s = 1/4, cn1 = 1/(1/8 + 1/8 + 1/2) = 4/3
    assert o.view("cn4l")[5] == np.float32(0.25) and o.view("cn4s")[5] == np.float32(0.25)
    assert o.view("cn1")[3, 4, 5] == np.float32(4.0 / 3.0)
    # stretched levels: cn4s/cn4l = dz2/dz1 and both reproduce 2/(dz*(dz1+dz2))
    dzs = o.view("dzs")
    for k in (22, 25, 29):
        r = float(o.view("cn4s")[k]) / float(o.view("cn4l")[k])
        assert abs(r - float(dzs[k]) / float(dzs[k - 1])) < 1e-6


# ---------------------------------------------------------------- SOR: closed-form delta response
@pytest.mark.parametrize("omega", [1.0, 1.5, 1.83])
def test_sor_delta_response(omega):
    """rhs = delta at one red cell, p = 0: after the red sweep p(c) = -omega*cn1*rhs and the
    black neighbours are still 0; after boundp1 and the black sweep each neighbour holds
    omega*cn1(n)*cn_dir(n)*p(c) with the coefficient of the direction pointing back to c."""
    ip, jp, kp = 12, 10, 40
    case = inputs.make_case(ip, jp, kp, seed=2)
    o = common.oracle_from_case(case, dtype=np.float64, omega=omega, pjuge=0.0)
    ci, cj, ck = 6, 4, 25  # a stretched level: cn4l != cn4s
    assert (ci + cj + ck) % 2 == 1  # nrd = 0 updates i = 1 + mod(k+j, 2): i+j+k odd
    rhs = o.view("rhs")
    rhs[ci, cj, ck] = 0.7
    o.call("sor_colour", 0)
    P = o.view("p")
    cn1 = o.view("cn1")
    pc = -omega * cn1[ci, cj, ck] * 0.7
    assert P[ci, cj, ck] == pytest.approx(pc, rel=1e-15)
    nz = np.argwhere(P[1:ip + 1, 1:jp + 1, 1:kp + 1] != 0)
    assert nz.shape[0] == 1
    o.call("boundp1")
    o.call("sor_colour", 1)
    V = {n: o.view(n) for n in ("cn2l", "cn2s", "cn3l", "cn3s", "cn4l", "cn4s")}
    checks = [((ci + 1, cj, ck), V["cn2s"][ci + 1]), ((ci - 1, cj, ck), V["cn2l"][ci - 1]),
              ((ci, cj + 1, ck), V["cn3s"][cj + 1]), ((ci, cj - 1, ck), V["cn3l"][cj - 1]),
              ((ci, cj, ck + 1), V["cn4s"][ck + 1]), ((ci, cj, ck - 1), V["cn4l"][ck - 1])]
    for (i, j, k), c in checks:
        assert P[i, j, k] == pytest.approx(omega * cn1[i, j, k] * c * pc, rel=1e-14), (i, j, k)
    assert V["cn4s"][ck + 1] != V["cn4l"][ck - 1]  # the vertical direction is really tested


def test_sor_overrelaxation_accelerates():
    """On a solvable Neumann/periodic problem the residual after 100 iterations is much
    smaller at omega = 1.5 than at omega = 1 (Young: the optimum lies above 1.5 here)."""
    ip, jp, kp = 16, 12, 16
    res = {}
    for omega in (1.0, 1.5):
        o = OracleState(ip, jp, kp, dtype=np.float64, omega=omega, pjuge=0.0, nmaxp=100)
        for n in ("dx1", "dy1", "dzn", "dzs", "dxs", "dys"):
            o.a[n][:] = 1.0
        o.init()
        r = o.view("rhs")
        rng = np.random.default_rng(0)
        src = rng.standard_normal((ip, jp, kp))
        r[1:ip + 1, 1:jp + 1, 1:kp + 1] = src - src.mean()
        o.call("press_sor")
        P = o.view("p")
        p = P[0:ip + 2, 0:jp + 2, 0:kp + 2]
        lap = (p[2:, 1:-1, 1:-1] + p[:-2, 1:-1, 1:-1] + p[1:-1, 2:, 1:-1] + p[1:-1, :-2, 1:-1]
               + p[1:-1, 1:-1, 2:] + p[1:-1, 1:-1, :-2] - 6 * p[1:-1, 1:-1, 1:-1])
        res[omega] = np.linalg.norm(lap - r[1:ip + 1, 1:jp + 1, 1:kp + 1])
    assert res[1.5] < 0.5 * res[1.0]


# ---------------------------------------------------------------- invariants of the physics
def _uniform_state(u0, ip=10, jp=8, kp=12):
    case = inputs.make_case(ip, jp, kp, seed=4, coverage=0.0)
    o = common.oracle_from_case(case, dtype=np.float64)
    o.a["u"][:] = u0
    return o


def test_uniform_flow_has_no_advection_or_diffusion():
    o = _uniform_state(3.0)
    o.call("velfg")
    ip, jp, kp = o.ip, o.jp, o.kp
    for n in ("f", "g", "h"):
        assert np.all(common.interior(n, o.a[n], ip, jp, kp) == 0.0), n
    o.call("les_calc_sm")
    assert np.all(common.interior("sm", o.a["sm"], ip, jp, kp) == 0.0)


def test_linear_shear_smagorinsky():
    """u = a*z (cell-centre heights), v = w = 0: dudz = a, all other gradients 0, so
    sm = (cs*delx1)^2 * |a| away from the ground and from i = 1, where the four-point average
    of les.f95:46 includes diu3(0,:,:), a halo vel2 never writes (0)."""
    ip, jp, kp = 10, 8, 30
    case = inputs.make_case(ip, jp, kp, seed=4, coverage=0.0)
    o = common.oracle_from_case(case, dtype=np.float64)
    a = -0.37
    z2 = o.a["z2"]
    dzn = o.view("dzn")
    U = o.view("u")
    for k in range(0, kp + 2):
        zc = (z2[k - 1] - 0.5 * dzn[k]) if k >= 1 else -0.5 * dzn[1]
        U[:, :, k] = a * zc
    o.call("velfg")
    o.call("les_calc_sm")
    sm = o.view("sm")
    delx1 = o.view("delx1")
    for k in range(2, kp):
        want = (0.1 * delx1[k]) ** 2 * abs(a)
        got = sm[2:ip + 1, 1:jp + 1, k]
        assert np.allclose(got, want, rtol=2e-6), k
        assert np.allclose(sm[1, 1:jp + 1, k], want / 2, rtol=2e-6), k


def test_adams_bashforth_closed_form():
    o = _uniform_state(0.0)
    rng = np.random.default_rng(1)
    for n in ("f", "fold"):
        o.a[n][:] = rng.standard_normal(o.a[n].size)
    f0 = common.interior("f", o.a["f"].copy(), o.ip, o.jp, o.kp)
    fo = o.view("fold")[:, :, :].copy()
    o.call("adam")
    assert np.allclose(common.interior("f", o.a["f"], o.ip, o.jp, o.kp), 1.5 * f0 - 0.5 * fo)
    assert np.array_equal(o.view("fold")[:, :, :], f0)


def test_press_leaves_zero_mean_pressure_and_mirror_halos():
    case = common.tiny_case()
    o = common.oracle_from_case(case, **common.REF_PHYS, **common.CFG_SOR, sum_mode=SUM_EXACT)
    o.ifdata()
    o.step(1)
    o.step(2)
    ip, jp, kp = case["dims"]
    P = o.view("p")
    vol = np.einsum("i,j,k->ijk", o.view("dx1")[1:ip + 1], o.view("dy1")[1:jp + 1],
                    o.view("dzn")[1:kp + 1]).astype(np.float64)
    pin = P[1:ip + 1, 1:jp + 1, 1:kp + 1].astype(np.float64)
    assert abs((vol * pin).sum() / vol.sum()) < 1e-6 * np.abs(pin).max()
    # boundp1 + boundp2: every halo cell is its mirror interior cell
    Mi = lambda a: 1 if a == 0 else (ip if a == ip + 1 else a)  # noqa: E731
    Mj = lambda b: jp if b == 0 else (1 if b == jp + 1 else b)  # noqa: E731
    Mk = lambda c: 1 if c == 0 else (kp if c == kp + 1 else c)  # noqa: E731
    for a in (0, 3, ip + 1):
        for b in (0, 2, jp + 1):
            for c in (0, 4, kp + 1):
                assert P[a, b, c] == P[Mi(a), Mj(b), Mk(c)]


def test_fp64_oracle_tracks_fp32():
    """The FP64 instantiation is the high-precision oracle: after a few steps the FP32 run
    differs from it by a small relative L2."""
    case = common.tiny_case()
    o32 = common.oracle_from_case(case, **common.REF_PHYS, **common.CFG_SOR, sum_mode=SUM_SEQ)
    o64 = o32.copy(np.float64)
    o32.ifdata()
    o64.ifdata()
    for n in range(1, 4):
        o32.step(n)
        o64.step(n)
    for n in ("u", "v", "w"):
        assert common.rel_l2(o32.a[n], o64.a[n]) < 1e-4, n
    assert common.rel_l2(o32.a["p"], o64.a["p"]) < 1e-3


def _press_residual(o, **over):
    o.call("press", **over)
    return o.scalars.resid_l2, o.scalars.refine_passes


def test_poisson_residual_pins():
    """The FP64 residual |rhs - L p|/|rhs| of the converged-solve mode: 1 for p = 0 and for any
    constant p (L annihilates constants, which a dropped term, a wrong sign or a wrong
    neighbour index breaks), and scale invariant."""
    case = common.tiny_case()
    o = common.oracle_from_case(case, np.float64, **common.REF_PHYS, **common.CFG_SOR)
    o.ifdata()
    o.step(1)
    base = o.copy()
    chk = dict(nmaxp=0, refine_tol=1e-300, refine_max=0)
    o.a["p"][:] = 0.0
    r0, _ = _press_residual(o, **chk)
    assert abs(r0 - 1.0) < 1e-14
    o = base.copy()
    o.a["p"][:] = 3.7
    r1, _ = _press_residual(o, **chk)
    assert abs(r1 - 1.0) < 1e-12
    o = base.copy()
    rng = np.random.default_rng(2)
    o.a["p"][:] = rng.standard_normal(o.a["p"].size)
    o2 = o.copy()
    o2.a["p"][:] *= 2.0
    o2.a["u"][:] *= 2.0
    o2.a["v"][:] *= 2.0
    o2.a["w"][:] *= 2.0
    o2.a["f"][:] *= 2.0
    o2.a["g"][:] *= 2.0
    o2.a["h"][:] *= 2.0
    ra, _ = _press_residual(o, **chk)
    rb, _ = _press_residual(o2, **chk)
    assert abs(ra - rb) < 1e-12 * max(ra, 1.0) and ra > 1e-3


def test_refinement_converges_and_continues_sor():
    """Refinement passes are further blocks of nmaxp SOR iterations, run until the FP64
    residual reaches the tolerance; the residual decreases with every block."""
    case = common.tiny_case()
    o = common.oracle_from_case(case, np.float64, **common.REF_PHYS, **common.CFG_SOR)
    o.ifdata()
    o.step(1)
    res = []
    for nm in (10, 20, 40):
        q = o.copy()
        r, _ = _press_residual(q, nmaxp=nm, refine_tol=1e-300, refine_max=0)
        res.append(r)
    assert res[0] > res[1] > res[2]
    q = o.copy()
    r, passes = _press_residual(q, nmaxp=10, refine_tol=1e-9, refine_max=500)
    assert r <= 1e-9 and passes > 0
    # the same as 10 * (passes + 1) plain iterations
    q2 = o.copy()
    r2, _ = _press_residual(q2, nmaxp=10 * (passes + 1), refine_tol=1e-300, refine_max=0)
    assert abs(r2 - r) <= 1e-6 * r


def _folded_laplacian(o, p, ip, jp, kp):
    """L p = sum_nb c_nb p_nb - p/cn1 (init.f95:82-107's coefficients from the oracle), with
    press.f95's boundary images: mirror across the i and k faces, periodic in j."""
    pe = np.concatenate([p[1:], p[-1:]], 0)
    pw = np.concatenate([p[:1], p[:-1]], 0)
    pn = np.roll(p, -1, 1)
    ps = np.roll(p, 1, 1)
    pu = np.concatenate([p[:, :, 1:], p[:, :, -1:]], 2)
    pd = np.concatenate([p[:, :, :1], p[:, :, :-1]], 2)
    a = o.a
    cn1 = a["cn1"].reshape((ip, jp, kp), order="F")
    return (a["cn2l"][:, None, None] * pe + a["cn2s"][:, None, None] * pw
            + a["cn3l"][None, :, None] * pn + a["cn3s"][None, :, None] * ps
            + a["cn4l"][None, None, :] * pu + a["cn4s"][None, None, :] * pd - p / cn1)


def _null_weight(o, ip, jp, kp):
    """(dxs(i-1)+dxs(i)) (dys(j-1)+dys(j)) (dzs(k-1)+dzs(k)); dxs(0:), dys(0:), dzs(-1:)."""
    dxs, dys, dzs = o.a["dxs"], o.a["dys"], o.a["dzs"]
    wx = dxs[0:ip] + dxs[1:ip + 1]
    wy = dys[0:jp] + dys[1:jp + 1]
    wz = dzs[1:kp + 1] + dzs[2:kp + 2]
    return wx[:, None, None] * wy[None, :, None] * wz[None, None, :]


def test_poisson_solvability_weight_and_converged_solve():
    """On a vertically stretched grid (kp = 30: spacing 2 m, x1.05 above 40 m).
    * w = (dxs(i-1)+dxs(i))(dys(j-1)+dys(j))(dzs(k-1)+dzs(k)) is the left null vector of the
      folded Poisson operator: sum w L p = 0 for any p (a wrong weight, e.g. the cell volume
      dx1 dy1 dzn, leaves a sum of the size of the terms).
    * press.f95's volume-mean removal leaves sum w rhs != 0 there, so no p solves L p = rhs;
      the converged-solve mode first removes rhs's w-mean and then reaches a tight residual
      (it stalls without that shift)."""
    ip, jp, kp = 16, 12, 30
    case = inputs.make_case(ip, jp, kp, seed=5, hmax=12.0, coverage=0.3)
    o = common.oracle_from_case(case, np.float64, **common.REF_PHYS, **common.CFG_SOR)
    dz = np.diff(o.a["dzs"][1:kp + 2])
    assert np.abs(dz).max() > 0.01  # the grid is stretched
    w = _null_weight(o, ip, jp, kp)
    rng = np.random.default_rng(4)
    for _ in range(3):
        p = rng.standard_normal((ip, jp, kp))
        lp = _folded_laplacian(o, p, ip, jp, kp)
        assert abs((w * lp).sum()) <= 1e-12 * np.abs(w * lp).sum()
        vol = (o.a["dx1"][2:ip + 2, None, None] * o.a["dy1"][None, 1:jp + 1, None]
               * o.a["dzn"][None, None, 2:kp + 2])
        assert abs((vol * lp).sum()) > 1e-6 * np.abs(vol * lp).sum()
    # the constants are the right null vector
    assert np.abs(_folded_laplacian(o, np.full((ip, jp, kp), 2.5), ip, jp, kp)).max() < 1e-12
    o.ifdata()
    o.step(1)
    base = o.copy()
    q = base.copy()
    q.call("press")  # press.f95 as written: the volume mean removed
    rhs = common.interior("rhs", q.a["rhs"], ip, jp, kp)
    # no p gets |rhs - L p| below rhs's distance to the range of L (the w-orthogonal vectors),
    # here above the tolerance reached below
    floor = abs((w * rhs).sum()) / np.linalg.norm(w) / np.linalg.norm(rhs)
    assert floor > 1e-10, floor
    q = base.copy()
    r, passes = _press_residual(q, nmaxp=50, omega=1.8, refine_tol=1e-10, refine_max=400)
    rhs = common.interior("rhs", q.a["rhs"], ip, jp, kp)
    assert abs((w * rhs).sum()) <= 1e-13 * np.abs(w * rhs).sum()
    assert r <= 1e-10, (r, passes)
    # the residual the oracle reports is that of the shifted rhs and the returned p
    p = common.interior("p", q.a["p"], ip, jp, kp)
    res = rhs - _folded_laplacian(o, p, ip, jp, kp)
    assert np.linalg.norm(res) <= 2e-10 * np.linalg.norm(rhs)
