"""Source generator of the adjoint form: the LES time step written once as a list of loop
statements (one Fortran loop nest or one assignment of a fused loop each, in the reference's
order and operation order), from which the forward kernels and their adjoint (transposed)
kernels are emitted as HIP templates over the floating-point type.

A statement assigns one array element per point of an index box,

    y(I,J,K) = F(x1(..), x2(..), ..., scalars)

where no point of the box reads an element another point of the box writes (true of every loop
of the step: the SOR sweeps are split by colour).  Its adjoint, run in reverse statement order,
takes yb = adj_y(I,J,K), zeroes it, and adds dF/dx * yb to the adjoint of every element F reads
(atomically: several points may read one element), keeping the contribution of y's own element.
The derivative of F is formed by reverse accumulation over F's expression tree at the forward
values of its operands; the step's nonlinear operands are final when the step ends except u, v
and w at the outflow (bondv1), which the adjoint reads from a snapshot taken before the outflow.

Run as `python -m les_halo_amd.adjgen` (les_halo_amd.build does it) to write
csrc/les_adj_gen.inc.
"""
import os

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "csrc", "les_adj_gen.inc")


# ---------------------------------------------------------------------------------------------
# index expressions (C strings over i, j, k, im, jm, km)
# ---------------------------------------------------------------------------------------------
class Ix:
    def __init__(self, s):
        self.s = s

    def __add__(self, o):
        return Ix("(%s + %s)" % (self.s, _ix(o)))

    def __sub__(self, o):
        return Ix("(%s - %s)" % (self.s, _ix(o)))

    def __radd__(self, o):
        return Ix("(%s + %s)" % (_ix(o), self.s))


def _ix(o):
    return o.s if isinstance(o, Ix) else str(o)


I, J, K = Ix("i"), Ix("j"), Ix("k")
IM, JM, KM = Ix("im"), Ix("jm"), Ix("km")


# ---------------------------------------------------------------------------------------------
# expression trees; Python's left-to-right evaluation of + - * / builds the same association
# as the C and Fortran expressions they are copied from
# ---------------------------------------------------------------------------------------------
class E:
    def __add__(self, o):
        return Op("+", self, _e(o))

    def __radd__(self, o):
        return Op("+", _e(o), self)

    def __sub__(self, o):
        return Op("-", self, _e(o))

    def __rsub__(self, o):
        return Op("-", _e(o), self)

    def __mul__(self, o):
        return Op("*", self, _e(o))

    def __rmul__(self, o):
        return Op("*", _e(o), self)

    def __truediv__(self, o):
        return Op("/", self, _e(o))

    def __rtruediv__(self, o):
        return Op("/", _e(o), self)

    def __neg__(self):
        return Neg(self)


def _e(o):
    return o if isinstance(o, E) else Num(o)


class Num(E):
    def __init__(self, v):
        self.v = float(v)


class Par(E):  # a run-time parameter (dt, ro, ...): constant
    def __init__(self, name):
        self.name = name


class Sc(E):  # a scalar the step computes (uout, rhsav, pav): active
    def __init__(self, name):
        self.name = name


# the step's scalars in the order of AdjScal (les_adj.hip); scal_add takes the position
SCALARS = ("uout", "rhsav", "pav")


class Ref(E):  # an array element; idx: tuple of C index strings
    def __init__(self, arr, idx):
        self.arr, self.idx = arr, idx


class Op(E):
    def __init__(self, op, a, b):
        self.op, self.a, self.b = op, a, b


class Neg(E):
    def __init__(self, a):
        self.a = a


class Sqrt(E):
    def __init__(self, a):
        self.a = a


class Arr:
    def __init__(self, name):
        self.name = name

    def __call__(self, *idx):
        return Ref(self.name, tuple(_ix(x) for x in idx))


# ---------------------------------------------------------------------------------------------
# arrays: Fortran bounds as (lo, hi-offset) per dimension, hi = lo-term + size symbol
# ---------------------------------------------------------------------------------------------
# 3-D: ((ilo, ihi_off), (jlo, jhi_off), (klo, khi_off)); hi = im/jm/km + off
A3 = {
    "p": ((0, 2), (0, 2), (0, 1)), "u": ((0, 1), (-1, 1), (0, 1)), "v": ((0, 1), (-1, 1), (0, 1)),
    "w": ((0, 1), (-1, 1), (-1, 1)), "usum": ((0, 0), (0, 0), (0, 0)),
    "vsum": ((0, 0), (0, 0), (0, 0)), "wsum": ((0, 0), (0, 0), (0, 0)),
    "f": ((0, 0), (0, 0), (0, 0)), "g": ((0, 0), (0, 0), (0, 0)), "h": ((0, 0), (0, 0), (0, 0)),
    "fold": ((1, 0), (1, 0), (1, 0)), "gold": ((1, 0), (1, 0), (1, 0)),
    "hold": ((1, 0), (1, 0), (1, 0)), "rhs": ((0, 1), (0, 1), (0, 1)),
    "sm": ((-1, 1), (-1, 1), (0, 1)),
    "bmask1": ((-1, 1), (0, 1), (0, 1)), "cmask1": ((0, 1), (-1, 1), (0, 1)),
    "dmask1": ((0, 1), (0, 1), (0, 1)), "cn1": ((1, 0), (1, 0), (1, 0)),
    "cd": ((1, 0), (1, 0), (1, 0)),
    "fd": ((1, 0), (1, 0), (1, 0)), "gd": ((1, 0), (1, 0), (1, 0)), "hd": ((1, 0), (1, 0), (1, 0)),
    "u_pre": ((0, 1), (-1, 1), (0, 1)), "v_pre": ((0, 1), (-1, 1), (0, 1)),
    "w_pre": ((0, 1), (-1, 1), (-1, 1)),
}
for q in range(1, 10):
    lo = -1 if q in (1, 5) else 0
    for base in ("diu", "nou", "cov"):
        A3["%s%d" % (base, q)] = ((lo, 2), (0, 2), (0, 2))
# 1-D: (lo, dim symbol, hi offset)
A1 = {"dx1": (-1, "im", 1), "dxs": (0, "im", 0), "dy1": (0, "jm", 1), "dys": (0, "jm", 0),
      "dzn": (-1, "km", 2), "dzs": (-1, "km", 2), "cn2l": (1, "im", 0), "cn2s": (1, "im", 0),
      "cn3l": (1, "jm", 0), "cn3s": (1, "jm", 0), "cn4l": (1, "km", 0), "cn4s": (1, "km", 0),
      "delx1": (1, "km", 0)}
# 2-D wind_profile(1:jm, 1:km)
A2 = {"wind": None}
# arrays whose adjoint is kept (the rest are inputs held fixed: grid, masks, coefficients)
PASSIVE = {"bmask1", "cmask1", "dmask1", "cn1", "dx1", "dxs", "dy1", "dys", "dzn", "dzs", "cn2l",
           "cn2s", "cn3l", "cn3s", "cn4l", "cn4s", "delx1"}
ARRAYS = list(A3) + list(A1) + list(A2)
SNAP = {"u": "u_pre", "v": "v_pre", "w": "w_pre"}


def flat(arr, idx):
    """C expression of the flat (Fortran column-major) offset of arr(idx)."""
    if arr in A3:
        (il, io), (jl, jo), (kl, ko) = A3[arr]
        ni, nj = "(im + %d)" % (io - il + 1), "(jm + %d)" % (jo - jl + 1)
        return "((long)(%s - (%d)) + (long)%s * ((long)(%s - (%d)) + (long)%s * (long)(%s - (%d))))" % (
            idx[0], il, ni, idx[1], jl, nj, idx[2], kl)
    if arr in A1:
        return "(long)(%s - (%d))" % (idx[0], A1[arr][0])
    return "((long)(%s - 1) + (long)jm * (long)(%s - 1))" % (idx[0], idx[1])


def size_expr(arr):
    if arr in A3:
        (il, io), (jl, jo), (kl, ko) = A3[arr]
        return "((long)(im + %d) * (long)(jm + %d) * (long)(km + %d))" % (
            io - il + 1, jo - jl + 1, ko - kl + 1)
    if arr in A1:
        lo, dim, off = A1[arr]
        return "(long)(%s + %d - (%d) + 1)" % (dim, off, lo)
    return "((long)jm * (long)km)"


# ---------------------------------------------------------------------------------------------
# statements
# ---------------------------------------------------------------------------------------------
class Stmt:
    def __init__(self, name, target, expr, k, j, i, colour=None, snap=False):
        self.name, self.target, self.expr = name, target, expr
        self.rk, self.rj, self.ri = [tuple(_ix(x) for x in r) for r in (k, j, i)]
        self.colour = colour  # None, 0 or 1: red-black nrd (points with (i+j+k+nrd) odd)
        self.snap = snap      # forward operand values of u, v, w from the pre-outflow snapshot


def build_step():
    """The step of main_ocl.f95:264-299 for n >= 2 (EXTERNAL_WIND_PROFILE, IFBF = 1), split into
    statements; 'OP' entries are the reductions and the snapshot the driver runs itself.  A
    per-cell drag coefficient cd multiplies feedbf's beta term (beta*cd*u*bmask; cd = 1 is the
    reference)."""
    u, v, w, p = Arr("u"), Arr("v"), Arr("w"), Arr("p")
    f, g, h = Arr("f"), Arr("g"), Arr("h")
    fold, gold, hold = Arr("fold"), Arr("gold"), Arr("hold")
    usum, vsum, wsum = Arr("usum"), Arr("vsum"), Arr("wsum")
    rhs, sm, cd, wind = Arr("rhs"), Arr("sm"), Arr("cd"), Arr("wind")
    bmask1, cmask1, dmask1, cn1 = Arr("bmask1"), Arr("cmask1"), Arr("dmask1"), Arr("cn1")
    dx1, dxs, dy1, dys, dzn, dzs = (Arr(n) for n in ("dx1", "dxs", "dy1", "dys", "dzn", "dzs"))
    cn2l, cn2s, cn3l, cn3s, cn4l, cn4s = (Arr(n) for n in ("cn2l", "cn2s", "cn3l", "cn3s",
                                                         "cn4l", "cn4s"))
    delx1 = Arr("delx1")
    diu = {q: Arr("diu%d" % q) for q in range(1, 10)}
    nou = {q: Arr("nou%d" % q) for q in range(1, 10)}
    cov = {q: Arr("cov%d" % q) for q in range(1, 10)}
    fd, gd, hd = Arr("fd"), Arr("gd"), Arr("hd")
    dt, ro, vn, alpha, beta, cs0, omega = (Par(n) for n in ("dt", "ro", "vn", "alpha", "beta",
                                                           "cs0", "omega"))
    uout, rhsav, pav = Sc("uout"), Sc("rhsav"), Sc("pav")
    two = Num(2.0)
    IN = ((1, KM), (1, JM), (1, IM))
    S = []

    def st(name, target, expr, k, j, i, **kw):
        S.append(Stmt(name, target, expr, k, j, i, **kw))

    # velnw.f95:27-59
    st("velnw_u", u(I, J, K), u(I, J, K) + dt * (f(I, J, K) - (-p(I, J, K) + p(I + 1, J, K)) / ro / dxs(I)), *IN)
    st("velnw_v", v(I, J, K), v(I, J, K) + dt * (g(I, J, K) - (-p(I, J, K) + p(I, J + 1, K)) / ro / dys(J)), *IN)
    st("velnw_w", w(I, J, K), w(I, J, K) + dt * (h(I, J, K) - (-p(I, J, K) + p(I, J, K + 1)) / ro / dzs(K)),
       (1, KM - 1), (1, JM), (1, IM))
    # bondv1_ext_wind_profile.f95:27-127 (n >= 2: no fill)
    st("inflow_u", u(I, J, K), wind(J, K), (1, KM), (1, JM), (0, 1))
    st("inflow_v", v(I, J, K), Num(0.0), (1, KM), (1, JM), (0, 1))
    st("inflow_w", w(I, J, K), Num(0.0), (1, KM), (1, JM), (0, 1))
    S.append("SNAP")
    S.append("UOUT")
    st("outflow_u", u(IM, J, K), u(IM, J, K) - dt * uout * (u(IM, J, K) - u(IM - 1, J, K)) / dxs(IM),
       (1, KM), (1, JM), (0, 0), snap=True)
    st("outflow_v", v(IM + 1, J, K), v(IM + 1, J, K) - dt * uout * (v(IM + 1, J, K) - v(IM, J, K)) / dxs(IM),
       (1, KM), (1, JM), (0, 0), snap=True)
    st("outflow_w", w(IM + 1, J, K), w(IM + 1, J, K) - dt * uout * (w(IM + 1, J, K) - w(IM, J, K)) / dxs(IM),
       (1, KM), (1, JM), (0, 0), snap=True)
    for a, kk in ((u, (0, KM + 1)), (v, (0, KM + 1)), (w, (0, KM))):
        st("perj0_" + a.name, a(I, 0, K), a(I, JM, K), kk, (0, 0), (0, IM + 1))
        st("perj1_" + a.name, a(I, JM + 1, K), a(I, 1, K), kk, (0, 0), (0, IM + 1))
    for a in (u, v):
        st("bot_" + a.name, a(I, J, 0), -a(I, J, 1), (0, 0), (0, JM + 1), (0, IM + 1))
        st("top_" + a.name, a(I, J, KM + 1), a(I, J, KM), (0, 0), (0, JM + 1), (0, IM + 1))
    # w(-1:im+1, -1:jm+1, 0 and km) = 0; i = -1 addresses the element before w(0, j, k) in memory
    st("bot_w", w(I, J, 0), Num(0.0), (0, 0), (-1, JM + 1), (-1, IM + 1))
    st("top_w", w(I, J, KM), Num(0.0), (0, 0), (-1, JM + 1), (-1, IM + 1))
    # vel2.f95:48-188
    st("nou1", nou[1](I, J, K), (u(I - 1, J, K) + u(I, J, K)) / two, *IN)
    st("diu1", diu[1](I, J, K), (-u(I - 1, J, K) + u(I, J, K)) / dx1(I), *IN)
    st("nou5", nou[5](I, J, K), (v(I, J - 1, K) + v(I, J, K)) / two, *IN)
    st("diu5", diu[5](I, J, K), (-v(I, J - 1, K) + v(I, J, K)) / dy1(J), *IN)
    st("nou9", nou[9](I, J, K), (w(I, J, K - 1) + w(I, J, K)) / two, *IN)
    st("diu9", diu[9](I, J, K), (-w(I, J, K - 1) + w(I, J, K)) / dzn(K), *IN)
    for q in (1, 5, 9):
        st("cov%d" % q, cov[q](I, J, K), nou[q](I, J, K) * diu[q](I, J, K), *IN)
    st("nou2", nou[2](I, J, K), (dx1(I + 1) * v(I, J - 1, K) + dx1(I) * v(I + 1, J - 1, K)) / (dx1(I) + dx1(I + 1)), *IN)
    st("diu2", diu[2](I, J, K), two * (-u(I, J - 1, K) + u(I, J, K)) / (dy1(J - 1) + dy1(J)), *IN)
    st("cov2", cov[2](I, J, K), nou[2](I, J, K) * diu[2](I, J, K), *IN)
    K1 = ((1, KM + 1), (1, JM), (1, IM))
    st("nou3", nou[3](I, J, K), (dx1(I + 1) * w(I, J, K - 1) + dx1(I) * w(I + 1, J, K - 1)) / (dx1(I) + dx1(I + 1)), *K1)
    st("diu3", diu[3](I, J, K), (-u(I, J, K - 1) + u(I, J, K)) / dzs(K - 1), *K1)
    st("cov3", cov[3](I, J, K), nou[3](I, J, K) * diu[3](I, J, K), *K1)
    st("nou4", nou[4](I, J, K), (dy1(J + 1) * u(I - 1, J, K) + dy1(J) * u(I - 1, J + 1, K)) / (dy1(J) + dy1(J + 1)), *IN)
    st("diu4", diu[4](I, J, K), two * (-v(I - 1, J, K) + v(I, J, K)) / (dx1(I - 1) + dx1(I)), *IN)
    st("cov4", cov[4](I, J, K), (nou[4](I, J, K) - Num(0.0)) * diu[4](I, J, K), *IN)
    st("nou6", nou[6](I, J, K), (dy1(J + 1) * w(I, J, K - 1) + dy1(J) * w(I, J + 1, K - 1)) / (dy1(J) + dy1(J + 1)), *K1)
    st("diu6", diu[6](I, J, K), (-v(I, J, K - 1) + v(I, J, K)) / dzs(K - 1), *K1)
    st("cov6", cov[6](I, J, K), nou[6](I, J, K) * diu[6](I, J, K), *K1)
    KL = ((1, KM - 1), (1, JM), (1, IM))
    st("nou7", nou[7](I, J, K), (dzn(K + 1) * u(I - 1, J, K) + dzn(K) * u(I - 1, J, K + 1)) / (dzn(K) + dzn(K + 1)), *KL)
    st("diu7", diu[7](I, J, K), two * (-w(I - 1, J, K) + w(I, J, K)) / (dx1(I - 1) + dx1(I)), *KL)
    st("cov7", cov[7](I, J, K), (nou[7](I, J, K) - Num(0.0)) * diu[7](I, J, K), *KL)
    st("nou8", nou[8](I, J, K), (dzn(K + 1) * v(I, J - 1, K) + dzn(K) * v(I, J - 1, K + 1)) / (dzn(K) + dzn(K + 1)), *KL)
    st("diu8", diu[8](I, J, K), two * (-w(I, J - 1, K) + w(I, J, K)) / (dy1(J - 1) + dy1(J)), *KL)
    st("cov8", cov[8](I, J, K), nou[8](I, J, K) * diu[8](I, J, K), *KL)
    # vel2.f95:125-188 halo copies
    for q, kk in ((1, (1, KM)), (4, (1, KM)), (7, (1, KM - 1))):
        for base in (nou, diu, cov):
            st("im1_%s%d" % (base[q].name[:3], q), base[q](IM + 1, J, K), base[q](IM, J, K), kk, (1, JM), (0, 0))
    for q, kk in ((2, (1, KM)), (5, (1, KM)), (8, (1, KM - 1))):
        for base in (nou, diu, cov):
            st("j0_%s%d" % (base[q].name[:3], q), base[q](I, 0, K), base[q](I, JM, K), kk, (0, 0), (1, IM))
        for base in (nou, diu, cov):
            st("j1_%s%d" % (base[q].name[:3], q), base[q](I, JM + 1, K), base[q](I, 1, K), kk, (0, 0), (1, IM))
    st("im1_diu2b", diu[2](IM + 1, J, K), diu[2](IM, J, K), (1, KM + 1), (1, JM + 1), (0, 0))
    st("im1_diu3b", diu[3](IM + 1, J, K), diu[3](IM, J, K), (1, KM + 1), (1, JM + 1), (0, 0))
    st("j0_diu4b", diu[4](I, 0, K), diu[4](I, JM, K), (1, KM + 1), (0, 0), (1, IM + 1))
    st("j0_diu6b", diu[6](I, 0, K), diu[6](I, JM, K), (1, KM + 1), (0, 0), (1, IM + 1))
    # velFG.f95:59-105
    covx1 = (dx1(I + 1) * cov[1](I, J, K) + dx1(I) * cov[1](I + 1, J, K)) / (dx1(I) + dx1(I + 1))
    covy1 = (cov[2](I, J, K) + cov[2](I, J + 1, K)) / two
    covz1 = (cov[3](I, J, K) + cov[3](I, J, K + 1)) / two
    dfu1 = (two * (-diu[1](I, J, K) + diu[1](I + 1, J, K)) / (dx1(I) + dx1(I + 1)) +
            (-diu[2](I, J, K) + diu[2](I, J + 1, K)) / dy1(J) + (-diu[3](I, J, K) + diu[3](I, J, K + 1)) / dzn(K))
    st("velfg_f", f(I, J, K), (-(covx1 + covy1 + covz1) + vn * dfu1), *IN)
    covx1 = (cov[4](I, J, K) + cov[4](I + 1, J, K)) / two
    covy1 = (dy1(J + 1) * cov[5](I, J, K) + dy1(J) * cov[5](I, J + 1, K)) / (dy1(J) + dy1(J + 1))
    covz1 = (cov[6](I, J, K) + cov[6](I, J, K + 1)) / two
    dfv1 = ((-diu[4](I, J, K) + diu[4](I + 1, J, K)) / dx1(I) +
            two * (-diu[5](I, J, K) + diu[5](I, J + 1, K)) / (dy1(J) + dy1(J + 1)) +
            (-diu[6](I, J, K) + diu[6](I, J, K + 1)) / dzn(K))
    st("velfg_g", g(I, J, K), (-(covx1 + covy1 + covz1) + vn * dfv1), *IN)
    covx1 = (cov[7](I, J, K) + cov[7](I + 1, J, K)) / two
    covy1 = (cov[8](I, J, K) + cov[8](I, J + 1, K)) / two
    covz1 = (dzn(K + 1) * cov[9](I, J, K) + dzn(K) * cov[9](I, J, K + 1)) / (dzn(K) + dzn(K + 1))
    dfw1 = ((-diu[7](I, J, K) + diu[7](I + 1, J, K)) / dx1(I) + (-diu[8](I, J, K) + diu[8](I, J + 1, K)) / dy1(J) +
            (-diu[9](I, J, K) + diu[9](I, J, K + 1)) / dzs(K))
    st("velfg_h", h(I, J, K), (-(covx1 + covy1 + covz1) + vn * dfw1), *KL)
    # feedbf.f95:41-65
    for s_, a, m in ((usum, u, bmask1), (vsum, v, cmask1), (wsum, w, dmask1)):
        st("fb_" + s_.name, s_(I, J, K), (s_(I, J, K) + a(I, J, K)) * m(I, J, K), *IN)
    for t_, s_, a, m in ((f, usum, u, bmask1), (g, vsum, v, cmask1), (h, wsum, w, dmask1)):
        st("fb_" + t_.name, t_(I, J, K),
           t_(I, J, K) + (alpha * s_(I, J, K) * dt + beta * cd(I, J, K) * a(I, J, K) * m(I, J, K)), *IN)
    # les.f95:40-56
    q4 = Num(0.25)
    d1 = diu
    dudxx1 = d1[1](I, J, K)
    dudyx1 = (d1[2](I - 1, J, K) + d1[2](I - 1, J + 1, K) + d1[2](I, J, K) + d1[2](I, J + 1, K)) * q4
    dudzx1 = (d1[3](I - 1, J, K) + d1[3](I - 1, J, K + 1) + d1[3](I, J, K) + d1[3](I, J, K + 1)) * q4
    dvdxx1 = (d1[4](I, J, K) + d1[4](I, J - 1, K) + d1[4](I + 1, J, K) + d1[4](I + 1, J - 1, K)) * q4
    dvdyx1 = d1[5](I, J, K)
    dvdzx1 = (d1[6](I, J - 1, K) + d1[6](I, J - 1, K + 1) + d1[6](I, J, K) + d1[6](I, J, K + 1)) * q4
    dwdxx1 = (d1[7](I, J, K) + d1[7](I, J, K - 1) + d1[7](I + 1, J, K) + d1[7](I + 1, J, K - 1)) * q4
    dwdyx1 = (d1[8](I, J, K) + d1[8](I, J, K - 1) + d1[8](I, J + 1, K) + d1[8](I, J + 1, K - 1)) * q4
    dwdzx1 = d1[9](I, J, K)
    cdk = cs0 * delx1(K)
    s = (two * (dudxx1 * dudxx1 + dvdyx1 * dvdyx1 + dwdzx1 * dwdzx1) +
         (dudyx1 + dvdxx1) * (dudyx1 + dvdxx1) + (dwdyx1 + dvdzx1) * (dwdyx1 + dvdzx1) +
         (dudzx1 + dwdxx1) * (dudzx1 + dwdxx1))
    st("calc_sm", sm(I, J, K), (cdk * cdk) * Sqrt(s), *IN)
    # boundsm.f95:13-32
    st("bsm_i0", sm(0, J, K), sm(1, J, K), (0, KM + 1), (-1, JM + 1), (0, 0))
    st("bsm_i1", sm(IM + 1, J, K), sm(IM, J, K), (0, KM + 1), (-1, JM + 1), (0, 0))
    st("bsm_j1", sm(I, JM + 1, K), sm(I, JM, K), (0, KM + 1), (0, 0), (0, IM + 1))
    st("bsm_j0", sm(I, 0, K), sm(I, 1, K), (0, KM + 1), (0, 0), (0, IM + 1))
    st("bsm_k0", sm(I, J, 0), -sm(I, J, 1), (0, 0), (-1, JM + 1), (0, IM + 1))
    st("bsm_k1", sm(I, J, KM + 1), sm(I, J, KM), (0, 0), (-1, JM + 1), (0, IM + 1))
    # les.f95:78-180
    def xi(a, b):  # (dx1(i+1)*a + dx1(i)*b)/(dx1(i)+dx1(i+1))
        return (dx1(I + 1) * a + dx1(I) * b) / (dx1(I) + dx1(I + 1))

    def xm(a, b):  # (dx1(i)*a + dx1(i-1)*b)/(dx1(i-1)+dx1(i))
        return (dx1(I) * a + dx1(I - 1) * b) / (dx1(I - 1) + dx1(I))

    def yi(a, b):
        return (dy1(J + 1) * a + dy1(J) * b) / (dy1(J) + dy1(J + 1))

    def ym(a, b):
        return (dy1(J) * a + dy1(J - 1) * b) / (dy1(J - 1) + dy1(J))

    S_ = lambda di, dj, dk: sm(I + di if di >= 0 else I - (-di), J + dj if dj >= 0 else J - (-dj),
                               K + dk if dk >= 0 else K - (-dk))
    evsx2 = S_(1, 0, 0)
    evsx1 = S_(0, 0, 0)
    evsy2 = (dy1(J + 1) * xi(S_(0, 0, 0), S_(1, 0, 0)) + dy1(J) * xi(S_(0, 1, 0), S_(1, 1, 0))) / (dy1(J) + dy1(J + 1))
    evsy1 = (dy1(J + 1) * xi(S_(0, -1, 0), S_(1, -1, 0)) + dy1(J) * xi(S_(0, 0, 0), S_(1, 0, 0))) / (dy1(J) + dy1(J + 1))
    evsz2 = (dzn(K + 1) * xi(S_(0, 0, 0), S_(1, 0, 0)) + dzn(K) * xi(S_(0, 0, 1), S_(1, 0, 1))) / (dzn(K) + dzn(K + 1))
    evsz1 = (dzn(K) * xi(S_(0, 0, -1), S_(1, 0, -1)) + dzn(K - 1) * xi(S_(0, 0, 0), S_(1, 0, 0))) / (dzn(K - 1) + dzn(K))
    visux2 = (evsx2) * two * d1[1](I + 1, J, K)
    visux1 = (evsx1) * two * d1[1](I, J, K)
    visuy2 = (evsy2) * (d1[2](I, J + 1, K) + d1[4](I + 1, J, K))
    visuy1 = (evsy1) * (d1[2](I, J, K) + d1[4](I + 1, J - 1, K))
    visuz2 = (evsz2) * (d1[3](I, J, K + 1) + d1[7](I + 1, J, K))
    visuz1 = (evsz1) * (d1[3](I, J, K) + d1[7](I + 1, J, K - 1))
    vfu = (visux2 - visux1) / dx1(I) + (visuy2 - visuy1) / dy1(J) + (visuz2 - visuz1) / dzn(K)
    st("visc_f", f(I, J, K), (f(I, J, K) + vfu), *IN)
    evsy2 = S_(0, 1, 0)
    evsy1 = S_(0, 0, 0)
    evsx2 = (dy1(J + 1) * xi(S_(0, 0, 0), S_(1, 0, 0)) + dy1(J) * xi(S_(0, 1, 0), S_(1, 1, 0))) / (dy1(J) + dy1(J + 1))
    evsx1 = (dy1(J + 1) * xm(S_(-1, 0, 0), S_(0, 0, 0)) + dy1(J) * xm(S_(-1, 1, 0), S_(0, 1, 0))) / (dy1(J) + dy1(J + 1))
    evsz2 = (dzn(K + 1) * xi(S_(0, 0, 0), S_(1, 0, 0)) + dzn(K) * xi(S_(0, 0, 1), S_(1, 0, 1))) / (dzn(K) + dzn(K + 1))
    evsz1 = (dzn(K) * xi(S_(0, 0, -1), S_(1, 0, -1)) + dzn(K - 1) * xi(S_(0, 0, 0), S_(1, 0, 0))) / (dzn(K - 1) + dzn(K))
    visvx2 = (evsx2) * (d1[2](I, J + 1, K) + d1[4](I + 1, J, K))
    visvx1 = (evsx1) * (d1[2](I - 1, J + 1, K) + d1[4](I, J, K))
    visvy2 = (evsy2) * two * d1[5](I, J + 1, K)
    visvy1 = (evsy1) * two * d1[5](I, J, K)
    visvz2 = (evsz2) * (d1[6](I, J, K + 1) + d1[8](I, J + 1, K))
    visvz1 = (evsz1) * (d1[6](I, J, K) + d1[8](I, J + 1, K - 1))
    vfv = (visvx2 - visvx1) / dx1(I) + (visvy2 - visvy1) / dy1(J) + (visvz2 - visvz1) / dzn(K)
    st("visc_g", g(I, J, K), (g(I, J, K) + vfv), *IN)
    evsz2 = S_(0, 0, 1)
    evsz1 = S_(0, 0, 0)
    evsx2 = (dzn(K + 1) * xi(S_(0, 0, 0), S_(1, 0, 0)) + dzn(K) * xi(S_(0, 0, 1), S_(1, 0, 1))) / (dzn(K) + dzn(K + 1))
    evsx1 = (dzn(K + 1) * xm(S_(-1, 0, 0), S_(0, 0, 0)) + dzn(K) * xm(S_(-1, 0, 1), S_(0, 0, 1))) / (dzn(K) + dzn(K + 1))
    evsy2 = (dzn(K + 1) * yi(S_(0, 0, 0), S_(0, 1, 0)) + dzn(K) * yi(S_(0, 0, 1), S_(0, 1, 1))) / (dzn(K) + dzn(K + 1))
    evsy1 = (dzn(K + 1) * ym(S_(0, -1, 0), S_(0, 0, 0)) + dzn(K) * ym(S_(0, -1, 1), S_(0, 0, 1))) / (dzn(K) + dzn(K + 1))
    viswx2 = (evsx2) * (d1[3](I, J, K + 1) + d1[7](I + 1, J, K))
    viswx1 = (evsx1) * (d1[3](I - 1, J, K + 1) + d1[7](I, J, K))
    viswy2 = (evsy2) * (d1[6](I, J, K + 1) + d1[8](I, J + 1, K))
    viswy1 = (evsy1) * (d1[6](I, J - 1, K + 1) + d1[8](I, J, K))
    viswz2 = (evsz2) * two * d1[9](I, J, K + 1)
    viswz1 = (evsz1) * two * d1[9](I, J, K)
    vfw = (viswx2 - viswx1) / dx1(I) + (viswy2 - viswy1) / dy1(J) + (viswz2 - viswz1) / dzn(K)
    st("visc_h", h(I, J, K), (h(I, J, K) + vfw), *IN)
    # adam.f95:35-43
    c15, c05 = Num(1.5), Num(0.5)
    for t_, o_, d_ in ((f, fold, fd), (g, gold, gd), (h, hold, hd)):
        st("adam_save_" + t_.name, d_(I, J, K), t_(I, J, K), *IN)
        st("adam_" + t_.name, t_(I, J, K), c15 * t_(I, J, K) - c05 * o_(I, J, K), *IN)
        st("adam_old_" + t_.name, o_(I, J, K), d_(I, J, K), *IN)
    # bondFG.f95:15-32
    st("bfg_f", f(0, J, K), f(1, J, K), (1, KM), (1, JM), (0, 0))
    st("bfg_g", g(I, 0, K), g(I, JM, K), (1, KM), (0, 0), (1, IM))
    st("bfg_h0", h(I, J, 0), Num(0.0), (0, 0), (1, JM), (1, IM))
    st("bfg_h1", h(I, J, KM), Num(0.0), (0, 0), (1, JM), (1, IM))
    # press.f95:49-82
    st("rhs_uvw", rhs(I, J, K), (-u(I - 1, J, K) + u(I, J, K)) / dx1(I) + (-v(I, J - 1, K) + v(I, J, K)) / dy1(J) +
       (-w(I, J, K - 1) + w(I, J, K)) / dzn(K), *IN)
    st("rhs_fgh", rhs(I, J, K), (f(I, J, K) - f(I - 1, J, K)) / dx1(I) + (g(I, J, K) - g(I, J - 1, K)) / dy1(J) +
       (h(I, J, K) - h(I, J, K - 1)) / dzn(K) + rhs(I, J, K) / dt, *IN)
    S.append("RHSAV")
    st("rhs_adj", rhs(I, J, K), rhs(I, J, K) - rhsav, *IN)
    S.append("SOR")
    S.append("PAV")
    st("pav_adj", p(I, J, K), p(I, J, K) - pav, *IN)
    S.extend(boundp1(p) + boundp2(p))
    return S


def boundp1(p):
    return [Stmt("bp1_i0", p(0, J, K), p(1, J, K), (0, KM + 1), (0, JM + 1), (0, 0)),
            Stmt("bp1_i1", p(IM + 1, J, K), p(IM, J, K), (0, KM + 1), (0, JM + 1), (0, 0)),
            Stmt("bp1_j0", p(I, 0, K), p(I, JM, K), (0, KM + 1), (0, 0), (0, IM + 1)),
            Stmt("bp1_j1", p(I, JM + 1, K), p(I, 1, K), (0, KM + 1), (0, 0), (0, IM + 1))]


def boundp2(p):
    return [Stmt("bp2_k0", p(I, J, 0), p(I, J, 1), (0, 0), (0, JM + 1), (0, IM + 1)),
            Stmt("bp2_k1", p(I, J, KM + 1), p(I, J, KM), (0, 0), (0, JM + 1), (0, IM + 1))]


def build_sor():
    """One SOR iteration of press.f95:90-104: colour nrd = 0, boundp1, colour 1, boundp1,
    boundp2."""
    p, rhs, cn1 = Arr("p"), Arr("rhs"), Arr("cn1")
    cn2l, cn2s, cn3l, cn3s, cn4l, cn4s = (Arr(n) for n in ("cn2l", "cn2s", "cn3l", "cn3s",
                                                         "cn4l", "cn4s"))
    omega = Par("omega")
    out = []
    for nrd in (0, 1):
        reltmp = omega * (cn1(I, J, K) * (cn2l(I) * p(I + 1, J, K) + cn2s(I) * p(I - 1, J, K) +
                                          cn3l(J) * p(I, J + 1, K) + cn3s(J) * p(I, J - 1, K) +
                                          cn4l(K) * p(I, J, K + 1) + cn4s(K) * p(I, J, K - 1) -
                                          rhs(I, J, K)) - p(I, J, K))
        out.append(Stmt("sor%d" % nrd, p(I, J, K), p(I, J, K) + reltmp, (1, KM), (1, JM), (1, IM),
                        colour=nrd))
        out.extend(boundp1(p))
    out.extend(boundp2(p))
    return out


# ---------------------------------------------------------------------------------------------
# emission
# ---------------------------------------------------------------------------------------------
def _leaves(e, acc):
    if isinstance(e, Ref):
        acc.setdefault((e.arr, e.idx), None)
    elif isinstance(e, Sc):
        acc.setdefault(("#" + e.name, ()), None)
    elif isinstance(e, Op):
        _leaves(e.a, acc)
        _leaves(e.b, acc)
    elif isinstance(e, (Neg, Sqrt)):
        _leaves(e.a, acc)
    return acc


def _load(arr, idx, fwd_alias=None):
    a = fwd_alias.get(arr, arr) if fwd_alias else arr
    return "A.%s[%s]" % (a, flat(arr, idx))


class Emitter:
    """Emits the forward value of every node as a local (n<id>) and, for the adjoint, the
    reverse accumulation into locals b<id>, then one scatter per distinct leaf."""

    def __init__(self):
        self.lines, self.n = [], 0
        self.ids = {}

    def new(self):
        self.n += 1
        return self.n

    def fwd(self, e, leafvar):
        """Forward value of e; leafvar maps a leaf key to the local holding its value."""
        if isinstance(e, Num):
            return "T(%r)" % e.v
        if isinstance(e, Par):
            return "P.%s" % e.name
        if isinstance(e, Sc):
            return leafvar[("#" + e.name, ())]
        if isinstance(e, Ref):
            return leafvar[(e.arr, e.idx)]
        if isinstance(e, Op):
            a, b = self.fwd(e.a, leafvar), self.fwd(e.b, leafvar)
            v = self.new()
            self.lines.append("const T n%d = %s %s %s;" % (v, a, e.op, b))
            self.ids[id(e)] = ("n%d" % v, a, b)
            return "n%d" % v
        if isinstance(e, Neg):
            a = self.fwd(e.a, leafvar)
            v = self.new()
            self.lines.append("const T n%d = -%s;" % (v, a))
            self.ids[id(e)] = ("n%d" % v, a, None)
            return "n%d" % v
        if isinstance(e, Sqrt):
            a = self.fwd(e.a, leafvar)
            v = self.new()
            self.lines.append("const T n%d = sqrt(%s);" % (v, a))
            self.ids[id(e)] = ("n%d" % v, a, None)
            return "n%d" % v
        raise TypeError(e)

    def rev(self, e, bar, leafbar):
        """Adds the adjoint 'bar' (a C expression) of e into its operands."""
        if isinstance(e, (Num, Par)):
            return
        if isinstance(e, Sc):
            leafbar.setdefault(("#" + e.name, ()), []).append(bar)
            return
        if isinstance(e, Ref):
            if e.arr in PASSIVE:
                return
            leafbar.setdefault((e.arr, e.idx), []).append(bar)
            return
        v, a, b = self.ids[id(e)]
        bv = "b%d" % self.new()
        self.lines.append("const T %s = %s;" % (bv, bar))
        if isinstance(e, Neg):
            self.rev(e.a, "-%s" % bv, leafbar)
        elif isinstance(e, Sqrt):
            self.rev(e.a, "(%s > T(0) ? %s / (T(2) * %s) : T(0))" % (a, bv, v), leafbar)
        elif e.op == "+":
            self.rev(e.a, bv, leafbar)
            self.rev(e.b, bv, leafbar)
        elif e.op == "-":
            self.rev(e.a, bv, leafbar)
            self.rev(e.b, "-%s" % bv, leafbar)
        elif e.op == "*":
            self.rev(e.a, "%s * %s" % (bv, b), leafbar)
            self.rev(e.b, "%s * %s" % (bv, a), leafbar)
        elif e.op == "/":
            self.rev(e.a, "%s / %s" % (bv, b), leafbar)
            self.rev(e.b, "-%s * %s / %s" % (bv, v, b), leafbar)


def _decode(s):
    """Kernel prologue: thread -> (i, j, k) of the statement's box (i fastest)."""
    (klo, khi), (jlo, jhi), (ilo, ihi) = s.rk, s.rj, s.ri
    return [
        "const int i0 = %s, ni = %s - i0 + 1;" % (ilo, ihi),
        "const int j0 = %s, nj = %s - j0 + 1;" % (jlo, jhi),
        "const int k0 = %s, nk = %s - k0 + 1;" % (klo, khi),
        "const long t = (long)blockIdx.x * 256 + threadIdx.x;",
        "bool act = ni > 0 && nj > 0 && nk > 0 && t < (long)ni * nj * nk;",
        "const long tt = act ? t : 0;",
        "const int i = i0 + (int)(tt % (ni > 0 ? ni : 1));",
        "const int j = j0 + (int)((tt / (ni > 0 ? ni : 1)) % (nj > 0 ? nj : 1));",
        "const int k = k0 + (int)(tt / ((long)(ni > 0 ? ni : 1) * (nj > 0 ? nj : 1)));",
    ] + (["act = act && (((i + j + k + %d) & 1) == 1);" % s.colour] if s.colour is not None else [])


def emit_stmt(s, sid):
    out = []
    leaves = list(_leaves(s.expr, {}))
    tgt = (s.target.arr, s.target.idx)
    # forward kernel
    em = Emitter()
    lv = {}
    for n, key in enumerate(leaves):
        lv[key] = "l%d" % n
    body = []
    for n, key in enumerate(leaves):
        if key[0].startswith("#"):
            body.append("const T l%d = S->%s;" % (n, key[0][1:]))
        else:
            body.append("const T l%d = %s;" % (n, _load(key[0], key[1])))
    val = em.fwd(s.expr, lv)
    out.append("template <typename T>\n__global__ __launch_bounds__(256) void fw_%s(AdjArr<T> A, "
               "const AdjScal<T>* S, AdjPar<T> P, int im, int jm, int km) {" % s.name)
    out += ["  " + l for l in _decode(s)]
    out.append("  if (!act) return;")
    out += ["  " + l for l in body + em.lines]
    out.append("  A.%s[%s] = %s;" % (tgt[0], flat(*tgt), val))
    out.append("}")
    # adjoint kernel
    em = Emitter()
    alias = SNAP if s.snap else None
    body = []
    for n, key in enumerate(leaves):
        if key[0].startswith("#"):
            body.append("const T l%d = act ? S->%s : T(0);" % (n, key[0][1:]))
        else:
            body.append("const T l%d = act ? %s : T(0);" % (n, _load(key[0], key[1], alias)))
    em.fwd(s.expr, lv)
    fwd_lines = em.lines
    em.lines = []
    leafbar = {}
    em.rev(s.expr, "yb", leafbar)
    out.append("template <typename T>\n__global__ __launch_bounds__(256) void bw_%s(AdjArr<T> A, "
               "AdjArr<T> B, const AdjScal<T>* S, AdjScal<T>* SB, AdjPar<T> P, int im, int jm, "
               "int km) {" % s.name)
    out += ["  " + l for l in _decode(s)]
    out.append("  const long yi = act ? %s : 0;" % flat(*tgt))
    out.append("  const T yb = act ? B.%s[yi] : T(0);" % tgt[0])
    out += ["  " + l for l in body + fwd_lines + em.lines]
    self_bar = None
    scal = []
    for key, bars in leafbar.items():
        sumv = " + ".join("(%s)" % b for b in bars)
        if key == tgt:
            self_bar = sumv
        elif key[0].startswith("#"):
            scal.append((key[0][1:], sumv))
        else:
            out.append("  if (act) atomicAdd(&B.%s[%s], %s);" % (key[0], flat(*key), sumv))
    out.append("  if (act) B.%s[yi] = %s;" % (tgt[0], self_bar if self_bar else "T(0)"))
    for name, sumv in scal:
        out.append("  scal_add(SB, %d, act ? (%s) : T(0));" % (SCALARS.index(name), sumv))
    out.append("}")
    return out


def _extent(s):
    (klo, khi), (jlo, jhi), (ilo, ihi) = s.rk, s.rj, s.ri
    return "((long)std::max(0, (%s) - (%s) + 1) * std::max(0, (%s) - (%s) + 1) * std::max(0, (%s) - (%s) + 1))" % (
        ihi, ilo, jhi, jlo, khi, klo)


def generate():
    step = build_step()
    sor = build_sor()
    stmts = [s for s in step if isinstance(s, Stmt)] + sor
    uniq, seen = [], {}
    for s in stmts:  # boundp1/boundp2 occur in the SOR iteration and after pav: one kernel each
        code = "\n".join(emit_stmt(s, 0))
        if s.name in seen:
            assert seen[s.name] == code, s.name
            continue
        seen[s.name] = code
        uniq.append(s)
    out = ["// Generated by les_halo_amd/adjgen.py from its statement list of the LES step; do not",
           "// edit.  fw_<name>: the statement; bw_<name>: its adjoint.", ""]
    out.append("template <typename T>\nstruct AdjArr {")
    for a in ARRAYS:
        out.append("  T* %s;" % a)
    out.append("};")
    out.append("inline long adj_array_size(int a, int im, int jm, int km) {\n  switch (a) {")
    for n, a in enumerate(ARRAYS):
        out.append("    case %d: return %s;" % (n, size_expr(a)))
    out.append("  }\n  return 0;\n}")
    out.append("constexpr int ADJ_NARRAYS = %d;" % len(ARRAYS))
    out.append("enum AdjArray { %s };" % ", ".join("ADJ_A_%s" % a for a in ARRAYS))
    out.append("static const char* const ADJ_ARRAY_NAMES[ADJ_NARRAYS] = {%s};" %
               ", ".join('"%s"' % a for a in ARRAYS))
    out.append("template <typename T>\ninline T*& adj_slot(AdjArr<T>& A, int a) {\n  switch (a) {")
    for n, a in enumerate(ARRAYS):
        out.append("    case %d: return A.%s;" % (n, a))
    out.append("  }\n  return A.%s;\n}" % ARRAYS[0])
    out.append("constexpr bool ADJ_PASSIVE[ADJ_NARRAYS] = {%s};" %
               ", ".join("true" if a in PASSIVE else "false" for a in ARRAYS))
    # arrays the step writes (statement targets and the pre-outflow snapshot); the others are
    # constants of the window
    written = {st.target.arr for st in stmts} | set(SNAP.values())
    out.append("constexpr bool ADJ_WRITTEN[ADJ_NARRAYS] = {%s};" %
               ", ".join("true" if a in written else "false" for a in ARRAYS))
    for s in uniq:
        out += emit_stmt(s, 0)
        out.append("")
    # statement table
    out.append("enum AdjStmt {")
    for s in uniq:
        out.append("  ADJ_%s," % s.name)
    out.append("  ADJ_NSTMT\n};")
    out.append("inline long adj_stmt_extent(int s, int im, int jm, int km) {\n  switch (s) {")
    for s in uniq:
        out.append("    case ADJ_%s: return %s;" % (s.name, _extent(s)))
    out.append("  }\n  return 0;\n}")
    out.append("template <typename T>\nvoid adj_launch_fw(int s, hipStream_t st, const AdjArr<T>& A, "
               "const AdjScal<T>* S, const AdjPar<T>& P, int im, int jm, int km) {")
    out.append("  const long n = adj_stmt_extent(s, im, jm, km);\n  if (n <= 0) return;")
    out.append("  const dim3 grid((unsigned)((n + 255) / 256));\n  switch (s) {")
    for s in uniq:
        out.append("    case ADJ_%s: hipLaunchKernelGGL(fw_%s<T>, grid, dim3(256), 0, st, A, S, P, im, jm, km); break;"
                   % (s.name, s.name))
    out.append("  }\n}")
    out.append("template <typename T>\nvoid adj_launch_bw(int s, hipStream_t st, const AdjArr<T>& A, "
               "const AdjArr<T>& B, const AdjScal<T>* S, AdjScal<T>* SB, const AdjPar<T>& P, int im, "
               "int jm, int km) {")
    out.append("  const long n = adj_stmt_extent(s, im, jm, km);\n  if (n <= 0) return;")
    out.append("  const dim3 grid((unsigned)((n + 255) / 256));\n  switch (s) {")
    for s in uniq:
        out.append("    case ADJ_%s: hipLaunchKernelGGL(bw_%s<T>, grid, dim3(256), 0, st, A, B, S, SB, P, im, jm, km); break;"
                   % (s.name, s.name))
    out.append("  }\n}")
    # the step as an op list (statement ids and markers)
    def ops_of(seq):
        r = []
        for s in seq:
            if isinstance(s, Stmt):
                r.append("ADJ_%s" % s.name)
            else:
                r.append("ADJ_OP_%s" % s)
        return r
    out.append("enum { ADJ_OP_SNAP = -1, ADJ_OP_UOUT = -2, ADJ_OP_RHSAV = -3, ADJ_OP_SOR = -4, "
               "ADJ_OP_PAV = -5 };")
    st_ops = ops_of(step)
    out.append("static const int ADJ_STEP_OPS[] = {%s};" % ", ".join(st_ops))
    out.append("constexpr int ADJ_STEP_NOPS = %d;" % len(st_ops))
    so = ops_of(sor)
    out.append("static const int ADJ_SOR_OPS[] = {%s};" % ", ".join(so))
    out.append("constexpr int ADJ_SOR_NOPS = %d;" % len(so))
    return "\n".join(out) + "\n"


def write(path=OUT):
    src = generate()
    old = None
    if os.path.exists(path):
        with open(path) as fh:
            old = fh.read()
    if old != src:
        with open(path, "w") as fh:
            fh.write(src)
    return path


if __name__ == "__main__":
    print(write())
