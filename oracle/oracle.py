"""ctypes front end of the CPU oracle (liboracle.so) - test infrastructure only.

OracleState holds every array of the reference main program (main_ocl.f95:40-167) in its
Fortran bounds, flat and column-major, in float32 (reference arithmetic) or float64 (the
high-precision oracle).  Each method runs the restated reference subroutine of the same name.
"""
import ctypes
import os
import subprocess

import numpy as np

from les_halo_amd import fields as fl

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")

_PTR_NAMES = ["p", "u", "v", "w", "usum", "vsum", "wsum", "f", "g", "h", "fold", "gold", "hold",
              "rhs", "sm"]
_TAIL_NAMES = ["amask1", "bmask1", "cmask1", "dmask1", "cn1", "dx1", "dxs", "dy1", "dys", "dzn",
               "dzs", "z2", "cn2l", "cn2s", "cn3l", "cn3s", "cn4l", "cn4s", "delx1", "wind",
               "fx", "fy", "fz", "dfu1", "dfv1", "dfw1"]


class OrcFields(ctypes.Structure):
    _fields_ = ([("ip", ctypes.c_int), ("jp", ctypes.c_int), ("kp", ctypes.c_int)]
                + [(n, ctypes.c_void_p) for n in _PTR_NAMES]
                + [("diu", ctypes.c_void_p * 9), ("nou", ctypes.c_void_p * 9),
                   ("cov", ctypes.c_void_p * 9)]
                + [(n, ctypes.c_void_p) for n in _TAIL_NAMES])


class OrcParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ["dt", "ro", "vn", "alpha", "beta", "cs0", "omega", "pjuge"]] + \
               [("nmaxp", ctypes.c_int), ("sum_mode", ctypes.c_int),
                ("refine_tol", ctypes.c_double), ("refine_max", ctypes.c_int),
                ("refine_omega", ctypes.c_double)]


class OrcScalars(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ["uout", "rhsav", "area", "pav", "pco", "sor"]] + \
               [("sor_iters", ctypes.c_int), ("resid_l2", ctypes.c_double),
                ("refine_passes", ctypes.c_int)]


AVE3 = ["aveu", "avev", "avew", "avep", "avel", "aveuu", "avevv", "aveww", "avesm", "avesmsm",
        "uwfx"]
AVE2 = ["avesu", "avesv", "avesw", "avesuu", "avesvv", "avesww", "uwfxs"]


class OrcAve(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in AVE3 + AVE2]


SUM_SEQ, SUM_EXACT = 0, 1
_STAGES = ["velnw", "bondv1", "vel2", "velfg", "feedbf", "les_calc_sm", "boundsm",
           "les_calc_visc", "les", "adam", "bondfg", "boundp1", "boundp2", "press_rhsav",
           "sor_colour", "press_sor", "press_pav_adj", "press", "step", "ifdata",
           "init_coeffs", "delx1"]

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        for s in _STAGES:
            for sfx in ("_f32", "_f64"):
                fn = getattr(L, "orc_" + s + sfx)
                fn.restype = ctypes.c_int
                fn.argtypes = [ctypes.POINTER(OrcFields), ctypes.POINTER(OrcParams), ctypes.c_int,
                               ctypes.POINTER(OrcScalars)]
        for sfx, ct in (("_f32", ctypes.c_float), ("_f64", ctypes.c_double)):
            fn = getattr(L, "orc_feedbfm" + sfx)
            fn.restype = ctypes.c_int
            fn.argtypes = [ctypes.POINTER(OrcFields), ctypes.c_void_p]
        L.orc_exact_sum_f32.restype = ctypes.c_float
        L.orc_exact_sum_f32.argtypes = [ctypes.c_void_p, ctypes.c_long]
        _lib = L
    return _lib


def exact_sum_f32(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    return np.float32(lib().orc_exact_sum_f32(x.ctypes.data, x.size))


_SCRATCH = {
    "nou1": [(-1, 2), (0, 2), (0, 2)], "nou5": [(-1, 2), (0, 2), (0, 2)],
    "cov1": [(-1, 2), (0, 2), (0, 2)], "cov5": [(-1, 2), (0, 2), (0, 2)],
    "fx": [(0, 0), (0, 0), (0, 0)], "fy": [(0, 0), (0, 0), (0, 0)], "fz": [(0, 0), (0, 0), (0, 0)],
}


def _scratch_size(name, ip, jp, kp):
    if name.startswith("nou") or name.startswith("cov"):
        lo0 = -1 if name[-1] in "15" else 0
        return (ip + 2 - lo0 + 1) * (jp + 3) * (kp + 3)
    if name in ("fx", "fy", "fz"):
        return (ip + 1) * (jp + 1) * (kp + 1)
    if name == "dfu1":
        return (ip + 1) * jp * kp
    if name == "dfv1":
        return ip * (jp + 1) * kp
    if name == "dfw1":
        return ip * jp * kp
    raise KeyError(name)


class OracleState:
    """All arrays of one domain (flat numpy, Fortran bounds) plus the physics parameters."""

    def __init__(self, ip, jp, kp, dtype=np.float32, **physics):
        self.ip, self.jp, self.kp = ip, jp, kp
        self.dtype = np.dtype(dtype)
        self.sfx = "_f32" if self.dtype == np.float32 else "_f64"
        self.a = {}
        for n in fl.FIELDS:
            self.a[n] = np.zeros(fl.size(n, ip, jp, kp), self.dtype)
        for q in range(1, 10):
            self.a["nou%d" % q] = np.zeros(_scratch_size("nou%d" % q, ip, jp, kp), self.dtype)
            self.a["cov%d" % q] = np.zeros(_scratch_size("cov%d" % q, ip, jp, kp), self.dtype)
        for n in ("fx", "fy", "fz", "dfu1", "dfv1", "dfw1"):
            self.a[n] = np.zeros(_scratch_size(n, ip, jp, kp), self.dtype)
        self.params = dict(dt=0.2, ro=1.1763, vn=1.583e-5, alpha=-10.0, beta=-1.0, cs0=0.1,
                           omega=1.0, pjuge=1e-4, nmaxp=50, sum_mode=SUM_SEQ,
                           refine_tol=0.0, refine_max=0, refine_omega=0.0)
        self.params.update(physics)
        self.scalars = OrcScalars()

    # -- data --
    def set(self, name, data):
        self.a[name][:] = np.asarray(data).reshape(-1).astype(self.dtype)

    def load_case(self, case):
        for n in fl.GRID + fl.MASKS + ["wind_profile"]:
            self.set(n, case[n])

    def view(self, name):
        return fl.FArray(name, self.ip, self.jp, self.kp, data=self.a[name])

    def copy(self, dtype=None):
        o = OracleState(self.ip, self.jp, self.kp, dtype or self.dtype, **self.params)
        for n, v in self.a.items():
            o.a[n][:] = v.astype(o.dtype)
        return o

    def _struct(self):
        s = OrcFields()
        s.ip, s.jp, s.kp = self.ip, self.jp, self.kp
        for n in _PTR_NAMES:
            setattr(s, n, self.a[n].ctypes.data)
        for q in range(9):
            s.diu[q] = self.a["diu%d" % (q + 1)].ctypes.data
            s.nou[q] = self.a["nou%d" % (q + 1)].ctypes.data
            s.cov[q] = self.a["cov%d" % (q + 1)].ctypes.data
        for n in _TAIL_NAMES:
            setattr(s, n, self.a["wind_profile" if n == "wind" else n].ctypes.data)
        return s

    def _params(self, **over):
        p = dict(self.params)
        p.update(over)
        q = OrcParams()
        for k in ["dt", "ro", "vn", "alpha", "beta", "cs0", "omega", "pjuge"]:
            # the FP32 oracle receives exactly the float32 value of each constant
            v = float(np.float32(p[k])) if self.dtype == np.float32 else float(p[k])
            setattr(q, k, v)
        q.nmaxp = int(p["nmaxp"])
        q.sum_mode = int(p["sum_mode"])
        q.refine_tol = float(p["refine_tol"])
        q.refine_max = int(p["refine_max"])
        # the correction blocks' omega as the FP32 path sees it
        ro = float(p["refine_omega"])
        q.refine_omega = float(np.float32(ro)) if self.dtype == np.float32 else ro
        return q

    def call(self, stage, n=0, **over):
        s = self._struct()
        q = self._params(**over)
        out = OrcScalars()
        r = getattr(lib(), "orc_" + stage + self.sfx)(ctypes.byref(s), ctypes.byref(q), n,
                                                      ctypes.byref(out))
        assert r == 0
        self.scalars = out
        return out

    def feedbfm(self, zbm):
        z = np.ascontiguousarray(zbm, dtype=self.dtype)
        s = self._struct()
        getattr(lib(), "orc_feedbfm" + self.sfx)(ctypes.byref(s), z.ctypes.data)

    def init(self):
        """init.f95:82-107 coefficients and les.f95:36 delx1 from the loaded grid."""
        self.call("init_coeffs")
        self.call("delx1")

    def ifdata(self, **over):
        return self.call("ifdata", 1, **over)

    def step(self, n, **over):
        return self.call("step", n, **over)

    def new_stats(self):
        """Zeroed aveflow statistics arrays (flat, i fastest)."""
        ip, jp, kp = self.ip, self.jp, self.kp
        return {n: np.zeros(ip * jp * kp if n in AVE3 else ip * kp, self.dtype) for n in AVE3 + AVE2}

    def aveflow(self, stats, n, n1, nmax):
        """aveflow.f95:42-157 on 'stats' (new_stats()) after step n."""
        st = OrcAve(*[stats[k].ctypes.data for k in AVE3 + AVE2])
        r = getattr(lib(), "orc_aveflow" + self.sfx)(ctypes.byref(self._struct()), ctypes.byref(st),
                                                   n, n1, nmax)
        assert r == 0
