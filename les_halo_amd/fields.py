"""Field table of the LES_Halo state: ids (les_field in include/les_hip.h) and the reference's
Fortran bounds (src/Fortran/main_ocl.f95:40-167), as functions of the interior sizes."""
import numpy as np

# name -> (id, bounds(ip,jp,kp) -> list of (lo, hi) per dimension)
_B3 = lambda a, b, c: (lambda ip, jp, kp: [(a[0], a[1](ip)), (b[0], b[1](jp)), (c[0], c[1](kp))])


def _r(lo, off):
    return (lo, (lambda n: n + off))


FIELDS = {
    "p": (0, _B3(_r(0, 2), _r(0, 2), _r(0, 1))),
    "u": (1, _B3(_r(0, 1), _r(-1, 1), _r(0, 1))),
    "v": (2, _B3(_r(0, 1), _r(-1, 1), _r(0, 1))),
    "w": (3, _B3(_r(0, 1), _r(-1, 1), _r(-1, 1))),
    "usum": (4, _B3(_r(0, 0), _r(0, 0), _r(0, 0))),
    "vsum": (5, _B3(_r(0, 0), _r(0, 0), _r(0, 0))),
    "wsum": (6, _B3(_r(0, 0), _r(0, 0), _r(0, 0))),
    "f": (7, _B3(_r(0, 0), _r(0, 0), _r(0, 0))),
    "g": (8, _B3(_r(0, 0), _r(0, 0), _r(0, 0))),
    "h": (9, _B3(_r(0, 0), _r(0, 0), _r(0, 0))),
    "fold": (10, _B3(_r(1, 0), _r(1, 0), _r(1, 0))),
    "gold": (11, _B3(_r(1, 0), _r(1, 0), _r(1, 0))),
    "hold": (12, _B3(_r(1, 0), _r(1, 0), _r(1, 0))),
    "rhs": (13, _B3(_r(0, 1), _r(0, 1), _r(0, 1))),
    "sm": (14, _B3(_r(-1, 1), _r(-1, 1), _r(0, 1))),
    "diu1": (15, _B3(_r(-1, 2), _r(0, 2), _r(0, 2))),
    "diu2": (16, _B3(_r(0, 2), _r(0, 2), _r(0, 2))),
    "diu3": (17, _B3(_r(0, 2), _r(0, 2), _r(0, 2))),
    "diu4": (18, _B3(_r(0, 2), _r(0, 2), _r(0, 2))),
    "diu5": (19, _B3(_r(-1, 2), _r(0, 2), _r(0, 2))),
    "diu6": (20, _B3(_r(0, 2), _r(0, 2), _r(0, 2))),
    "diu7": (21, _B3(_r(0, 2), _r(0, 2), _r(0, 2))),
    "diu8": (22, _B3(_r(0, 2), _r(0, 2), _r(0, 2))),
    "diu9": (23, _B3(_r(0, 2), _r(0, 2), _r(0, 2))),
    "amask1": (24, _B3(_r(0, 1), _r(0, 1), _r(0, 1))),
    "bmask1": (25, _B3(_r(-1, 1), _r(0, 1), _r(0, 1))),
    "cmask1": (26, _B3(_r(0, 1), _r(-1, 1), _r(0, 1))),
    "dmask1": (27, _B3(_r(0, 1), _r(0, 1), _r(0, 1))),
    "cn1": (28, _B3(_r(1, 0), _r(1, 0), _r(1, 0))),
    "cn2l": (29, lambda ip, jp, kp: [(1, ip)]),
    "cn2s": (30, lambda ip, jp, kp: [(1, ip)]),
    "cn3l": (31, lambda ip, jp, kp: [(1, jp)]),
    "cn3s": (32, lambda ip, jp, kp: [(1, jp)]),
    "cn4l": (33, lambda ip, jp, kp: [(1, kp)]),
    "cn4s": (34, lambda ip, jp, kp: [(1, kp)]),
    "dx1": (35, lambda ip, jp, kp: [(-1, ip + 1)]),
    "dxs": (36, lambda ip, jp, kp: [(0, ip)]),
    "dy1": (37, lambda ip, jp, kp: [(0, jp + 1)]),
    "dys": (38, lambda ip, jp, kp: [(0, jp)]),
    "dzn": (39, lambda ip, jp, kp: [(-1, kp + 2)]),
    "dzs": (40, lambda ip, jp, kp: [(-1, kp + 2)]),
    "z2": (41, lambda ip, jp, kp: [(1, kp + 2)]),
    "delx1": (42, lambda ip, jp, kp: [(1, kp)]),
    "wind_profile": (43, lambda ip, jp, kp: [(1, jp), (1, kp)]),
}
NFIELDS = 44

STATE_3D = ["p", "u", "v", "w", "usum", "vsum", "wsum", "f", "g", "h", "fold", "gold", "hold",
            "rhs", "sm"] + ["diu%d" % q for q in range(1, 10)]
MASKS = ["amask1", "bmask1", "cmask1", "dmask1"]
GRID = ["dx1", "dxs", "dy1", "dys", "dzn", "dzs", "z2"]
COEFFS = ["cn1", "cn2l", "cn2s", "cn3l", "cn3s", "cn4l", "cn4s", "delx1"]


def bounds(name, ip, jp, kp):
    return FIELDS[name][1](ip, jp, kp)


def field_id(name):
    return FIELDS[name][0]


def shape(name, ip, jp, kp):
    return tuple(hi - lo + 1 for lo, hi in bounds(name, ip, jp, kp))


def size(name, ip, jp, kp):
    return int(np.prod(shape(name, ip, jp, kp)))


class FArray:
    """A Fortran-bounded view over a flat column-major float array: a[i, j, k] uses Fortran
    indices.  .flat is the contiguous storage passed across the C ABI."""

    def __init__(self, name, ip, jp, kp, data=None, dtype=np.float32):
        self.name = name
        self.b = bounds(name, ip, jp, kp)
        shp = tuple(hi - lo + 1 for lo, hi in self.b)
        if data is None:
            data = np.zeros(int(np.prod(shp)), dtype=dtype)
        self.flat = data
        self.arr = data.reshape(shp, order="F")

    def _idx(self, key):
        if not isinstance(key, tuple):
            key = (key,)
        out = []
        for k, (lo, hi) in zip(key, self.b):
            if isinstance(k, slice):
                start = (k.start if k.start is not None else lo) - lo
                stop = (k.stop if k.stop is not None else hi + 1) - lo
                out.append(slice(start, stop, k.step))
            else:
                out.append(k - lo)
        return tuple(out)

    def __getitem__(self, key):
        return self.arr[self._idx(key)]

    def __setitem__(self, key, val):
        self.arr[self._idx(key)] = val

    def interior(self, ip, jp, kp):
        return self[1:ip + 1, 1:jp + 1, 1:kp + 1]
