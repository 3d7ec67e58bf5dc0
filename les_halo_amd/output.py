"""Snapshot output of the reference time loop through the C ABI (les_output.cpp): the NetCDF
files of module_LES_write_netcdf.f95 and the anime files of anime.f95."""
import ctypes

import numpy as np

from . import fields as fl
from ._lib import lib


def _arr(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float32).reshape(-1))


class NetCDFOutput:
    """init_netcdf_file / write_to_netcdf_file / close_netcdf_file
    (module_LES_write_netcdf.f95:111-386): six classic-format files under `directory`."""

    def __init__(self, directory, ip, jp, kp, ntimes=20):
        self.dims = (ip, jp, kp)
        h = ctypes.c_void_p()
        r = lib().les_netcdf_create(str(directory).encode(), ip, jp, kp, ntimes, ctypes.byref(h))
        if r != 0:
            raise RuntimeError("les_netcdf_create failed (%d)" % r)
        self.h = h

    def write(self, les, t):
        """Time level t (1-based) of p, u, v, w, usum, vsum, wsum from an LES handle."""
        r = lib().les_netcdf_write(self.h, les.h, int(t))
        if r != 0:
            raise RuntimeError("les_netcdf_write failed (%d)" % r)

    def write_arrays(self, t, p, u, v, w, usum, vsum, wsum):
        """The same from host arrays in the Fortran bounds."""
        ip, jp, kp = self.dims
        a = [_arr(x) for x in (p, u, v, w, usum, vsum, wsum)]
        for x, nm in zip(a, ("p", "u", "v", "w", "usum", "vsum", "wsum")):
            if x.size != fl.size(nm, ip, jp, kp):
                raise ValueError("%s has %d elements, expected %d" % (nm, x.size, fl.size(nm, ip, jp, kp)))
        r = lib().les_netcdf_write_arrays(self.h, int(t), *[x.ctypes.data for x in a])
        if r != 0:
            raise RuntimeError("les_netcdf_write_arrays failed (%d)" % r)

    def close(self):
        if self.h:
            lib().les_netcdf_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def anime(les, n, n0, nmax, data22, data23):
    """anime.f95:26-49 from an LES handle (writes only at n == n0, n == nmax or n % 1000 == 0)."""
    r = lib().les_anime(les.h, int(n), int(n0), int(nmax), str(data22).encode(), str(data23).encode())
    if r != 0:
        raise RuntimeError("les_anime failed (%d)" % r)


def anime_arrays(n, n0, nmax, ip, jp, kp, dx1, dy1, z2, u, v, w, amask1, data22, data23):
    """anime.f95:26-49 from host arrays in the Fortran bounds."""
    a = [_arr(x) for x in (dx1, dy1, z2, u, v, w, amask1)]
    for x, nm in zip(a, ("dx1", "dy1", "z2", "u", "v", "w", "amask1")):
        if x.size != fl.size(nm, ip, jp, kp):
            raise ValueError("%s has %d elements, expected %d" % (nm, x.size, fl.size(nm, ip, jp, kp)))
    r = lib().les_anime_arrays(int(n), int(n0), int(nmax), ip, jp, kp, *[x.ctypes.data for x in a],
                               str(data22).encode(), str(data23).encode())
    if r != 0:
        raise RuntimeError("les_anime_arrays failed (%d)" % r)
