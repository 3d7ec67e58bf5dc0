"""Synthetic inputs of the benchmark configurations, through the library's deterministic
host generators (les_gen_* in include/les_hip.h)."""
import numpy as np

from . import fields as fl
from ._lib import lib, ptr


def gen_grid(ip, jp, kp, dx=4.0, dy=4.0, dz0=2.0, z_flat=40.0, ratio=1.05):
    g = {n: np.zeros(fl.size(n, ip, jp, kp), np.float32) for n in fl.GRID}
    r = lib().les_gen_grid(ip, jp, kp, dx, dy, dz0, z_flat, ratio, ptr(g["dx1"]), ptr(g["dxs"]),
                           ptr(g["dy1"]), ptr(g["dys"]), ptr(g["dzn"]), ptr(g["dzs"]), ptr(g["z2"]))
    assert r == 0, r
    return g


def gen_terrain(ip, jp, seed, dx=4.0, dy=4.0, coverage=0.30, wmin=8.0, wmax=40.0, hmax=60.0):
    zbm = np.zeros((ip + 3) * (jp + 3), np.float32)
    r = lib().les_gen_terrain(ip, jp, dx, dy, seed, coverage, wmin, wmax, hmax, ptr(zbm))
    assert r == 0, r
    return zbm


def gen_masks(ip, jp, kp, zbm, grid):
    m = {n: np.zeros(fl.size(n, ip, jp, kp), np.float32) for n in fl.MASKS}
    r = lib().les_gen_masks(ip, jp, kp, ptr(zbm), ptr(grid["z2"]), ptr(grid["dzn"]), ptr(m["amask1"]),
                            ptr(m["bmask1"]), ptr(m["cmask1"]), ptr(m["dmask1"]))
    assert r == 0, r
    return m


def gen_inflow(jp, kp, grid, ustar=0.2, z0=0.1, noise=0.01, seed=0x1F10):
    wind = np.zeros(jp * kp, np.float32)
    r = lib().les_gen_inflow(jp, kp, ptr(grid["z2"]), ptr(grid["dzn"]), ustar, z0, noise, seed,
                             ptr(wind))
    assert r == 0, r
    return wind


def make_case(ip, jp, kp, seed=0x4C455331, ustar=0.2, dx=4.0, dy=4.0, dz0=2.0, z_flat=40.0,
              ratio=1.05, coverage=0.30, hmax=60.0, noise=0.01):
    """Grid, terrain masks and inflow of one configuration (seeded)."""
    grid = gen_grid(ip, jp, kp, dx, dy, dz0, z_flat, ratio)
    zbm = gen_terrain(ip, jp, seed, dx, dy, coverage, 8.0, 40.0, hmax)
    masks = gen_masks(ip, jp, kp, zbm, grid)
    wind = gen_inflow(jp, kp, grid, ustar=ustar, noise=noise, seed=seed ^ 0x5EED)
    case = dict(grid)
    case.update(masks)
    case["wind_profile"] = wind
    case["zbm"] = zbm
    case["dims"] = (ip, jp, kp)
    return case


# Physics of the benchmark configurations (dt, Smagorinsky Cs, SOR) and the reference's
# constants (set.f95:78-82)
CONFIG_PHYSICS = dict(dt=0.05, ro=1.1763, vn=1.583e-5, alpha=-10.0, beta=-1.0, cs0=0.1,
                      omega=1.5, pjuge=0.0, nmaxp=50)


def ensemble_specs(n=64, seed=0x454E5331):
    """Member sizes and inflow friction velocities of the batched-ensemble configuration: ip
    and jp uniform over {96, 128, ..., 320}, kp over {60, 90}, u* uniform in [0.1, 0.5] m/s,
    one terrain seed per member (numpy PCG64 stream from 'seed')."""
    rng = np.random.default_rng(seed)
    sizes = np.arange(96, 321, 32)
    out = []
    for q in range(n):
        ip, jp = (int(x) for x in rng.choice(sizes, 2))
        kp = int(rng.choice([60, 90]))
        out.append(dict(dims=(ip, jp, kp), seed=int(rng.integers(1, 2 ** 63)),
                        ustar=float(rng.uniform(0.1, 0.5))))
    return out


def gis_zbm(path, ipmax, jpmax):
    """zbm(-1:ipmax+1,-1:jpmax+1), zero except the 100 x 100 block of a GIS height file placed
    as feedbfm.f95:31-37 does (les_gis_read_zbm)."""
    zbm = np.zeros((ipmax + 3) * (jpmax + 3), np.float32)
    r = lib().les_gis_read_zbm(path.encode(), ipmax, jpmax, zbm.ctypes.data)
    if r != 0:
        raise ValueError("les_gis_read_zbm(%s) failed (%d)" % (path, r))
    return zbm
