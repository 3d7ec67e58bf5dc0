"""Shared helpers of the test-suite: seeded states, bitwise comparison, the oracle chain."""
import hashlib

import numpy as np

from les_halo_amd import fields as fl
from les_halo_amd import inputs

# reference constants used by the golden vectors (set.f95:78-82, press.f95:44-46)
REF_PHYS = dict(dt=0.05, ro=1.1763, vn=1.583e-5, alpha=-10.0, beta=-1.0)
REF_SOR = dict(omega=1.0, pjuge=1e-4, nmaxp=50)
# the benchmark configuration: 50 SOR iterations at omega 1.5, no early exit
CFG_SOR = dict(omega=1.5, pjuge=0.0, nmaxp=50)

DUMP_ORDER = (["p", "u", "v", "w", "usum", "vsum", "wsum", "f", "g", "h", "fold", "gold", "hold",
               "rhs", "sm"] + ["diu%d" % q for q in range(1, 10)]
              + ["cn1", "cn2l", "cn2s", "cn3l", "cn3s", "cn4l", "cn4s", "delx1"])
RANDOM_FIELDS = {"p": 10.0, "u": 3.0, "v": 1.0, "w": 1.0, "usum": 1.0, "vsum": 1.0, "wsum": 1.0,
                 "f": 1.0, "g": 1.0, "h": 1.0, "fold": 1.0, "gold": 1.0, "hold": 1.0, "rhs": 1.0,
                 "sm": 0.1, "diu1": 0.5, "diu2": 0.5, "diu3": 0.5, "diu4": 0.5, "diu5": 0.5,
                 "diu6": 0.5, "diu7": 0.5, "diu8": 0.5, "diu9": 0.5}


def splitmix_uniform(seed, n):
    """n uniforms in [-1, 1) from SplitMix64 (vectorised, platform independent)."""
    with np.errstate(over="ignore"):
        idx = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
    return (2.0 * u - 1.0).astype(np.float32)


def random_state(ip, jp, kp, seed):
    """A full state of seeded random values (every array including halos)."""
    st = {}
    for q, (n, scale) in enumerate(RANDOM_FIELDS.items()):
        st[n] = (np.float32(scale) * splitmix_uniform(seed * 1000 + q, fl.size(n, ip, jp, kp))).astype(np.float32)
    return st


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float32).tobytes()).hexdigest()[:32]


def bits_equal(a, b):
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def mismatch_report(a, b, name):
    a = np.ascontiguousarray(a, np.float32).ravel()
    b = np.ascontiguousarray(b, np.float32).ravel()
    bad = a.view(np.uint32) != b.view(np.uint32)
    idx = np.flatnonzero(bad)[:5]
    return "%s: %d/%d differ, first flat idx %s, a=%s b=%s, max|d|=%g" % (
        name, int(bad.sum()), a.size, idx.tolist(), a[idx].tolist(), b[idx].tolist(),
        float(np.max(np.abs(a.astype(np.float64) - b.astype(np.float64)))) if a.size else 0.0)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (d if d > 0 else 1.0))


def oracle_from_case(case, dtype=np.float32, **params):
    from oracle.oracle import OracleState
    ip, jp, kp = case["dims"]
    o = OracleState(ip, jp, kp, dtype=dtype, **params)
    o.load_case(case)
    o.init()
    return o


def interior(name, flat, ip, jp, kp):
    return fl.FArray(name, ip, jp, kp, data=np.asarray(flat)).interior(ip, jp, kp)


def tiny_case(ip=16, jp=12, kp=10, seed=7):
    # lower buildings on a tiny grid so that some columns are solid at several levels
    return inputs.make_case(ip, jp, kp, seed=seed, hmax=12.0, coverage=0.3)
