"""The drop-in entry points in the reference host program's call pattern
(main_ocl.f95:198-253): the state after the init sequence is handed to
initialise_LES_kernel (TEMPL:18-30) once, then run_LES_kernel(n, nmax, wind_profile)
(TEMPL:219-230) runs each step, and the state is read back.

* A Fortran host (tests/dropin/dropin_main.f90) compiled with the library's Fortran binding
  (les_halo_amd/fortran/module_les_hip.f90), as the reference's main program would be;
* the same entry points through ctypes (les_initialise_LES_kernel, les_run_LES_kernel,
  les_finalise_LES_kernel), twice in one process (a second initialise replaces the default
  handle).

Both run the library's default configuration, the reference's press.f95 settings (omega 1,
early exit at 1e-4, 50 iterations), and must equal the oracle in exact-sum mode bit for bit.
"""
import ctypes
import os
import subprocess
import tempfile

import numpy as np
import pytest

import common
from les_halo_amd import fields as fl
from oracle.oracle import SUM_EXACT

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "dropin", "_build", "dropin_main")

# initialise_LES_kernel's array arguments, in order (TEMPL:18-30, external wind profile form)
INIT_ARGS = (["p", "u", "v", "w", "usum", "vsum", "wsum", "f", "g", "h", "fold", "gold", "hold"]
             + ["diu%d" % q for q in range(1, 10)]
             + ["amask1", "bmask1", "cmask1", "dmask1", "cn1", "cn2l", "cn2s", "cn3l", "cn3s",
                "cn4l", "cn4s", "rhs", "sm", "dxs", "dys", "dzs", "dx1", "dy1", "dzn",
                "wind_profile"])
OUT_ORDER = ["p", "u", "v", "w", "usum", "vsum", "wsum", "f", "g", "h", "fold", "gold", "hold",
             "rhs", "sm"]
DT = 0.05


def _default_physics(dims):
    """The library's default configuration (set.f95's constants, press.f95's SOR settings),
    which the drop-in entry points run, with the host's dt."""
    from les_halo_amd._lib import LesConfig, lib
    cfg = LesConfig()
    assert lib().les_default_config(ctypes.byref(cfg), *dims) == 0
    return dict(dt=DT, ro=cfg.ro, vn=cfg.vn, alpha=cfg.alpha, beta=cfg.beta, cs0=cfg.cs0,
                omega=cfg.omega, pjuge=cfg.pjuge, nmaxp=cfg.nmaxp)


def _after_init(dims, seed):
    """The host's state after set/grid/init/ifdata (main_ocl.f95:181-196), from the oracle."""
    case = common.tiny_case(*dims, seed=seed)
    o = common.oracle_from_case(case, **_default_physics(dims), sum_mode=SUM_EXACT)
    o.ifdata()
    return case, o


def test_fortran_host_binary_built():
    """build() compiled and linked the Fortran host against libles_hip.so (no GPU needed)."""
    if not os.path.exists(EXE):
        pytest.skip("tests/dropin/_build/dropin_main not built")
    out = subprocess.run(["ldd", EXE], capture_output=True, text=True).stdout
    assert "libles_hip.so" in out and "not found" not in out.split("libles_hip.so")[1].split("\n")[0]


@pytest.mark.gpu
@pytest.mark.parametrize("dims,seed,nsteps", [((16, 12, 10), 7, 6), ((40, 32, 20), 3, 4)])
def test_fortran_host(dims, seed, nsteps):
    if not os.path.exists(EXE):
        pytest.skip("tests/dropin/_build/dropin_main not built")
    ip, jp, kp = dims
    case, o = _after_init(dims, seed)
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(np.asarray([ip, jp, kp, nsteps], np.int32).tobytes())
            f.write(np.float32(DT).tobytes())
            for n in INIT_ARGS:
                f.write(np.ascontiguousarray(o.a[n], np.float32).tobytes())
        r = subprocess.run([EXE, fin, fout], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        raw = np.fromfile(fout, np.float32)
    for n in range(1, nsteps + 1):
        o.step(n)
    off, bad = 0, []
    for name in OUT_ORDER:
        cnt = fl.size(name, ip, jp, kp)
        got = raw[off:off + cnt]
        off += cnt
        if not common.bits_equal(got, o.a[name]):
            bad.append(common.mismatch_report(got, o.a[name], name))
    assert off == raw.size
    assert not bad, "\n".join(bad)


def _c_args(o):
    from les_halo_amd._lib import ptr
    arrs = [np.ascontiguousarray(o.a[n], np.float32) for n in INIT_ARGS]
    return arrs, [ptr(a) for a in arrs]


@pytest.mark.gpu
def test_ctypes_host_twice():
    from les_halo_amd._lib import lib, ptr
    L = lib()
    fn = L.les_initialise_LES_kernel
    assert len(fn.argtypes) == len(INIT_ARGS) + 4
    for dims, seed, nsteps in (((16, 12, 10), 5, 5), ((14, 10, 8), 6, 3)):
        ip, jp, kp = dims
        case, o = _after_init(dims, seed)
        arrs, ptrs = _c_args(o)
        assert fn(*ptrs, DT, ip, jp, kp) == 0
        wind = arrs[INIT_ARGS.index("wind_profile")]
        for n in range(1, nsteps + 1):
            assert L.les_run_LES_kernel(n, nsteps, ptr(wind)) == 0
            o.step(n)
        h = L.les_default_handle()
        assert h
        for name in ("u", "v", "w", "p", "f", "sm", "rhs"):
            got = np.empty(fl.size(name, ip, jp, kp), np.float32)
            assert L.les_download(ctypes.c_void_p(h), fl.field_id(name), ptr(got)) == 0
            assert common.bits_equal(got, o.a[name]), common.mismatch_report(got, o.a[name], name)
    assert L.les_finalise_LES_kernel() == 0
    assert not L.les_default_handle()
    assert L.les_run_LES_kernel(1, 1, None) != 0  # no default handle any more
    assert L.les_finalise_LES_kernel() == 0        # finalising twice is harmless
