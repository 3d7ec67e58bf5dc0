"""Host-side mirror of the reference's GPU host API for the LES_Halo time step.

`LES` owns one device domain through libles_hip.so and exposes the reference's kernel-level
names (run_LES_kernel's states, src/OpenCL/module_LES_combined_ocl_TEMPL_V2.f95:453-1058, and
the CPU subroutines of src/Fortran/main_ocl.f95:264-299).  Arrays cross the boundary in the
reference's Fortran bounds and layouts (les_halo_amd.fields).  Every call raises LESError on a
non-zero status, with the library's message.
"""
import ctypes

import numpy as np

from . import fields as fl
from ._lib import (ERR_NAMES, LES_PREC_FP32, LES_PREC_MIXED, LES_SUM_EXACT, LES_SUM_REFERENCE,
                   LesConfig, LesScalars, lib, ptr)


class LESError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s: %s" % (ERR_NAMES.get(code, code), msg))
        self.code = code


class LES:
    """One ip x jp x kp domain on the GPU."""

    def __init__(self, ip, jp, kp, dt=0.2, ro=1.1763, vn=None, alpha=-10.0, beta=-1.0, cs0=0.1,
                 omega=1.0, pjuge=1e-4, nmaxp=50, precision=LES_PREC_FP32, refine_tol=1e-7,
                 refine_max=8, device=0, refine_omega=0.0, refine_solver=0,
                 sum_order=LES_SUM_EXACT):
        L = lib()
        cfg = LesConfig()
        L.les_default_config(ctypes.byref(cfg), ip, jp, kp)
        cfg.dt, cfg.ro, cfg.alpha, cfg.beta, cfg.cs0 = dt, ro, alpha, beta, cs0
        if vn is not None:
            cfg.vn = vn
        cfg.omega, cfg.pjuge, cfg.nmaxp, cfg.precision = omega, pjuge, nmaxp, precision
        cfg.refine_tol, cfg.refine_max, cfg.device = refine_tol, refine_max, device
        cfg.refine_omega = refine_omega
        cfg.refine_solver = refine_solver
        cfg.sum_order = sum_order
        self.cfg = cfg
        self.dims = (ip, jp, kp)
        h = ctypes.c_void_p()
        r = L.les_create(ctypes.byref(cfg), ctypes.byref(h))
        if r != 0:
            raise LESError(r, "les_create failed (%dx%dx%d)" % (ip, jp, kp))
        self.h = h

    # ---- lifetime ----
    def close(self):
        if getattr(self, "h", None):
            lib().les_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def last_error(self):
        msg = lib().les_last_error(self.h)
        return msg.decode() if msg else ""

    def _chk(self, r, what):
        if r != 0:
            msg = lib().les_last_error(self.h)
            raise LESError(r, "%s: %s" % (what, msg.decode() if msg else ""))

    # ---- data ----
    def upload(self, name, data):
        ip, jp, kp = self.dims
        a = np.ascontiguousarray(np.asarray(data, dtype=np.float32).reshape(-1))
        if a.size != fl.size(name, ip, jp, kp):
            raise ValueError("%s: expected %d values, got %d" % (name, fl.size(name, ip, jp, kp), a.size))
        self._chk(lib().les_upload(self.h, fl.field_id(name), ptr(a)), "upload " + name)

    def download(self, name):
        ip, jp, kp = self.dims
        a = np.empty(fl.size(name, ip, jp, kp), np.float32)
        self._chk(lib().les_download(self.h, fl.field_id(name), ptr(a)), "download " + name)
        return a

    def device_ptr(self, name):
        p = ctypes.c_void_p()
        n = ctypes.c_int64()
        self._chk(lib().les_device_ptr(self.h, fl.field_id(name), ctypes.byref(p), ctypes.byref(n)),
                  "device_ptr")
        return p.value, n.value

    def load_case(self, case, coefficients=True):
        """Uploads grid, masks and inflow of a synthetic case; computes cn*/delx1."""
        for n in fl.GRID + fl.MASKS + ["wind_profile"]:
            self.upload(n, case[n])
        if coefficients:
            self.init_coefficients()

    AVE_NAMES = ["aveu", "avev", "avew", "avep", "avel", "aveuu", "avevv", "aveww", "avesm",
                 "avesmsm", "uwfx", "avesu", "avesv", "avesw", "avesuu", "avesvv", "avesww", "uwfxs"]

    def aveflow(self, n, n1, nmax):
        """aveflow.f95:42-157 after step n (les_aveflow)."""
        self._chk(lib().les_aveflow(self.h, n, n1, nmax), "aveflow")

    def aveflow_stats(self):
        """The eighteen statistics arrays (flat, i fastest)."""
        ip, jp, kp = self.dims
        out = {}
        for q, n in enumerate(self.AVE_NAMES):
            a = np.empty(ip * jp * kp if q < 11 else ip * kp, np.float32)
            self._chk(lib().les_aveflow_download(self.h, q, a.ctypes.data), "aveflow_download")
            out[n] = a
        return out

    def set_terrain(self, zbm):
        """feedbfm.f95:42-63 on the device from zbm(-1:ip+1,-1:jp+1) (z2, dzn uploaded)."""
        z = np.ascontiguousarray(zbm, dtype=np.float32)
        self._chk(lib().les_set_terrain(self.h, z.ctypes.data), "set_terrain")

    def init_coefficients(self):
        self._chk(lib().les_init_coefficients(self.h), "init_coefficients")

    def write_restart(self, data20, data21, n=0, time=0.0, with_header=False):
        """data20/data21 in the reference's unformatted records (press.f95:192-199,
        adam.f95:24-29); with_header makes data20 a data30 restart file."""
        self._chk(lib().les_write_restart(self.h, data20.encode(), data21.encode(), int(n),
                                          float(time), 1 if with_header else 0), "write_restart")

    def read_restart(self, data30, data31):
        """ifdata.f95:227-247 (ICAL = 1): interiors of u..wsum, fold..hold; returns (n, time)."""
        n, t = ctypes.c_int(), ctypes.c_float()
        self._chk(lib().les_read_restart(self.h, data30.encode(), data31.encode(), ctypes.byref(n),
                                         ctypes.byref(t)), "read_restart")
        return n.value, t.value

    def state(self, names=None):
        return {n: self.download(n) for n in (names or fl.STATE_3D)}

    def sync(self):
        self._chk(lib().les_sync(self.h), "sync")

    def scalars(self):
        s = LesScalars()
        self._chk(lib().les_get_scalars(self.h, ctypes.byref(s)), "get_scalars")
        return {k: getattr(s, k) for k, _ in LesScalars._fields_}

    # ---- reference stages (names of run_LES_kernel's states and of main_ocl's calls) ----
    def _call(self, name, *args):
        self._chk(getattr(lib(), "les_" + name)(self.h, *args), name)

    def velnw(self):
        self._call("velnw")

    def bondv1_init_uvw(self, n):
        self._call("bondv1_init_uvw", n)

    def velnw__bondv1_init_uvw(self, n):
        self._call("velnw__bondv1_init_uvw", n)

    def bondv1_calc_uout(self):
        self._call("bondv1_calc_uout")

    def bondv1_calc_uvw(self):
        self._call("bondv1_calc_uvw")

    def bondv1(self, n):
        """bondv1_ext_wind_profile.f95 as a whole."""
        self.bondv1_init_uvw(n)
        self.bondv1_calc_uout()
        self.bondv1_calc_uvw()

    def velfg(self):
        self._call("velfg")

    def feedbf(self):
        self._call("feedbf")

    def les_calc_sm(self):
        self._call("les_calc_sm")

    def velfg__feedbf__les_calc_sm(self):
        self._call("velfg__feedbf__les_calc_sm")

    def les_bound_sm(self):
        self._call("les_bound_sm")

    def les_calc_visc(self):
        self._call("les_calc_visc")

    def adam(self):
        self._call("adam")

    def les_calc_visc__adam(self):
        self._call("les_calc_visc__adam")

    def les(self):
        self._call("les")

    def bondfg(self):
        self._call("bondfg")

    def boundp1(self):
        self._call("boundp1")

    def boundp2(self):
        self._call("boundp2")

    def press_rhsav(self):
        self._call("press_rhsav")

    def press_sor(self, nrd):
        self._call("press_sor", nrd)

    def press_sor_solve(self):
        self._call("press_sor_solve")

    def press_pav(self):
        self._call("press_pav")

    def press_adj(self):
        self._call("press_adj")

    def press_boundp(self):
        self._call("press_boundp")

    def press(self):
        self._call("press")

    def step(self, n):
        self._call("step", n)

    def run(self, n0, nsteps):
        self._chk(lib().les_run(self.h, n0, nsteps), "run")

    def ifdata(self):
        self._call("ifdata")

    # ---- halo exchange ----
    def halo_size(self, names, h_w=1, correct=0):
        ip, jp, kp = self.dims
        ext = fl.shape(names[0], ip, jp, kp)
        return int(lib().les_halo_size(len(names), ext[0], ext[1], ext[2], h_w, correct))

    def exchange_halo_read(self, names, dev_buffer, h_w=1, correct=0):
        ids = (ctypes.c_int * len(names))(*[fl.field_id(n) for n in names])
        self._chk(lib().les_exchange_halo_read(self.h, ids, len(names), h_w, correct,
                                               ctypes.c_void_p(dev_buffer)), "exchange_halo_read")

    def exchange_halo_write(self, names, dev_buffer, h_w=1):
        ids = (ctypes.c_int * len(names))(*[fl.field_id(n) for n in names])
        self._chk(lib().les_exchange_halo_write(self.h, ids, len(names), h_w,
                                                ctypes.c_void_p(dev_buffer)), "exchange_halo_write")

    def halo_read_arrays(self, dev_arrays, im, jm, km, dev_buffer, h_w=1, correct=0):
        """exchange_<v_dim>_halo_read on caller-owned device arrays (addresses), enqueued on
        the handle's stream."""
        ptrs = (ctypes.c_void_p * len(dev_arrays))(*dev_arrays)
        self._chk(lib().les_halo_read_arrays(self.h, ptrs, len(dev_arrays), im, jm, km, h_w,
                                             correct, ctypes.c_void_p(dev_buffer)),
                  "halo_read_arrays")

    def halo_write_arrays(self, dev_arrays, im, jm, km, dev_buffer, h_w=1):
        """exchange_<v_dim>_halo_write on caller-owned device arrays (addresses)."""
        ptrs = (ctypes.c_void_p * len(dev_arrays))(*dev_arrays)
        self._chk(lib().les_halo_write_arrays(self.h, ptrs, len(dev_arrays), im, jm, km, h_w,
                                              ctypes.c_void_p(dev_buffer)), "halo_write_arrays")

    KERNELS = {"sor_sweep": 0, "velfg": 1, "visc_adam": 2, "calc_sm": 3, "rhs": 4, "velnw": 5,
               "step": 6, "sor_iter": 7}

    def autotune(self, reps=20):
        """Picks the fused SOR iteration's workgroup tile and the stencil kernels' workgroup
        shapes by timing every candidate on this grid (les_autotune); returns the choice."""
        from ._lib import LesTuning
        t = LesTuning()
        self._chk(lib().les_autotune(self.h, reps, ctypes.byref(t)), "autotune")
        names = ("velnw", "calc_sm", "visc_adam")
        return {"sor_tile_j": t.sor_tile_j, "sor_tile_k": t.sor_tile_k,
                "sor_iter_ms": float(t.sor_iter_ms),
                "stencil_shapes": {n: [t.stencil_bx[q], t.stencil_by[q]] for q, n in enumerate(names)},
                "stencil_ms": {n: float(t.stencil_ms[q]) for q, n in enumerate(names)}}

    def set_sor_tile(self, tile_j, tile_k=0):
        """SOR iteration form: tile_j = 0 two colour sweeps, else the fused launch's tile."""
        self._chk(lib().les_set_sor_tile(self.h, tile_j, tile_k), "set_sor_tile")

    def time_kernel(self, which, reps=20):
        """Average milliseconds per launch of one kernel (HIP events on the handle's stream)."""
        out = ctypes.c_float()
        self._chk(lib().les_time_kernel(self.h, self.KERNELS[which], reps, ctypes.byref(out)),
                  "time_kernel")
        return float(out.value)

    def reduce_sum_exact(self, dev_ptr, n):
        out = ctypes.c_float()
        self._chk(lib().les_reduce_sum_exact(self.h, ctypes.c_void_p(dev_ptr), n,
                                             ctypes.byref(out)), "reduce_sum_exact")
        return np.float32(out.value)


class Ensemble:
    """Independent LES members of any sizes advanced together (les_ensemble_*): batched (one
    launch per stage over every member's workgroups, in one captured graph per step) when the
    members allow it and batched=True, else one graph launch per member on its own stream,
    so the members run concurrently.  The members stay owned by the caller and must outlive
    the ensemble."""

    def __init__(self, members, batched=True):
        self.members = list(members)
        arr = (ctypes.c_void_p * len(self.members))(*[m.h for m in self.members])
        e = ctypes.c_void_p()
        r = lib().les_ensemble_create(arr, len(self.members), ctypes.byref(e))
        if r != 0:
            raise LESError(r, "les_ensemble_create failed")
        self.e = e
        lib().les_ensemble_set_mode(self.e, 1 if batched else 0)

    @property
    def batched(self):
        """True when the last run used the batched launches."""
        return bool(lib().les_ensemble_batched(self.e))

    def run(self, n0, nsteps):
        r = lib().les_ensemble_run(self.e, n0, nsteps)
        if r != 0:
            msgs = "; ".join(m.last_error() for m in self.members if m.last_error())
            raise LESError(r, "les_ensemble_run: %s" % msgs)

    def sync(self):
        r = lib().les_ensemble_sync(self.e)
        if r != 0:
            raise LESError(r, "les_ensemble_sync")

    def close(self):
        if getattr(self, "e", None):
            lib().les_ensemble_destroy(self.e)
            self.e = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["LES", "Ensemble", "LESError", "LES_PREC_FP32", "LES_PREC_MIXED", "LES_SUM_EXACT",
           "LES_SUM_REFERENCE"]


class TiledLES:
    """One ip x jp x kp domain as sub_i x sub_j tiles on the GPU (les_tiled_*): the same
    stages as LES, with face exchanges between the tiles; upload/download take whole-domain
    arrays."""

    def __init__(self, ip, jp, kp, sub_i=2, sub_j=2, dt=0.2, ro=1.1763, vn=None, alpha=-10.0,
                 beta=-1.0, cs0=0.1, omega=1.0, nmaxp=50, device=0):
        L = lib()
        cfg = LesConfig()
        L.les_default_config(ctypes.byref(cfg), ip, jp, kp)
        cfg.dt, cfg.ro, cfg.alpha, cfg.beta, cfg.cs0 = dt, ro, alpha, beta, cs0
        if vn is not None:
            cfg.vn = vn
        cfg.omega, cfg.pjuge, cfg.nmaxp, cfg.device = omega, 0.0, nmaxp, device
        cfg.sub_i, cfg.sub_j = sub_i, sub_j
        self.cfg = cfg
        self.dims = (ip, jp, kp)
        h = ctypes.c_void_p()
        r = L.les_tiled_create(ctypes.byref(cfg), ctypes.byref(h))
        if r != 0:
            raise LESError(r, "les_tiled_create failed (%dx%dx%d as %dx%d)" % (ip, jp, kp, sub_i, sub_j))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().les_tiled_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _chk(self, r, what):
        if r != 0:
            msg = lib().les_tiled_last_error(self.h)
            raise LESError(r, "%s: %s" % (what, msg.decode() if msg else ""))

    def upload(self, name, data):
        ip, jp, kp = self.dims
        a = np.ascontiguousarray(np.asarray(data, dtype=np.float32).reshape(-1))
        if a.size != fl.size(name, ip, jp, kp):
            raise ValueError("%s: expected %d values, got %d" % (name, fl.size(name, ip, jp, kp), a.size))
        self._chk(lib().les_tiled_upload(self.h, fl.field_id(name), ptr(a)), "upload " + name)

    def download(self, name):
        ip, jp, kp = self.dims
        a = np.empty(fl.size(name, ip, jp, kp), np.float32)
        self._chk(lib().les_tiled_download(self.h, fl.field_id(name), ptr(a)), "download " + name)
        return a

    def load_case(self, case, coefficients=True):
        for n in fl.GRID + fl.MASKS + ["wind_profile"]:
            self.upload(n, case[n])
        if coefficients:
            self._chk(lib().les_tiled_init_coefficients(self.h), "init_coefficients")

    def ifdata(self):
        self._chk(lib().les_tiled_ifdata(self.h), "ifdata")

    def step(self, n):
        self._chk(lib().les_tiled_step(self.h, int(n)), "step")

    def run(self, n0, nsteps):
        """Steps n0 .. n0+nsteps-1 (graph replays for n >= 2)."""
        self._chk(lib().les_tiled_run(self.h, int(n0), int(nsteps)), "run")

    def sync(self):
        self._chk(lib().les_tiled_sync(self.h), "sync")
