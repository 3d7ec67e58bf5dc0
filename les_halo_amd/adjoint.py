"""The adjoint form over the C ABI (les_adj_* in include/les_hip.h): gradients of a windowed
objective of the LES step with respect to the inflow wind profile and a per-cell drag
coefficient field, and a torch.autograd.Function exposing them.

    A = AdjointLES.from_les(L, fp64=True)      # state, grid, masks, coefficients of an LES
    J = A.window(nsteps, ckpt, probes)         # forward window with checkpoints, then reverse
    A.gradient("wind_profile"), A.gradient("cd")

The objective J is the mean of u over the probe points (i, j, k) and the window's steps, u
taken after each step.  The drag coefficient multiplies feedbf's beta term (beta*cd*u*mask,
feedbf.f95:53-55); cd = 1 is the reference's step.
"""
import ctypes

import numpy as np

from . import fields as fl
from ._lib import LesConfig, lib, ptr
from .les import LESError

LES_ADJ_CD = 100
# state the step carries (its adjoint is the gradient with respect to the window's start)
CARRIED = ["u", "v", "w", "p", "f", "g", "h", "fold", "gold", "hold", "usum", "vsum", "wsum"]
# everything the step reads: the carried state, grid, masks, coefficients, controls
INPUTS = CARRIED + ["rhs", "sm"] + ["diu%d" % q for q in range(1, 10)] + [
    "bmask1", "cmask1", "dmask1", "cn1", "cn2l", "cn2s", "cn3l", "cn3s", "cn4l", "cn4s", "dx1",
    "dxs", "dy1", "dys", "dzn", "dzs", "delx1", "wind_profile"]


def _fid(name):
    return LES_ADJ_CD if name == "cd" else fl.field_id(name)


def _size(name, dims):
    ip, jp, kp = dims
    return ip * jp * kp if name == "cd" else fl.size(name, ip, jp, kp)


def volume_sums(dx1, dy1, dzn, dims, fp64):
    """press.f95:67-74 / 125-132: area = pco = the sequential sum of dx1*dy1*dzn over the
    interior (FP32 in the reference's order, or FP64)."""
    ip, jp, kp = dims
    t = np.float64 if fp64 else np.float32
    dx = np.asarray(dx1, t)[2:ip + 2]   # dx1(1:ip), dx1 from -1
    dy = np.asarray(dy1, t)[1:jp + 1]   # dy1(1:jp), dy1 from 0
    dz = np.asarray(dzn, t)[2:kp + 2]   # dzn(1:kp), dzn from -1
    if fp64:
        return float(np.sum(dx[None, None, :] * dy[None, :, None] * dz[:, None, None]))
    area = np.float32(0.0)
    for k in range(kp):
        for j in range(jp):
            for i in range(ip):
                area = np.float32(area + np.float32(np.float32(dx[i] * dy[j]) * dz[k]))
    return float(area)


class AdjointLES:
    """One les_adj handle (FP32 or FP64)."""

    def __init__(self, ip, jp, kp, fp64=True, dt=0.05, ro=1.1763, vn=1.583e-5, alpha=-10.0,
                 beta=-1.0, cs0=0.1, omega=1.5, nmaxp=50, device=0):
        L = lib()
        cfg = LesConfig()
        L.les_default_config(ctypes.byref(cfg), ip, jp, kp)
        cfg.dt, cfg.ro, cfg.vn, cfg.alpha, cfg.beta, cfg.cs0 = dt, ro, vn, alpha, beta, cs0
        cfg.omega, cfg.nmaxp, cfg.device = omega, nmaxp, device
        self.dims = (ip, jp, kp)
        self.fp64 = bool(fp64)
        h = ctypes.c_void_p()
        r = L.les_adj_create(ctypes.byref(cfg), 1 if fp64 else 0, ctypes.byref(h))
        if r != 0:
            raise LESError(r, "les_adj_create failed (%dx%dx%d)" % (ip, jp, kp))
        self.h = h

    @classmethod
    def from_les(cls, L, fp64=True, area=None):
        """A handle holding the current state, grid, masks and coefficients of LES L (the
        state after some steps of the plain path, for instance); cd = 1.  area: the volume sum
        (default: L's, or the FP64 sum)."""
        c = L.cfg
        A = cls(*L.dims, fp64=fp64, dt=c.dt, ro=c.ro, vn=c.vn, alpha=c.alpha, beta=c.beta,
                cs0=c.cs0, omega=c.omega, nmaxp=c.nmaxp)
        for n in INPUTS:
            A.upload(n, L.download(n).astype(np.float64))
        if area is None and not fp64:
            area = L.scalars()["area"]  # the plain path's FP32 sum, once a press has run
        if not area:
            area = volume_sums(L.download("dx1"), L.download("dy1"), L.download("dzn"), L.dims, fp64)
        A.set_sums(area, area)
        return A

    def close(self):
        if getattr(self, "h", None):
            lib().les_adj_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, r, what):
        if r != 0:
            msg = lib().les_adj_last_error(self.h)
            raise LESError(r, "%s: %s" % (what, msg.decode() if msg else ""))

    def upload(self, name, data):
        a = np.ascontiguousarray(np.asarray(data, dtype=np.float64).reshape(-1))
        if a.size != _size(name, self.dims):
            raise ValueError("%s: expected %d values, got %d" % (name, _size(name, self.dims), a.size))
        self._chk(lib().les_adj_upload(self.h, _fid(name), ptr(a)), "upload " + name)

    def download(self, name):
        a = np.empty(_size(name, self.dims), np.float64)
        self._chk(lib().les_adj_download(self.h, _fid(name), ptr(a)), "download " + name)
        return a

    def upload_adjoint(self, name, data):
        a = np.ascontiguousarray(np.asarray(data, dtype=np.float64).reshape(-1))
        self._chk(lib().les_adj_upload_adjoint(self.h, _fid(name), ptr(a)), "upload_adjoint " + name)

    def gradient(self, name):
        """The adjoint of a carried field or control (after window: dJ/d(name))."""
        a = np.empty(_size(name, self.dims), np.float64)
        self._chk(lib().les_adj_download_adjoint(self.h, _fid(name), ptr(a)), "gradient " + name)
        return a

    def set_sums(self, area, pco):
        self._chk(lib().les_adj_set_sums(self.h, float(area), float(pco)), "set_sums")

    def zero_adjoints(self):
        self._chk(lib().les_adj_zero_adjoints(self.h), "zero_adjoints")

    def step(self, nsteps=1):
        self._chk(lib().les_adj_step(self.h, nsteps), "step")

    def reverse_step(self):
        """One step's vector-Jacobian product: the adjoints of the state after the step (as
        uploaded) become those of the current state; the controls' adjoints accumulate."""
        self._chk(lib().les_adj_reverse_step(self.h), "reverse_step")

    def set_slots(self, enable):
        """Whether window keeps per-step copies of the segment recomputation (default) or
        recomputes each step again before its transpose (less memory, the same gradient)."""
        self._chk(lib().les_adj_set_slots(self.h, 1 if enable else 0), "set_slots")

    def set_transposes(self, hand):
        """hand True (default): the hand-written fused transposes where they exist; False: every
        statement's generated transpose."""
        self._chk(lib().les_adj_set_transposes(self.h, 1 if hand else 0), "set_transposes")

    def window(self, nsteps, ckpt, probes):
        """Runs the window (checkpoint every ckpt steps) and its adjoint; returns J."""
        pr = np.ascontiguousarray(np.asarray(probes, dtype=np.int32).reshape(-1, 3))
        J = ctypes.c_double()
        self._chk(lib().les_adj_window(self.h, nsteps, ckpt, ptr(pr), pr.shape[0], ctypes.byref(J)),
                  "window")
        return J.value


def probe_points(case, n=16, zlo=5.0, zhi=30.0, seed=7):
    """n seeded interior probe cells (i, j, k), 1-based, whose centre height z2(k) - dzn(k)/2
    lies in [zlo, zhi] m and which are outside the buildings (amask1 = 1)."""
    ip, jp, kp = case["dims"]
    z2 = np.asarray(case["z2"], np.float64)[:kp]       # z2(1:kp), from 1
    dzn = np.asarray(case["dzn"], np.float64)[2:kp + 2]
    zc = z2 - 0.5 * dzn
    ks = [k + 1 for k in range(kp) if zlo <= zc[k] <= zhi] or [1]
    am = np.asarray(case["amask1"], np.float32).reshape(kp + 2, jp + 2, ip + 2)
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(1000 * n):
        if len(out) == n:
            break
        i, j, k = int(rng.integers(1, ip + 1)), int(rng.integers(1, jp + 1)), int(rng.choice(ks))
        if am[k, j, i] > 0 and (i, j, k) not in out:
            out.append((i, j, k))
    if len(out) < n:
        raise ValueError("fewer than %d probe cells in the height range" % n)
    return np.asarray(out, np.int32)


def _torch_function():
    import torch

    class WindowObjective(torch.autograd.Function):
        """J(wind_profile, cd) of an AdjointLES window: forward runs the window from the handle's
        state (restored before every call) and stores the gradient, backward scales it."""

        @staticmethod
        def forward(ctx, wind, cd, adj, state, nsteps, ckpt, probes):
            for n, a in state.items():
                adj.upload(n, a)
            adj.upload("wind_profile", wind.detach().cpu().double().numpy())
            adj.upload("cd", cd.detach().cpu().double().numpy())
            J = adj.window(nsteps, ckpt, probes)
            # gradients and J live where the inputs live (host or device tensors)
            gw = torch.from_numpy(adj.gradient("wind_profile")).to(
                device=wind.device, dtype=wind.dtype).reshape(wind.shape)
            gc = torch.from_numpy(adj.gradient("cd")).to(
                device=cd.device, dtype=cd.dtype).reshape(cd.shape)
            ctx.save_for_backward(gw, gc)
            return torch.tensor(J, dtype=wind.dtype, device=wind.device)

        @staticmethod
        def backward(ctx, gJ):
            gw, gc = ctx.saved_tensors
            return gJ * gw, gJ * gc, None, None, None, None, None

    return WindowObjective


def window_objective(wind, cd, adj, nsteps, ckpt, probes, state):
    """torch.autograd entry: J as a differentiable function of the wind profile (jp x kp
    tensor) and the drag field (ip*jp*kp tensor, k slowest); state: the carried fields
    (name -> array) the window starts from."""
    return _torch_function().apply(wind, cd, adj, state, nsteps, ckpt, probes)
