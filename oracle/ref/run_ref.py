"""Runs the reference Fortran physics (built by oracle/ref/build_ref.sh into oracle/_ref/) on
inputs written in ref_harness.f95's raw stream format - test infrastructure only, used in this
container to produce the golden vectors under tests/golden/."""
import os
import subprocess
import tempfile

import numpy as np

from les_halo_amd import fields as fl

HERE = os.path.dirname(os.path.abspath(__file__))
REF_OUT = os.path.join(os.path.dirname(HERE), "_ref")
REF_SRC = os.environ.get("LES_REFERENCE_SRC", "/root/reference/src/Fortran")

DUMP_ORDER = (["p", "u", "v", "w", "usum", "vsum", "wsum", "f", "g", "h", "fold", "gold", "hold",
               "rhs", "sm"] + ["diu%d" % q for q in range(1, 10)]
              + ["cn1", "cn2l", "cn2s", "cn3l", "cn3s", "cn4l", "cn4s", "delx1"])


def tag_of(ip, jp, kp, cfg=False):
    """Build tag of a grid; cfg selects the build with the benchmark's SOR parameters
    (omega 1.5, no early exit, 50 iterations) in place of press.f95:44-46's."""
    if (ip, jp, kp) == (150, 150, 90):
        tag = "ref150"
    elif (ip, jp, kp) == (300, 300, 90):
        tag = "ref300"
    else:
        tag = "%dx%dx%d" % (ip, jp, kp)
    return tag + ("cfg" if cfg else "")


def available():
    return os.path.isdir(REF_SRC)


def binary(ip, jp, kp, cfg=False):
    """The reference build for this grid; built on demand where the reference sources exist
    (this container), else whatever was built before (the GPU box gets the built files)."""
    tag = tag_of(ip, jp, kp, cfg)
    exe = os.path.join(REF_OUT, tag, "ref_harness")
    recipe = [os.path.join(HERE, "ref_harness.f95"), os.path.join(HERE, "build_ref.sh")]
    if not available():
        if not os.path.exists(exe):
            raise FileNotFoundError("reference build %s absent and no reference sources" % exe)
        return exe
    if not os.path.exists(exe) or max(map(os.path.getmtime, recipe)) > os.path.getmtime(exe):
        subprocess.run([os.path.join(HERE, "build_ref.sh"), tag], check=True,
                       stdout=subprocess.DEVNULL)
    return exe


def _f32(x):
    return np.asarray(x, dtype=np.float32).reshape(-1).tobytes()


def _i32(*xs):
    return np.asarray(xs, dtype=np.int32).tobytes()


def _header(mode, ip, jp, kp, phys, case):
    b = _i32(mode, ip, jp, kp)
    b += _f32([phys["dt"], phys["ro"], phys["vn"], phys["alpha"], phys["beta"]])
    b += b"".join(_f32(case[n]) for n in ["dx1", "dy1", "dzn", "dzs", "dxs", "dys", "z2"])
    b += b"".join(_f32(case[n]) for n in ["amask1", "bmask1", "cmask1", "dmask1", "wind_profile"])
    return b


def _parse(raw, ip, jp, kp, ndumps):
    sizes = [fl.size(n, ip, jp, kp) for n in DUMP_ORDER]
    per = sum(sizes)
    arr = np.frombuffer(raw, dtype=np.float32)
    assert arr.size == per * ndumps, (arr.size, per, ndumps)
    out = []
    off = 0
    for _ in range(ndumps):
        d = {}
        for n, s in zip(DUMP_ORDER, sizes):
            d[n] = arr[off:off + s].copy()
            off += s
        out.append(d)
    return out


def _run(exe, payload, cwd=None, want_stdout=False):
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(payload)
        r = subprocess.run([exe, fin, fout], check=True, cwd=cwd, stdout=subprocess.PIPE, text=True)
        with open(fout, "rb") as f:
            raw = f.read()
        return (raw, r.stdout) if want_stdout else raw


def timed_steps(exe, case, phys, nsteps):
    """Runs init + nsteps steps of the reference without dumps; returns the seconds the
    reference's time loop took (system_clock around the loop, ref_harness.f95)."""
    ip, jp, kp = case["dims"]
    payload = _header(0, ip, jp, kp, phys, case) + _i32(nsteps, 0)
    _, out = _run(exe, payload, want_stdout=True)
    for line in out.splitlines():
        if "steps_seconds=" in line:
            return float(line.split("=")[1])
    raise RuntimeError("no timing in reference output: %r" % out[-500:])


def trajectory(case, phys, nsteps, dump_steps, cfg=False, names=None):
    """Mode 0: init + ifdata sequence + nsteps steps; returns one state dict per dump step
    (only the arrays in names, when given).  cfg: the build with the benchmark's SOR
    parameters."""
    ip, jp, kp = case["dims"]
    payload = _header(0, ip, jp, kp, phys, case) + _i32(nsteps, len(dump_steps)) + _i32(*dump_steps)
    out = _parse(_run(binary(ip, jp, kp, cfg), payload), ip, jp, kp, len(dump_steps))
    if names is not None:
        out = [{n: d[n] for n in names} for d in out]
    return out


AVE3 = ["aveu", "avev", "avew", "avep", "avel", "aveuu", "avevv", "aveww", "avesm", "avesmsm",
        "uwfx"]
AVE2 = ["avesu", "avesv", "avesw", "avesuu", "avesvv", "avesww", "uwfxs"]


def aveflow_stats(case, phys, nsteps, n1):
    """Mode 3: init sequence, nsteps steps with aveflow.f95 after each (nmax = nsteps);
    returns the statistics arrays (flat, i fastest)."""
    ip, jp, kp = case["dims"]
    payload = _header(3, ip, jp, kp, phys, case) + _i32(nsteps, 0) + _i32(n1)
    raw = np.frombuffer(_run(binary(ip, jp, kp), payload), np.float32)
    out, pos = {}, 0
    for n in AVE3 + AVE2:
        cnt = ip * jp * kp if n in AVE3 else ip * kp
        out[n] = raw[pos:pos + cnt].copy()
        pos += cnt
    assert pos == raw.size, (pos, raw.size)
    return out


def stage(case, phys, state, stage_id, n):
    """Mode 1: one reference stage on a full state (dict of flat float32 arrays)."""
    ip, jp, kp = case["dims"]
    payload = _header(1, ip, jp, kp, phys, case) + _i32(stage_id, n)
    payload += b"".join(_f32(state[nm]) for nm in DUMP_ORDER)
    return _parse(_run(binary(ip, jp, kp), payload), ip, jp, kp, 1)[0]


STAGE_IDS = {"velnw": 1, "bondv1": 2, "velfg": 3, "feedbf": 4, "les": 5, "adam": 6, "press": 7,
             "boundp1": 8, "boundp2": 9, "bondfg": 10, "boundsm": 11, "step": 12, "anime": 13}


def anime_files(case, phys, state, n):
    """Mode 1, stage 13: anime.f95 on a full state; returns the bytes of data22 and data23."""
    ip, jp, kp = case["dims"]
    payload = _header(1, ip, jp, kp, phys, case) + _i32(STAGE_IDS["anime"], n)
    payload += b"".join(_f32(state[nm]) for nm in DUMP_ORDER)
    with tempfile.TemporaryDirectory() as td:
        _run(binary(ip, jp, kp), payload, cwd=td)
        with open(os.path.join(td, "anime22.bin"), "rb") as f:
            d22 = f.read()
        with open(os.path.join(td, "anime23.bin"), "rb") as f:
            d23 = f.read()
    return d22, d23


def tokyo_masks(ip, jp, kp, dzn, z2, zbm):
    """Mode 2: feedbfm.f95 with its own GIS read (needs the 150x150 reference build)."""
    b = _i32(2, ip, jp, kp) + _f32(dzn) + _f32(z2) + _f32(zbm)
    raw = _run(binary(ip, jp, kp), b, cwd=REF_SRC)
    arr = np.frombuffer(raw, dtype=np.float32)
    sizes = [(ip + 3) * (jp + 3)] + [fl.size(n, ip, jp, kp) for n in fl.MASKS]
    out, off = {}, 0
    for n, s in zip(["zbm"] + fl.MASKS, sizes):
        out[n] = arr[off:off + s].copy()
        off += s
    return out
