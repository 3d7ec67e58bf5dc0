"""Writes the golden vectors of the reference CPU physics into tests/golden/.

Runs ONLY in the build container, where the reference sources exist: it builds the reference
Fortran with oracle/ref/build_ref.sh and runs it through oracle/ref/ref_harness.f95 on seeded
inputs.  The fixtures hold SHA-256 digests and L2 norms of every output array (plus a few
sampled values and the terrain of the reference's GIS file), so that the tests can pin the
oracle bit for bit without the reference present.

    python tests/golden/make_ref_vectors.py [--big]
    python tests/golden/make_ref_vectors.py --cfg   (only the benchmark-SOR entries, *_cfg)

The --cfg entries come from the reference built with the benchmark's SOR parameters (omega 1.5,
no early exit, 50 iterations; build_ref.sh <tag>cfg edits only press.f95:44-46): a 16x12x10
trajectory, the odd-jp grid, 150x150x90 after 1, 5 and 10 steps, and u, v, w, p of 300x300x90
after 10 and 100 steps (the grid of the benchmark, the comparison the GPU tests repeat).
"""
import argparse
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from les_halo_amd import inputs  # noqa: E402
from oracle.oracle import OracleState  # noqa: E402
from oracle.ref import run_ref  # noqa: E402
import common  # noqa: E402


def digest(state):
    return {n: [common.sha(state[n]), float(np.linalg.norm(state[n].astype(np.float64)))]
            for n in run_ref.DUMP_ORDER}


def full_state_for_stage(case, seed):
    ip, jp, kp = case["dims"]
    st = common.random_state(ip, jp, kp, seed)
    o = OracleState(ip, jp, kp)
    o.load_case(case)
    o.init()
    for n in ["cn1", "cn2l", "cn2s", "cn3l", "cn3s", "cn4l", "cn4s", "delx1"]:
        st[n] = o.a[n].copy()
    return st


STAGE_CASES = [("velnw", 0), ("bondv1", 1), ("bondv1", 3), ("velfg", 0), ("feedbf", 0),
               ("les", 0), ("adam", 0), ("press", 0), ("boundp1", 0), ("boundp2", 0),
               ("bondfg", 0), ("boundsm", 0), ("step", 1), ("step", 4)]


def cfg_vectors(out):
    """The *_cfg entries (reference built with omega 1.5, pjuge 0, nmaxp 50)."""
    phys = dict(common.REF_PHYS)
    case = common.tiny_case()
    dumps = [0, 1, 2, 5, 12]
    traj = run_ref.trajectory(case, phys, 12, dumps, cfg=True)
    out["tiny_trajectory_cfg"] = {"dims": list(case["dims"]), "seed": 7, "hmax": 12.0,
                                  "coverage": 0.3, "sor": common.CFG_SOR,
                                  "dumps": {str(n): digest(s) for n, s in zip(dumps, traj)}}
    print("tiny cfg trajectory done", flush=True)
    case_odd = common.tiny_case(14, 11, 8, seed=9)
    traj = run_ref.trajectory(case_odd, phys, 3, [0, 3], cfg=True)
    out["odd_trajectory_cfg"] = {"dims": list(case_odd["dims"]), "seed": 9, "hmax": 12.0,
                                 "coverage": 0.3, "sor": common.CFG_SOR,
                                 "dumps": {str(n): digest(s) for n, s in zip([0, 3], traj)}}
    print("odd cfg trajectory done", flush=True)
    keep = ("u", "v", "w", "p", "sm", "rhs", "usum", "f")
    ip, jp, kp = 150, 150, 90
    case = inputs.make_case(ip, jp, kp)
    dumps = [1, 5, 10]
    traj = run_ref.trajectory(case, phys, 10, dumps, cfg=True, names=keep)
    out["ref150_trajectory_cfg"] = {
        "dims": [ip, jp, kp], "seed": 0x4C455331, "sor": common.CFG_SOR,
        "dumps": {str(n): {k: [common.sha(s[k]), float(np.linalg.norm(s[k].astype(np.float64)))]
                           for k in keep} for n, s in zip(dumps, traj)}}
    print("150 cfg trajectory done", flush=True)
    ip, jp, kp = 300, 300, 90
    case = inputs.make_case(ip, jp, kp)
    dumps = [10, 100]
    keep = ("u", "v", "w", "p")
    traj = run_ref.trajectory(case, phys, 100, dumps, cfg=True, names=keep)
    out["ref300_trajectory_cfg"] = {
        "dims": [ip, jp, kp], "seed": 0x4C455331, "sor": common.CFG_SOR,
        "dumps": {str(n): {k: [common.sha(s[k]), float(np.linalg.norm(s[k].astype(np.float64)))]
                           for k in keep} for n, s in zip(dumps, traj)}}
    print("300 cfg trajectory done", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--big", action="store_true", help="also the 150x150x90 reference trajectory")
    ap.add_argument("--cfg", action="store_true",
                    help="only (re)write the entries of the benchmark-SOR reference build")
    args = ap.parse_args()
    if args.cfg:
        path = os.path.join(HERE, "ref_vectors.json")
        out = json.load(open(path))
        cfg_vectors(out)
        with open(path, "w") as f:
            json.dump(out, f, indent=1, sort_keys=True)
        print("wrote", path)
        return
    phys = dict(common.REF_PHYS)
    out = {"generator": "tests/golden/make_ref_vectors.py",
           "reference_build": "oracle/ref/build_ref.sh (flang -O2 -ffp-contract=off)",
           "phys": phys, "sor": common.REF_SOR}

    # 1. tiny trajectory (init + 12 steps)
    case = common.tiny_case()
    dumps = [0, 1, 2, 5, 12]
    traj = run_ref.trajectory(case, phys, 12, dumps)
    out["tiny_trajectory"] = {"dims": list(case["dims"]), "seed": 7, "hmax": 12.0,
                              "coverage": 0.3, "dumps": {str(n): digest(s) for n, s in zip(dumps, traj)}}
    print("tiny trajectory done")

    # 2. single stages on a seeded random state
    stages = {}
    for name, n in STAGE_CASES:
        seed = 11 + len(stages)
        st = full_state_for_stage(case, seed)
        res = run_ref.stage(case, phys, st, run_ref.STAGE_IDS[name], n)
        stages["%s_n%d" % (name, n)] = {"stage": name, "n": n, "seed": seed, "out": digest(res)}
    out["tiny_stages"] = {"dims": list(case["dims"]), "seed": 7, "stages": stages}
    print("stages done")

    # 3. an odd-jm grid (red-black colouring across the periodic seam differs)
    case_odd = common.tiny_case(14, 11, 8, seed=9)
    dumps_odd = [0, 3]
    traj = run_ref.trajectory(case_odd, phys, 3, dumps_odd)
    out["odd_trajectory"] = {"dims": list(case_odd["dims"]), "seed": 9, "hmax": 12.0,
                             "coverage": 0.3,
                             "dumps": {str(n): digest(s) for n, s in zip(dumps_odd, traj)}}
    print("odd trajectory done")

    # 4. aveflow statistics over a short run (aveflow.f95 after every step)
    st = run_ref.aveflow_stats(case, phys, 6, 3)
    out["aveflow"] = {"dims": list(case["dims"]), "seed": 7, "hmax": 12.0, "coverage": 0.3,
                      "nsteps": 6, "n1": 3,
                      "stats": {n: [common.sha(v), float(np.linalg.norm(v.astype(np.float64)))]
                                for n, v in st.items()}}
    print("aveflow done")

    # 5. the reference's own grid size and GIS terrain
    if args.big:
        ip, jp, kp = 150, 150, 90
        case = inputs.make_case(ip, jp, kp)
        dumps = [1, 10]
        traj = run_ref.trajectory(case, phys, 10, dumps)
        out["ref150_trajectory"] = {"dims": [ip, jp, kp], "seed": 0x4C455331,
                                    "dumps": {str(n): digest(s) for n, s in zip(dumps, traj)}}
        print("150 trajectory done")
        grid = inputs.gen_grid(ip, jp, kp)
        tk = run_ref.tokyo_masks(ip, jp, kp, grid["dzn"], grid["z2"],
                                 np.zeros((ip + 3) * (jp + 3), np.float32))
        np.savez_compressed(os.path.join(HERE, "tokyo_150x150x90.npz"), zbm=tk["zbm"],
                            dzn=grid["dzn"], z2=grid["z2"],
                            **{n: np.packbits(tk[n].astype(np.uint8)) for n in
                               ["amask1", "bmask1", "cmask1", "dmask1"]})
        out["tokyo_masks"] = {n: common.sha(tk[n]) for n in ["zbm", "amask1", "bmask1", "cmask1",
                                                              "dmask1"]}
        print("tokyo masks done")
    else:
        old = os.path.join(HERE, "ref_vectors.json")
        if os.path.exists(old):
            prev = json.load(open(old))
            for key in ("ref150_trajectory", "tokyo_masks"):
                if key in prev:
                    out[key] = prev[key]

    with open(os.path.join(HERE, "ref_vectors.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote", os.path.join(HERE, "ref_vectors.json"))


if __name__ == "__main__":
    main()
