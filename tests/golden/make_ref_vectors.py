"""Writes the golden vectors of the reference CPU physics into tests/golden/.

Runs ONLY in the build container, where the reference sources exist: it builds the reference
Fortran with oracle/ref/build_ref.sh and runs it through oracle/ref/ref_harness.f95 on seeded
inputs.  The fixtures hold SHA-256 digests and L2 norms of every output array (plus a few
sampled values and the terrain of the reference's GIS file), so that the tests can pin the
oracle bit for bit without the reference present.

    python tests/golden/make_ref_vectors.py [--big]
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--big", action="store_true", help="also the 150x150x90 reference trajectory")
    args = ap.parse_args()
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
