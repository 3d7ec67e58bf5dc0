#!/usr/bin/env python3
"""Benchmark of the LES_Halo time step on one MI355X.

Workload: the plain LES step at 300x300x90 FP32 (dx = dy = 4 m, dz = 2 m stretched x1.05 above
40 m, seeded synthetic buildings covering 30 % of the ground, log-law inflow u* = 0.2 m/s with
1 % noise, dt = 0.05 s, Smagorinsky Cs = 0.1, exactly 50 red-black SOR iterations at omega 1.5).
A step is one pass of the whole time step (velnw .. press) over the domain, replayed from a
captured HIP graph with every input already resident in HBM.

Prints one JSON line: value = M cell-steps/s of the whole job; roofline = the dominant kernel
(the fused SOR iteration, or one colour sweep where the grid does not allow it) timed live with
HIP events on the library's stream, against the HBM peak, priced at the algorithmic bytes of
one SOR iteration at the stage boundaries of press.f95 (two colour sweeps, 16 B per cell; the
fused kernel's own compulsory 12 B per cell are reported beside it); ensemble = the batched
leg: the 64 uneven members of the ensemble configuration advanced together, against the same
members run one after another; cpu_baseline = the reference Fortran time loop (built from the
reference sources by oracle/ref/build_ref.sh with the benchmark's SOR settings, serial) on a
bounded sample of the same workload, or the FP32 CPU oracle when the reference build is absent.

    python bench.py [--gpus 1] [--steps 100] [--warmup 5] [--ensemble 64]
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
# algorithmic FP32 bytes per interior cell-step: velnw 40 + velfg/feedbf/calc_sm 89 + visc/adam 88
# + rhs 28 + 50 SOR iterations at 16 B (two 8 B colour sweeps) + pav/adj 12
B_STEP = 1057
SOR_SWEEP_B_PER_CELL = 8       # one colour sweep: r p 4 + r rhs 2 + w p 2
SOR_ITER_B_PER_CELL = 16       # one SOR iteration at press.f95's stage boundaries: two sweeps
SOR_ITER_FUSED_B_PER_CELL = 12  # the fused iteration's compulsory bytes: r p 4 + r rhs 4 + w p 4
METRIC = "M cell-steps/s (50 SOR iters/step) + % HBM peak, 1 GPU; batched vs plain"
GRID = (300, 300, 90)
PHYS = dict(dt=0.05, ro=1.1763, vn=1.583e-5, alpha=-10.0, beta=-1.0, cs0=0.1, omega=1.5,
            pjuge=0.0, nmaxp=50)


def traffic_from_profiles(kernel):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3 --pmc summary
    (profiles/sor_300x300x90_traffic.json, written from FETCH_SIZE and WRITE_SIZE runs)."""
    path = os.path.join(HERE, "profiles", "sor_300x300x90_traffic.json")
    if not os.path.exists(path):
        return None
    try:
        ent = json.load(open(path)).get(kernel)
        return float(ent["hbm_bytes_per_launch"]) if ent else None
    except (ValueError, KeyError, TypeError):
        return None


def host_cpu():
    """The host's CPU model and the number of CPUs this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, len(os.sched_getaffinity(0))


def cpu_baseline(case, nsteps):
    """The reference's serial time loop on nsteps steps of the same workload (the build with
    the benchmark's SOR settings: omega 1.5, exactly 50 iterations)."""
    from oracle.ref import run_ref
    ip, jp, kp = case["dims"]
    exe = os.path.join(HERE, "oracle", "_ref", run_ref.tag_of(ip, jp, kp, cfg=True), "ref_harness")
    phys = {k: PHYS[k] for k in ("dt", "ro", "vn", "alpha", "beta")}
    model, ncpu = host_cpu()
    if os.path.exists(exe):
        secs = run_ref.timed_steps(exe, case, phys, nsteps)
        kind = "reference"
        sample = ("reference Fortran time loop (flang -O2, press.f95 with omega 1.5 and no early "
                  "exit: 50 SOR iterations per step), %d steps after the init sequence at "
                  "%dx%dx%d, serial" % (nsteps, ip, jp, kp))
    else:
        from oracle.oracle import OracleState
        o = OracleState(ip, jp, kp, **PHYS)
        o.load_case(case)
        o.init()
        o.ifdata()
        t0 = time.perf_counter()
        for n in range(1, nsteps + 1):
            o.step(n)
        secs = time.perf_counter() - t0
        kind = "port"
        sample = "FP32 C++ oracle, %d steps at %dx%dx%d" % (nsteps, ip, jp, kp)
    return {"value": round(ip * jp * kp * nsteps / secs / 1e6, 3), "unit": "M cell-steps/s",
            "cores": 1, "kind": kind, "sample": sample, "seconds": round(secs, 3),
            "host_cpu": model, "host_nproc": ncpu}


CALIB_COPY_BYTES = 256 << 20  # calibration copy: 256 MiB, far larger than the 32 MiB of L2


def profile_sor(L, reps):
    """For rocprofv3 --pmc runs: REPS colour sweeps (k_sor_split), REPS fused iterations of
    each tile shape (k_sor_iter, 10 levels per workgroup) and one device copy of CALIB_COPY_BYTES (read and written once each, its source
    evicted from L2 by the copy itself), whose FETCH_SIZE/WRITE_SIZE calibrate the counters."""
    import torch
    ip, jp, kp = GRID
    n = CALIB_COPY_BYTES // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    ms = L.time_kernel("sor_sweep", reps)
    # every fused variant with a fixed tile depth (10 levels), so that each counter run
    # profiles the same dispatches whatever tile the autotuner chose
    iter_us = {}
    for tj in (4, 8, 16):
        L.set_sor_tile(tj, 10)
        iter_us["k_sor_iter<%d, %d>" % (tj, 256 if tj == 4 else 512)] = round(
            L.time_kernel("sor_iter", reps) * 1e3, 3)
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"launches": reps,
                      "sweep_bytes_per_launch": SOR_SWEEP_B_PER_CELL * ip * jp * kp,
                      "iter_bytes_per_launch": SOR_ITER_B_PER_CELL * ip * jp * kp,
                      "copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n,
                      "sweep_avg_us": round(ms * 1e3, 3), "iter_avg_us": iter_us}),
          flush=True)


def ensemble_bench(nmembers, steps, warmup):
    """The batched-ensemble configuration: nmembers members of uneven sizes (ip, jp in
    96..320, kp 60 or 90, own terrain seed and u*), the same physics as the plain step, each
    autotuned.  Aggregate cell-steps/s of the members advanced together (les_ensemble_run) in
    both forms, every member's graph on its own stream and the batched launches (one launch per
    stage over all members), against the same members run one after another (les_run each);
    value is the faster form."""
    import torch
    from les_halo_amd import LES, Ensemble, inputs
    members, cells = [], 0
    for spec in inputs.ensemble_specs(nmembers):
        ip, jp, kp = spec["dims"]
        L = LES(ip, jp, kp, **PHYS)
        L.load_case(inputs.make_case(ip, jp, kp, seed=spec["seed"], ustar=spec["ustar"]))
        L.ifdata()
        L.autotune(5)
        L.run(1, warmup)
        members.append(L)
        cells += ip * jp * kp
    for L in members:
        L.sync()
    n0 = 1 + warmup
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for L in members:
        L.run(n0, steps)
        L.sync()
    t_seq = time.perf_counter() - t0
    times, n = {}, n0 + steps
    for form, batched in (("concurrent_graphs", False), ("batched_launches", True)):
        E = Ensemble(members, batched=batched)
        E.run(n, 2)  # warm-up (captures the form's graphs)
        E.sync()
        n += 2
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        E.run(n, steps)
        E.sync()
        times[form] = time.perf_counter() - t0
        n += steps
        if batched and not E.batched:
            times.pop(form)
        E.close()
    agg = lambda t: cells * steps / t / 1e6  # noqa: E731
    best = min(times, key=times.get)
    t_ens = times[best]
    return {"members": nmembers, "cells": cells, "steps": steps,
            "value": round(agg(t_ens), 2), "unit": "M cell-steps/s", "form": best,
            "forms": {f: round(agg(t), 2) for f, t in times.items()},
            "sequential_value": round(agg(t_seq), 2), "speedup_vs_sequential": round(t_seq / t_ens, 3),
            "ms_per_step_all_members": round(t_ens / steps * 1e3, 3)}


def tiled_bench(steps, case):
    """The plain step as 2 x 2 tiles of 150 x 150 x 90 on the one GPU (les_tiled: face
    exchanges between the tiles after every stage; SOR colour sweeps on each tile's split copy,
    each storing its tiles' face cells into the neighbours' halos), timed as graph replays, then
    checked bitwise against one handle on the whole domain after the same steps."""
    import numpy as np
    import torch
    from les_halo_amd import LES, TiledLES
    ip, jp, kp = GRID
    sor = {k: PHYS[k] for k in ("omega", "nmaxp")}
    phys = {k: PHYS[k] for k in ("dt", "ro", "vn", "alpha", "beta", "cs0")}
    T = TiledLES(ip, jp, kp, sub_i=2, sub_j=2, **phys, **sor)
    T.load_case(case)
    T.ifdata()
    T.run(1, 2)
    T.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    T.run(3, steps)
    T.sync()
    secs = time.perf_counter() - t0
    L = LES(ip, jp, kp, **phys, **sor, pjuge=0.0)
    L.load_case(case)
    L.ifdata()
    L.run(1, 2 + steps)
    same = all(np.array_equal(T.download(n).view(np.uint32), L.download(n).view(np.uint32))
               for n in ("u", "v", "w", "p"))
    T.close()
    L.close()
    return {"tiles": [2, 2], "tile_grid": [ip // 2, jp // 2, kp], "steps": steps,
            "ms_per_step": round(secs / steps * 1e3, 3),
            "value": round(ip * jp * kp * steps / secs / 1e6, 2), "unit": "M cell-steps/s",
            "bitwise_equal_single_domain_uvwp": bool(same), "steps_compared": 2 + steps}


def adjoint_bench(steps, ckpt=20, nprobes=16):
    """The adjoint configuration: a window of STEPS steps at 150 x 150 x 90 FP32 from the plain
    path's state after the init sequence and two steps, drag field 0.5, objective the mean u at
    16 seeded probes 5-30 m high; checkpoints every ckpt steps.  Times the gradient (forward
    window, segment and step recomputation, reverse) against the plain step and against the
    adjoint form's own forward step."""
    import numpy as np
    import torch
    from les_halo_amd import LES, inputs
    from les_halo_amd.adjoint import AdjointLES, probe_points
    ip, jp, kp = 150, 150, 90
    case = inputs.make_case(ip, jp, kp)
    L = LES(ip, jp, kp, **PHYS)
    L.load_case(case)
    L.ifdata()
    L.run(1, 2)
    L.sync()
    A = AdjointLES.from_les(L, fp64=False)
    A.upload("cd", np.full(ip * jp * kp, 0.5))
    probes = probe_points(case, n=nprobes)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    L.run(3, steps)
    L.sync()
    t_plain = time.perf_counter() - t0
    L.close()
    state = {n: A.download(n) for n in ("u", "v", "w", "p", "f", "g", "h", "fold", "gold", "hold",
                                        "usum", "vsum", "wsum")}
    A.step(1)  # warm-up of every kernel
    for n, a in state.items():
        A.upload(n, a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    A.step(steps)
    A.download("u")
    t_fwd = time.perf_counter() - t0
    for n, a in state.items():
        A.upload(n, a)
    t0 = time.perf_counter()
    J = A.window(steps, ckpt, probes)
    t_grad = time.perf_counter() - t0
    gw, gc = A.gradient("wind_profile"), A.gradient("cd")
    A.close()
    return {"grid": [ip, jp, kp], "dtype": "f32", "steps": steps, "checkpoint_every": ckpt,
            "probes": nprobes, "objective": J, "grad_wind_norm": float(np.linalg.norm(gw)),
            "grad_cd_norm": float(np.linalg.norm(gc)),
            "plain_forward_s": round(t_plain, 4), "adjoint_form_forward_s": round(t_fwd, 4),
            "gradient_s": round(t_grad, 4),
            "gradient_over_plain_forward": round(t_grad / t_plain, 2),
            "gradient_over_own_forward": round(t_grad / t_fwd, 2)}


def mixed_bench(steps, tol=1e-7, refine_max=200, refine_omega=1.95, refine_solver=1):
    """The mixed-precision configuration: 600x600x90, stencils and SOR sweeps in FP32, rhs mean,
    residual and refined pressure in FP64 (LES_PREC_MIXED), refinement to a relative L2
    residual of tol.  Steps run eagerly (the host reads the residual between passes)."""
    import torch
    from les_halo_amd import LES, inputs
    from les_halo_amd.les import LES_PREC_MIXED
    ip, jp, kp = 600, 600, 90
    L = LES(ip, jp, kp, **PHYS, precision=LES_PREC_MIXED, refine_tol=tol, refine_max=refine_max,
            refine_omega=refine_omega, refine_solver=refine_solver)
    L.load_case(inputs.make_case(ip, jp, kp))
    L.ifdata()
    L.run(1, 2)
    L.sync()
    passes, resid, cg_iters = [], [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for n in range(3, 3 + steps):
        L.step(n)
        s = L.scalars()
        passes.append(s["refine_passes"])
        resid.append(s["resid_l2"])
        cg_iters.append(s["cg_iters"])
    L.sync()
    secs = time.perf_counter() - t0
    return {"grid": [ip, jp, kp], "steps": steps, "ms_per_step": round(secs / steps * 1e3, 3),
            "value": round(ip * jp * kp * steps / secs / 1e6, 2), "unit": "M cell-steps/s",
            "refine_tol": tol, "refine_max": refine_max, "refine_omega": refine_omega,
            "refine_solver": ["sor", "pcg"][refine_solver],
            "refine_passes": passes, "cg_iterations": cg_iters,
            "resid_l2_last": resid[-1], "resid_l2_max": max(resid)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-steps", type=int, default=12)
    ap.add_argument("--sor-tile", type=int, nargs=2, default=None, metavar=("TJ", "TK"),
                    help="set the fused SOR tile by hand instead of autotuning (profiling runs: "
                         "then every launch of the kernel is a step's or the live timing's)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--mixed", type=int, default=0, metavar="STEPS",
                    help="also run the mixed-precision 600x600x90 configuration for STEPS steps")
    ap.add_argument("--refine-solver", type=int, default=1, choices=(0, 1),
                    help="mixed precision corrections: 0 SOR sweeps, 1 preconditioned CG")
    ap.add_argument("--refine-omega", type=float, default=None,
                    help="relaxation of the mixed-precision correction solves")
    ap.add_argument("--ensemble", type=int, default=64, metavar="N",
                    help="the batched leg: the ensemble configuration with N members (0: skip)")
    ap.add_argument("--tiled", type=int, default=0, metavar="STEPS",
                    help="also run the 2x2-tiles configuration for STEPS steps")
    ap.add_argument("--adjoint", type=int, default=0, metavar="STEPS",
                    help="also time the adjoint form's gradient over a STEPS-step window")
    ap.add_argument("--profile-sor", type=int, default=0, metavar="REPS",
                    help="counter runs: set up the workload, launch REPS SOR colour sweeps and "
                         "one device copy of the same algorithmic bytes (calibration), no timing")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs the LES step on one GPU only (no multi-GPU path); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from les_halo_amd import LES, inputs
    ip, jp, kp = GRID
    case = inputs.make_case(ip, jp, kp)
    L = LES(ip, jp, kp, **PHYS)
    L.load_case(case)
    L.ifdata()
    if args.sor_tile:
        L.set_sor_tile(*args.sor_tile)
        tuning = {"sor_tile_j": args.sor_tile[0], "sor_tile_k": args.sor_tile[1],
                  "sor_iter_ms": None}
    else:
        tuning = L.autotune(20)  # picks the fused SOR iteration's tile; outside the timed region
    L.run(1, args.warmup)
    L.sync()
    if args.profile_sor:
        profile_sor(L, args.profile_sor)
        return
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    L.run(1 + args.warmup, args.steps)
    L.sync()
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    cells = ip * jp * kp
    ms_per_step = secs / args.steps * 1e3
    value = cells * args.steps / secs / 1e6

    # dominant kernel (the fused SOR iteration when the grid allows it, else the colour
    # sweep), timed live with HIP events on the library's stream
    compulsory = None
    if tuning["sor_tile_j"] > 0:
        kern_ms = L.time_kernel("sor_iter", 100)
        bytes_launch = SOR_ITER_B_PER_CELL * cells
        compulsory = SOR_ITER_FUSED_B_PER_CELL * cells
        # the kernel's name as rocprofv3 reports it (tile rows, workgroup size; the variant
        # that is not a solve's first iteration)
        tj = tuning["sor_tile_j"]
        kname = "k_sor_iter<%d, %d, false>" % (tj, 256 if tj == 4 else 512)
    else:
        kern_ms = L.time_kernel("sor_sweep", 200)
        bytes_launch = SOR_SWEEP_B_PER_CELL * cells
        kname = "k_sor_split<true, false>"
    achieved = bytes_launch / (kern_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic_from_profiles(kname.replace(", false>", ">")), "kernel": kname,
                "bytes_per_launch": bytes_launch, "avg_us": round(kern_ms * 1e3, 3),
                "tile": [tuning["sor_tile_j"], tuning["sor_tile_k"]],
                "tuned_iter_us": (round(tuning["sor_iter_ms"] * 1e3, 3)
                                  if tuning["sor_iter_ms"] is not None else None)}
    if compulsory:
        # the fused kernel's own compulsory traffic (both colours of p and rhs read once, both
        # colours of p written once), priced the same way
        ach_c = compulsory / (kern_ms * 1e-3) / 1e9
        roofline["compulsory_bytes_per_launch"] = compulsory
        roofline["compulsory_achieved"] = round(ach_c, 1)
        roofline["compulsory_frac"] = round(ach_c / HBM_PEAK_GBS, 4)
    step_hbm_frac = B_STEP * cells / (ms_per_step * 1e-3) / (HBM_PEAK_GBS * 1e9)
    autotuned = {"sor_tile": [tuning["sor_tile_j"], tuning["sor_tile_k"]]}
    if "stencil_shapes" in tuning:
        autotuned["stencil_shapes"] = tuning["stencil_shapes"]

    line = {"metric": METRIC, "value": round(value, 2), "unit": "M cell-steps/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (seeded terrain and inflow)",
            "config": {"workload": "LES_Halo plain step, 300x300x90 FP32, 50 red-black SOR "
                                   "iterations (omega 1.5), single domain",
                       "grid": list(GRID), "sor_iters": 50, "omega": 1.5, "dt": 0.05,
                       "parallelism": "single domain, 1 GPU"},
            "step_hbm_frac": round(step_hbm_frac, 4), "step_bytes_per_cell": B_STEP,
            "autotuned": autotuned, "roofline": roofline}
    if args.tiled:
        line["tiled"] = tiled_bench(args.tiled, case)
    if args.adjoint:
        line["adjoint"] = adjoint_bench(args.adjoint)
    if args.mixed:
        romega = args.refine_omega
        if romega is None:  # SOR corrections over-relax hard; the CG preconditioner uses omega
            romega = 1.95 if args.refine_solver == 0 else 0.0
        line["mixed"] = mixed_bench(args.mixed, refine_omega=romega,
                                    refine_solver=args.refine_solver)
    if args.ensemble:
        del L
        line["ensemble"] = ensemble_bench(args.ensemble, args.steps, args.warmup)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(case, args.cpu_steps)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
