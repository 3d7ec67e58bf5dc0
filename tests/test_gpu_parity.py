"""HIP path against the oracle, bit for bit.

Each kernel-level entry point of libles_hip.so runs on a seeded random state (every array,
halos included) and must reproduce the oracle's FP32 restatement of the same reference stage
exactly; the oracle runs with exact (order-independent) reductions, the semantics of the HIP
reductions.  The oracle itself is pinned to the reference Fortran by test_oracle_reference.py.
"""
import numpy as np
import pytest

import common
from les_halo_amd import fields as fl
from oracle.oracle import SUM_EXACT

pytestmark = pytest.mark.gpu

STATE = [n for n in common.DUMP_ORDER]


def _gpu(case, **sor):
    from les_halo_amd import LES
    ip, jp, kp = case["dims"]
    L = LES(ip, jp, kp, **common.REF_PHYS, **sor)
    L.load_case(case)
    return L


def _pair(case, seed=None, **sor):
    L = _gpu(case, **sor)
    o = common.oracle_from_case(case, **common.REF_PHYS, **sor, sum_mode=SUM_EXACT)
    if seed is not None:
        ip, jp, kp = case["dims"]
        for name, arr in common.random_state(ip, jp, kp, seed).items():
            L.upload(name, arr)
            o.set(name, arr)
    return L, o


def _compare(L, o, tag, names=STATE):
    bad = []
    for n in names:
        g = L.download(n)
        if not common.bits_equal(g, o.a[n]):
            bad.append(common.mismatch_report(g, o.a[n], n))
    assert not bad, "%s:\n  %s" % (tag, "\n  ".join(bad))


@pytest.fixture(scope="module")
def case():
    return common.tiny_case()


@pytest.fixture(scope="module")
def case_odd():
    return common.tiny_case(14, 11, 8, seed=9)


def test_coefficients_match_init(case):
    L, o = _pair(case)
    _compare(L, o, "init.f95 coefficients", ["cn1", "cn2l", "cn2s", "cn3l", "cn3s", "cn4l",
                                             "cn4s", "delx1"])


# (gpu calls, oracle calls) per stage; oracle calls are (stage, n)
STAGES = {
    "velnw": (lambda L: L.velnw(), [("velnw", 0)]),
    "bondv1_n1": (lambda L: L.bondv1(1), [("bondv1", 1)]),
    "bondv1_n3": (lambda L: L.bondv1(3), [("bondv1", 3)]),
    "state1_n2": (lambda L: (L.velnw__bondv1_init_uvw(2), L.bondv1_calc_uout(), L.bondv1_calc_uvw()),
                  [("velnw", 0), ("bondv1", 2)]),
    "velfg": (lambda L: L.velfg(), [("velfg", 0)]),
    "feedbf": (lambda L: L.feedbf(), [("feedbf", 0)]),
    "state4": (lambda L: L.velfg__feedbf__les_calc_sm(),
               [("velfg", 0), ("feedbf", 0), ("les_calc_sm", 0)]),
    "les_calc_sm": (lambda L: L.les_calc_sm(), [("les_calc_sm", 0)]),
    "les_bound_sm": (lambda L: L.les_bound_sm(), [("boundsm", 0)]),
    "les_calc_visc": (lambda L: L.les_calc_visc(), [("les_calc_visc", 0)]),
    "les": (lambda L: L.les(), [("les", 0)]),
    "adam": (lambda L: L.adam(), [("adam", 0)]),
    "state6": (lambda L: L.les_calc_visc__adam(), [("les_calc_visc", 0), ("adam", 0)]),
    "bondfg": (lambda L: L.bondfg(), [("bondfg", 0)]),
    "boundp1": (lambda L: L.boundp1(), [("boundp1", 0)]),
    "boundp2": (lambda L: L.boundp2(), [("boundp2", 0)]),
    "press_rhsav": (lambda L: L.press_rhsav(), [("press_rhsav", 0)]),
    "press_sor_red": (lambda L: L.press_sor(0), [("sor_colour", 0), ("boundp1", 0)]),
    "press_sor_black": (lambda L: L.press_sor(1), [("sor_colour", 1), ("boundp1", 0)]),
    "press_sor_k": (lambda L: L.press_sor(2), [("boundp2", 0)]),
    "press_sor_solve": (lambda L: L.press_sor_solve(), [("press_sor", 0)]),
    "press_pav_adj_boundp": (lambda L: (L.press_pav(), L.press_adj(), L.press_boundp()),
                             [("press_pav_adj", 0)]),
    "press": (lambda L: L.press(), [("press", 0)]),
    "step_n1": (lambda L: L.step(1), [("step", 1)]),
    "step_n5": (lambda L: L.step(5), [("step", 5)]),
}


@pytest.mark.parametrize("sor", ["ref", "cfg"])
@pytest.mark.parametrize("stage", sorted(STAGES))
def test_stage_bitwise(case, stage, sor):
    sorp = common.REF_SOR if sor == "ref" else common.CFG_SOR
    if sor == "cfg" and "press" not in stage and "step" not in stage:
        pytest.skip("SOR parameters only matter for press")
    gpu_call, orc_calls = STAGES[stage]
    L, o = _pair(case, seed=101 + sorted(STAGES).index(stage), **sorp)
    print("stage", stage, sor, flush=True)
    gpu_call(L)
    for name, n in orc_calls:
        o.call(name, n)
    _compare(L, o, stage)
    s = L.scalars()
    if stage.startswith("bondv1") or stage.startswith("state1") or stage.startswith("step"):
        assert np.float32(s["uout"]) == np.float32(o.scalars.uout) or stage.startswith("step")
    if stage == "press_rhsav":
        assert np.float32(s["rhsav"]) == np.float32(o.scalars.rhsav)
        assert np.float32(s["area"]) == np.float32(o.scalars.area)
    if stage == "press_sor_solve" and sor == "ref":
        assert s["sor_iters"] == o.scalars.sor_iters
        assert np.float32(s["sor"]) == np.float32(o.scalars.sor)
    L.close()


def test_ifdata_from_zero(case):
    L, o = _pair(case)
    L.ifdata()
    o.ifdata()
    _compare(L, o, "ifdata")


@pytest.mark.parametrize("sor", ["ref", "cfg"])
def test_trajectory_eager(case, sor):
    sorp = common.REF_SOR if sor == "ref" else common.CFG_SOR
    L, o = _pair(case, **sorp)
    L.ifdata()
    o.ifdata()
    for n in range(1, 13):
        L.step(n)
        o.step(n)
        if n in (1, 2, 5, 12):
            _compare(L, o, "eager step %d" % n)


def test_trajectory_odd_jm(case_odd):
    L, o = _pair(case_odd, **common.REF_SOR)
    L.ifdata()
    o.ifdata()
    for n in range(1, 4):
        L.step(n)
        o.step(n)
    _compare(L, o, "odd jm step 3")


@pytest.mark.parametrize("sor", ["ref", "cfg"])
def test_trajectory_odd_km(sor):
    """An odd number of levels: kernels that take two or four levels a thread (visc+adam, the
    residual-summing SOR sweeps) end on a lone level whose partner lies past kp."""
    case = common.tiny_case(16, 12, 9, seed=13)
    sorp = common.REF_SOR if sor == "ref" else common.CFG_SOR
    L, o = _pair(case, **sorp)
    L.ifdata()
    o.ifdata()
    for n in range(1, 5):
        L.step(n)
        o.step(n)
    _compare(L, o, "odd km step 4")


def _stretched(case, fx=1.03, fy=0.97):
    """The case on a horizontal grid with varying spacings (dx1, dy1 scaled by powers of fx, fy
    over a 5-cell cycle) and dxs, dys from them as grid.f95:67-82 forms them."""
    c = dict(case)
    ip, jp, kp = c["dims"]
    dx1 = c["dx1"].copy()  # dx1(-1:ip+1)
    dy1 = c["dy1"].copy()  # dy1(0:jp+1)
    for q in range(dx1.size):
        dx1[q] = np.float32(dx1[q] * np.float32(fx ** (q % 5)))
    for q in range(dy1.size):
        dy1[q] = np.float32(dy1[q] * np.float32(fy ** (q % 5)))
    dxs = np.array([dx1[i + 1] / np.float32(2) + dx1[i + 2] / np.float32(2) for i in range(ip + 1)],
                   np.float32)
    dys = np.array([dy1[j] / np.float32(2) + dy1[j + 1] / np.float32(2) for j in range(jp + 1)],
                   np.float32)
    c.update(dx1=dx1, dy1=dy1, dxs=dxs, dys=dys)
    return c


@pytest.mark.parametrize("grid", ["uniform5x3", "stretched"])
@pytest.mark.parametrize("sor", ["ref", "cfg"])
def test_trajectory_horizontal_spacings(grid, sor):
    """Horizontal spacings that are not powers of two (5 m by 3 m: true divisions along i and j
    in every stencil kernel) and spacings that vary along i and j (also the SOR's general
    coefficients instead of the uniform-grid form)."""
    from les_halo_amd import inputs
    case = inputs.make_case(16, 12, 10, seed=17, hmax=12.0, coverage=0.3, dx=5.0, dy=3.0)
    if grid == "stretched":
        case = _stretched(case)
    sorp = common.REF_SOR if sor == "ref" else common.CFG_SOR
    L, o = _pair(case, **sorp)
    L.ifdata()
    o.ifdata()
    for n in range(1, 5):
        L.step(n)
        o.step(n)
    _compare(L, o, "%s grid step 4" % grid)


def test_graph_replay_equals_eager(case):
    """les_run replays a captured HIP graph; several replays with eager work in between."""
    A, o = _pair(case, **common.CFG_SOR)
    B = _gpu(case, **common.CFG_SOR)
    A.ifdata()
    B.ifdata()
    o.ifdata()
    for n in range(1, 4):
        A.step(n)
        o.step(n)
    B.run(1, 3)
    _compare(B, o, "graph after 3")
    for rep in range(3):
        n0 = 4 + 2 * rep
        B.run(n0, 2)
        B.sync()
        _ = B.download("u")
        for n in (n0, n0 + 1):
            o.step(n)
        _compare(B, o, "graph replay %d" % rep, ["u", "v", "w", "p", "f", "g", "h", "sm", "rhs"])


def test_halo_size_matches_reference_buffers():
    """TEMPL:111-119 halo buffer sizes (h_w = 1) from les_halo_size."""
    from les_halo_amd._lib import lib
    ip, jp, kp = 300, 300, 90
    L = lib()
    assert L.les_halo_size(2, ip + 3, jp + 3, kp + 2, 1, 0) == 4 * (ip + jp + 6 - 2) * (kp + 2)


def test_exact_sum_gpu_vs_oracle():
    from les_halo_amd import LES
    from oracle.oracle import exact_sum_f32
    ip, jp, kp = 40, 30, 20
    L = LES(ip, jp, kp)
    n = fl.size("p", ip, jp, kp)
    ptr, cnt = L.device_ptr("p")
    rng = np.random.default_rng(5)
    vecs = [
        rng.standard_normal(n).astype(np.float32),
        np.concatenate([np.float32([1e30, 1.0, -1e30]), np.zeros(n - 3, np.float32)]),
        (rng.standard_normal(n) * np.exp(rng.uniform(-80, 80, n))).astype(np.float32),
        np.full(n, np.float32(1e-45)),
    ]
    half = rng.standard_normal(n // 2).astype(np.float32)
    cancel = np.zeros(n, np.float32)
    cancel[:n // 2] = half
    cancel[n // 2:n // 2 * 2] = -half[::-1]
    vecs.append(cancel)  # exactly zero
    for q, x in enumerate(vecs):
        x = np.ascontiguousarray(x.astype(np.float32))
        L.upload("p", x)
        got = L.reduce_sum_exact(ptr, n)
        want = exact_sum_f32(x)
        assert got.view(np.uint32) == want.view(np.uint32), (q, got, want)


@pytest.mark.slow
def test_full_size_one_step():
    """300x300x90 (the benchmark grid): init and one step, every array bit for bit."""
    from les_halo_amd import inputs
    case = inputs.make_case(300, 300, 90)
    L, o = _pair(case, **common.CFG_SOR)
    L.ifdata()
    o.ifdata()
    L.step(1)
    o.step(1)
    _compare(L, o, "300x300x90 step 1", ["u", "v", "w", "p", "f", "g", "h", "sm", "rhs", "usum",
                                         "fold", "diu1", "diu5", "diu9"])


def _consistent_pair(case, seed, **sor):
    """Random state whose p halos are mirror-consistent (as every press leaves them)."""
    L, o = _pair(case, seed=seed, **sor)
    L.press_boundp()
    o.call("boundp1")
    o.call("boundp2")
    return L, o


@pytest.mark.parametrize("sor", ["ref", "cfg"])
def test_split_press_bitwise(case, sor):
    """press on the colour-split solve (consistent halos, even jp, separable cn1)."""
    sorp = common.REF_SOR if sor == "ref" else common.CFG_SOR
    L, o = _consistent_pair(case, 77, **sorp)
    L.press()
    o.call("press")
    _compare(L, o, "split press")
    s = L.scalars()
    assert np.float32(s["pav"]) == np.float32(o.scalars.pav)
    assert np.float32(s["rhsav"]) == np.float32(o.scalars.rhsav)


def test_split_sor_solve_bitwise(case):
    L, o = _consistent_pair(case, 78, **common.CFG_SOR)
    L.press_sor_solve()
    o.call("press_sor")
    _compare(L, o, "split sor solve")


def test_nonseparable_cn1_falls_back(case):
    """A cn1 that is not 1/((ax+by)+cz) runs the in-place solve and stays exact."""
    L, o = _consistent_pair(case, 79, **common.CFG_SOR)
    cn1 = o.a["cn1"].copy()
    cn1[5] *= np.float32(1.25)
    L.upload("cn1", cn1)
    o.set("cn1", cn1)
    L.press()
    o.call("press")
    _compare(L, o, "non-separable cn1 press")


def test_run_after_inconsistent_upload(case):
    """les_run from arbitrary p halos: first step in place, then the captured graph."""
    L, o = _pair(case, **common.CFG_SOR)
    L.ifdata()
    o.ifdata()
    for n in range(1, 3):
        L.step(n)
        o.step(n)
    p = o.a["p"].copy()
    p[:7] = np.float32(3.5)  # break some halo mirrors
    L.upload("p", p)
    o.set("p", p)
    L.run(3, 4)
    for n in range(3, 7):
        o.step(n)
    _compare(L, o, "run after inconsistent upload")


@pytest.mark.parametrize("dims,tile", [((40, 32, 20), (8, 10)), ((37, 18, 23), (4, 5)),
                                       ((37, 18, 23), (16, 23)), ((40, 32, 20), (4, 45)),
                                       ((33, 18, 13), (4, 6)), ((34, 20, 11), (8, 4))])
def test_fused_sor_iteration(dims, tile):
    """The fused iteration (both colours per launch, ping-pong iterates) with a given tile
    reproduces the oracle's trajectory bit for bit (50 iterations at omega 1.5, no early exit),
    eagerly and from the captured graph.  The widths put the last cell of a colour row (i = ip
    or ip - 1) at every position of its four-cell group; at ip = 33 the rows starting at i = 2
    end in a group of padding only."""
    case = common.tiny_case(*dims, seed=21)
    L, o = _pair(case, **common.CFG_SOR)
    L.set_sor_tile(*tile)
    L.ifdata()
    o.ifdata()
    for n in range(1, 4):
        L.step(n)
        o.step(n)
    _compare(L, o, "fused SOR tile %s" % (tile,), ["u", "v", "w", "p", "rhs", "sm"])
    L.run(4, 2)
    for n in (4, 5):
        o.step(n)
    _compare(L, o, "fused SOR, graph replay", ["u", "v", "w", "p"])


def test_autotune_keeps_results(case):
    L, o = _pair(case, **common.CFG_SOR)
    t = L.autotune(2)
    assert t["sor_tile_j"] in (0, 4, 8, 16) and t["sor_iter_ms"] > 0, t
    L.ifdata()
    o.ifdata()
    L.run(1, 3)
    for n in range(1, 4):
        o.step(n)
    _compare(L, o, "after autotune %s" % (t,), ["u", "v", "w", "p"])


def test_aveflow_statistics_bitwise(case):
    """aveflow.f95:42-157 on the device after every step equals the oracle's statistics."""
    L, o = _pair(case, **common.CFG_SOR)
    L.ifdata()
    o.ifdata()
    st = o.new_stats()
    nsteps, n1 = 6, 3
    for n in range(1, nsteps + 1):
        L.step(n)
        L.aveflow(n, n1, nsteps)
        o.step(n)
        o.aveflow(st, n, n1, nsteps)
    got = L.aveflow_stats()
    bad = [common.mismatch_report(got[n], st[n], n) for n in st if not common.bits_equal(got[n], st[n])]
    assert not bad, "\n".join(bad)


def test_resident_split_solve_interleavings(case):
    """The resident split solve (fused iterations: the split copy of p carried from step to
    step, rhs kept in the split layout and copied to F.rhs only when read) against the oracle
    through every way of leaving and re-entering it: graph replays, eager steps, reads of rhs
    between steps, an upload of p, a lone SOR colour, and p handed out by address (which ends
    the resident solve for the handle)."""
    L, o = _pair(case, **common.CFG_SOR)
    L.set_sor_tile(4, 5)
    L.ifdata()
    o.ifdata()
    names = ["u", "v", "w", "p", "rhs", "sm", "f", "usum"]

    def advance(n0, k, eager=False):
        if eager:
            for n in range(n0, n0 + k):
                L.step(n)
        else:
            L.run(n0, k)
        for n in range(n0, n0 + k):
            o.step(n)

    advance(1, 4)                      # n = 1 eagerly, a scattering step, graph replays
    _compare(L, o, "resident: first steps", names)
    advance(5, 2, eager=True)          # eager steps on the carried copy
    _compare(L, o, "resident: eager steps", names)
    advance(7, 3)                      # replays again after rhs was read
    _compare(L, o, "resident: replays after a read of rhs", names)
    ip, jp, kp = case["dims"]
    pv = fl.FArray("p", ip, jp, kp, data=o.a["p"].copy())
    pv[5, 4, 3] = pv[5, 4, 3] + np.float32(0.25)  # an interior value away from the faces
    p = pv.flat
    L.upload("p", p)
    o.set("p", p)
    advance(10, 3)                     # the upload makes the next solve scatter again
    _compare(L, o, "resident: after an upload of p", names)
    L.press_sor(0)                     # a lone colour sweep on F.p, then steps
    o.call("sor_colour", 0)
    o.call("boundp1")
    advance(13, 2)
    _compare(L, o, "resident: after a lone colour sweep", names)
    L.device_ptr("p")                  # p handed out: the handle leaves the resident solve
    advance(15, 3)
    _compare(L, o, "after p was handed out", names)
