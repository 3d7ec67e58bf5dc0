import os
import sys
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from concurrent.futures import TimeoutError as FutureTimeout

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and libles_hip.so")
    config.addinivalue_line("markers", "slow: longer CPU test")
    config.addinivalue_line("markers", "ref300: compares with the reference Fortran's 100-step "
                                       "300x300x90 runs, started in the background when collected")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle
    return oracle.lib()


# ---- the reference's 300x300x90 trajectories, computed beside the other tests ----------------
# A 100-step run of the serial reference at 300x300x90 takes minutes of one CPU core; the runs
# start as soon as a test that needs them is collected and proceed in background threads (each
# one a child process) while the GPU tests before them run.
_REF_POOL = None
_REF_JOBS = {}
_REF_LOCK = threading.Lock()


def _ref300_job(cfg):
    from les_halo_amd import inputs
    from oracle.ref import run_ref
    import common
    case = inputs.make_case(300, 300, 90)
    return run_ref.trajectory(case, common.REF_PHYS, 100, [10, 100], cfg=cfg,
                              names=("u", "v", "w", "p"))


def ref300_binary(cfg):
    return os.path.join(ROOT, "oracle", "_ref", "ref300cfg" if cfg else "ref300", "ref_harness")


def _start_ref300(cfg):
    global _REF_POOL
    with _REF_LOCK:
        if cfg in _REF_JOBS:
            return
        if _REF_POOL is None:
            _REF_POOL = ThreadPoolExecutor(max_workers=2, thread_name_prefix="ref300")
        _REF_JOBS[cfg] = (_REF_POOL.submit(_ref300_job, cfg), time.time())


def pytest_collection_finish(session):
    for item in session.items:
        m = item.get_closest_marker("ref300")
        if m is not None:
            cfg = bool(m.kwargs.get("cfg", False))
            if os.path.exists(ref300_binary(cfg)):
                _start_ref300(cfg)


def pytest_sessionfinish(session, exitstatus):
    if _REF_POOL is not None:
        _REF_POOL.shutdown(wait=False, cancel_futures=True)


@pytest.fixture
def ref300(request, capsys):
    """ref300(cfg) -> [state after 10 steps, state after 100 steps] (u, v, w, p) of the reference
    Fortran at 300x300x90 (cfg: the build with the benchmark's SOR parameters), waiting for the
    background run with a progress line every 30 s."""
    def get(cfg):
        if not os.path.exists(ref300_binary(cfg)):
            pytest.skip("reference build %s absent" % ref300_binary(cfg))
        _start_ref300(cfg)
        fut, t0 = _REF_JOBS[cfg]
        while True:
            try:
                return fut.result(timeout=30)
            except FutureTimeout:
                with capsys.disabled():
                    print("\n  [waiting for the reference run (%s): %.0f s since it started]"
                          % ("omega 1.5, no early exit" if cfg else "press.f95 defaults",
                             time.time() - t0), flush=True)
    return get
