"""TSR_PATH_GRAPH: Arnoldi windows captured into HIP graphs and replayed give the same
integration as the same path without graphs, bit for bit (same kernels, same
arguments), on the small workload (learned controller and classical law) and on a batch
of scenarios whose windows stop at different dimensions; replays outnumber captures."""
import ctypes as C
import time

import numpy as np
import pytest

import twoscalerich_amd as T
from twoscalerich_amd import _ffi as F
from twoscalerich_amd import controller as CT

pytestmark = pytest.mark.gpu


def _run(p, path, tend, learned):
    s = T.Solver(p, p.options())
    s.set_path(path)
    if learned:
        CT.attach(s)
    s.reset()
    s.set_state(p.initial(), 0.0)
    s.run(min(tend, 1e-9), 2)          # kernel tuning before the timed run
    s.reset()
    s.set_state(p.initial(), 0.0)
    T.lib().tsr_device_sync()
    t0 = time.perf_counter()
    rc, st, _ = s.run(tend, 10 ** 6)
    T.lib().tsr_device_sync()
    wall = time.perf_counter() - t0
    assert rc == 0, T.last_error()
    u, t = s.get_state()
    n_l, n_c = C.c_longlong(), C.c_longlong()
    T.lib().tsr_solver_graph_stats(s.h, C.byref(n_l), C.byref(n_c))
    s.close()
    return u, t, st, wall, n_l.value, n_c.value


@pytest.mark.parametrize("learned", [False, True], ids=["classical", "learned"])
def test_graph_windows_match_small_workload(learned):
    p = T.Problem.config(F.CFG_PLAIN_SMALL)
    tend = p.tend()
    ua, ta, sa, wa, _, _ = _run(p, F.PATH_BATCHED, tend, learned)
    ub, tb, sb, wb, nl, nc = _run(p, F.PATH_BATCHED | F.PATH_GRAPH, tend, learned)
    print(f"[graph] small workload ({'learned' if learned else 'classical'}): {sa.steps} steps; "
          f"without graphs {wa:.3f} s, with {wb:.3f} s ({nl} window replays, {nc} captures)", flush=True)
    assert ta == tb == tend and sa.steps == sb.steps and sa.fevals == sb.fevals
    assert np.array_equal(ua, ub)
    assert nl > nc > 0
    p.close()


def test_graph_windows_match_scenario_batch():
    p = T.Problem.config(F.CFG_SCENARIOS, 9, 8, 4)
    tend = 1e-3
    ua, _, sa, wa, _, _ = _run(p, F.PATH_BATCHED, tend, False)
    ub, _, sb, wb, nl, nc = _run(p, F.PATH_BATCHED | F.PATH_GRAPH, tend, False)
    print(f"[graph] 4 scenarios: {sa.steps} solver steps; without graphs {wa:.3f} s, with {wb:.3f} s "
          f"({nl} replays, {nc} captures)", flush=True)
    assert sa.steps == sb.steps and np.array_equal(ua, ub) and nl > 0
    p.close()


def test_graph_needs_deferred_windows():
    p = T.Problem.config(F.CFG_PLAIN_SMALL, 7, 5)
    s = T.Solver(p)
    assert T.lib().tsr_solver_set_path(s.h, F.PATH_DCGS2 | F.PATH_GRAPH) == 1
    assert "DEFERRED" in T.last_error()
    s.close()
    p.close()
