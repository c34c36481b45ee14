"""The batched solver path on the ragged and multi-scenario workloads, and what the
learned controller gains over the classical law.

  * ragged cells (micro grids of 11..31 nodes, 1-3 inclusions per cell) and a batch of
    independent scenarios with sampled soils: the batched path integrates to a common
    time and agrees with the plain path (the reference algorithm) and, for the batch,
    with each scenario integrated alone;
  * the learned controller reaches the same time with fewer RHS evaluations than the
    classical law on the same (batched) path, and its solution agrees with the plain
    path's to the accuracy both have against a tight-tolerance solution.
"""
import numpy as np
import pytest

import twoscalerich_amd as T
from twoscalerich_amd import _ffi as F
from twoscalerich_amd import controller as CT

pytestmark = pytest.mark.gpu


def _run(p, path, tend, controller="classical", opt=None, max_steps=100000):
    s = T.Solver(p, opt or p.options())
    s.set_path(path)
    if controller == "learned":
        CT.attach(s)
    s.reset()
    s.set_state(p.initial(), 0.0)
    rc, st, _ = s.run(tend, max_steps)
    assert rc == 0, T.last_error()
    u, t = s.get_state()
    s.close()
    return u, st


def _rel2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_ragged_batched_matches_plain():
    p = T.Problem.config(F.CFG_RAGGED, macro_n=21)
    assert p.view.contents.n_tmpl == 4
    tend = 2e-4
    ua, sa = _run(p, F.PATH_PLAIN, tend)
    ub, sb = _run(p, F.PATH_BATCHED, tend)
    d = _rel2(ub, ua)
    print(f"[batched] ragged: plain {sa.steps} steps / {sa.fevals} RHS, batched {sb.steps} / {sb.fevals}, "
          f"rel diff {d:.2e}", flush=True)
    assert d <= 1e-6
    p.close()


def test_scenario_batch_matches_single_scenarios():
    S = 4
    pb = T.Problem.config(F.CFG_SCENARIOS, macro_n=21, micro_r=11, n_scen=S)
    v = pb.view.contents
    offs = [v.scen[q].u_offset for q in range(S)] + [pb.N]
    tend = 1e-3
    ub, sb = _run(pb, F.PATH_BATCHED, tend)
    up, sp = _run(pb, F.PATH_PLAIN, tend)
    print(f"[batched] {S} scenarios: batched {sb.steps} solver steps / {sb.fevals} RHS, plain {sp.steps} / {sp.fevals}",
          flush=True)
    for q in range(S):
        a, b = ub[offs[q]:offs[q + 1]], up[offs[q]:offs[q + 1]]
        d = _rel2(a, b)
        print(f"[batched]   scenario {q}: rel diff batched vs plain {d:.2e}", flush=True)
        assert d <= 1e-6
    # the first scenario of the batch is the single-scenario problem of the same seed
    p1 = T.Problem.config(F.CFG_SCENARIOS, macro_n=21, micro_r=11, n_scen=1)
    u1, _ = _run(p1, F.PATH_BATCHED, tend)
    d = _rel2(ub[offs[0]:offs[1]], u1)
    print(f"[batched]   scenario 0 alone vs in the batch: {d:.2e}", flush=True)
    # the batch cuts every scenario's deferred-check window at the earliest prediction,
    # so the DCGS2 corrections past convergence (O(eps) terms) may differ
    assert d <= 1e-9
    p1.close()
    pb.close()


def test_scenario_batch_same_on_host_pool_and_one_thread():
    """A batch's per-scenario host work (the phi evaluations of the stopping rule, the step
    and the error estimate) spread over the host pool gives the same states, bit for bit,
    and the same step and RHS counts as on the stepping thread alone, on the batched path
    with the learned controller and on the plain path."""
    L = T.lib()
    S = 8
    pb = T.Problem.config(F.CFG_SCENARIOS, macro_n=21, micro_r=11, n_scen=S)
    tend = 5e-4
    try:
        for path, ctl in ((F.PATH_BATCHED, "learned"), (F.PATH_PLAIN, "classical")):
            nthr = L.tsr_host_threads(0)
            ua, sa = _run(pb, path, tend, controller=ctl)
            assert L.tsr_host_threads(1) == 1
            ub, sb = _run(pb, path, tend, controller=ctl)
            print(f"[batched] {S} scenarios, path {path}: {nthr} host threads {sa.steps} steps / {sa.fevals} RHS / "
                  f"{sa.wall_s:.2f} s, one thread {sb.steps} / {sb.fevals} / {sb.wall_s:.2f} s", flush=True)
            assert (sa.steps, sa.fevals, sa.rejects) == (sb.steps, sb.fevals, sb.rejects)
            assert np.array_equal(ua, ub)
    finally:
        L.tsr_host_threads(0)
        pb.close()


def test_scenario_batch_on_fp32_basis():
    """The FP32 basis copies with scenarios that stop their Arnoldi windows at different
    dimensions (each keeps the copy of the column it stopped at): every scenario of the
    batch stays within FP32 rounding of the FP64-basis batched path."""
    S = 4
    pb = T.Problem.config(F.CFG_SCENARIOS, macro_n=21, micro_r=11, n_scen=S)
    v = pb.view.contents
    offs = [v.scen[q].u_offset for q in range(S)] + [pb.N]
    tend = 1e-3
    u64, s64 = _run(pb, F.PATH_BATCHED, tend)
    u32, s32 = _run(pb, F.PATH_BATCHED | F.PATH_F32BASIS, tend)
    print(f"[batched] {S} scenarios: FP64 basis {s64.steps} solver steps / {s64.fevals} RHS, "
          f"FP32 basis {s32.steps} / {s32.fevals}", flush=True)
    for q in range(S):
        d = _rel2(u32[offs[q]:offs[q + 1]], u64[offs[q]:offs[q + 1]])
        print(f"[batched]   scenario {q}: FP32 vs FP64 basis {d:.2e}", flush=True)
        assert d <= 1e-6
    pb.close()


def test_learned_controller_gain():
    p = T.Problem.config(F.CFG_PLAIN)
    tend = 0.1
    uc, sc = _run(p, F.PATH_BATCHED, tend)
    ul, sl = _run(p, F.PATH_BATCHED, tend, controller="learned")
    up, _ = _run(p, F.PATH_PLAIN, tend)
    d_c, d_l = _rel2(uc, up), _rel2(ul, up)
    print(f"[batched] plain workload to t={tend}: classical {sc.steps} steps / {sc.fevals} RHS / {sc.wall_s:.1f} s, "
          f"learned {sl.steps} / {sl.fevals} / {sl.wall_s:.1f} s; rel diff vs plain: classical {d_c:.2e}, "
          f"learned {d_l:.2e}", flush=True)
    # measured on one MI355X: 25,711 vs 27,659 RHS evaluations, 10.2 vs 11.5 s.  The gain is
    # asserted in RHS evaluations (deterministic); the wall times are printed only, since
    # they depend on the box
    assert sl.fevals < 0.97 * sc.fevals
    assert sl.rejects <= sc.rejects + 5
    # both differ from the plain path by its own global error (~3e-6 against a
    # tight-tolerance solution at this time, profiles/)
    assert d_l <= 1e-5 and d_c <= 1e-5
    p.close()


def test_accuracy_gate_against_tight_tolerance():
    """The metric's accuracy condition on the plain GPU-scale workload: at rtol = atol =
    1e-7 the plain path and the learned + batched path both land within 1e-6 (relative,
    max norm) of a plain-path solution at rtol = atol = 1e-10, at t = 0.005 d (every run
    from InitStep 1e-9, as the tight run needs)."""
    p = T.Problem.config(F.CFG_PLAIN)
    tend = 0.005

    def opts(tol):
        o = p.options()
        o.RelTol = o.AbsTol = tol
        o.InitStep, o.MinStep = 1e-9, 1e-16
        return o

    uref, sref = _run(p, F.PATH_PLAIN, tend, opt=opts(1e-10), max_steps=10 ** 6)
    up, sp = _run(p, F.PATH_PLAIN, tend, opt=opts(1e-7))
    ul, sl = _run(p, F.PATH_BATCHED, tend, controller="learned", opt=opts(1e-7))
    scale = np.max(np.abs(uref))
    ep, el = np.max(np.abs(up - uref)) / scale, np.max(np.abs(ul - uref)) / scale
    print(f"[batched] accuracy at t={tend}: reference {sref.steps} steps; rtol 1e-7: plain {sp.steps} steps "
          f"max-rel {ep:.2e}, learned+batched {sl.steps} steps max-rel {el:.2e}", flush=True)
    assert ep <= 1e-6 and el <= 1e-6
    p.close()
