"""The batched solver path (TSR_PATH_*) against the plain path and the CPU oracle.

  * deferred stopping-rule checks give the plain path's result bit for bit;
  * DCGS2 agrees with CGS2 to rounding;
  * the linearised FD product agrees with the full one to O(eps), reproduces G(u)
    exactly at a zero perturbation, and its linearisation error is second order in the
    perturbation (a wrong derivative would make it first order);
  * integrations to t_end of the small workload by every path (the FP32-basis one too)
    land within 1e-6 of the oracle's solution.
"""
import numpy as np
import pytest

import twoscalerich_amd as T
from twoscalerich_amd import _ffi as F

from helpers import rel_inf

pytestmark = pytest.mark.gpu


def _solver(cfg=F.CFG_PLAIN_SMALL, path=F.PATH_PLAIN):
    p = T.Problem.config(cfg)
    s = T.Solver(p, p.options())
    s.set_path(path)
    s.reset()
    s.set_state(p.initial(), 0.0)
    return p, s


def _run(s, tend, steps):
    rc, st, recs = s.run(tend, steps, max_records=steps)
    assert rc == 0, T.last_error()
    u, t = s.get_state()
    return u, t, recs[:st.steps], st


def test_deferred_checks_match_plain_bitwise():
    p, a = _solver()
    ua, ta, ra, _ = _run(a, p.tend(), 40)
    _, b = _solver(path=F.PATH_DEFERRED)
    ub, tb, rb, stb = _run(b, p.tend(), 40)
    print(f"[paths] deferred: t={tb:.4e} fevals={stb.fevals}", flush=True)
    assert ta == tb
    assert [r.m for r in ra] == [r.m for r in rb]
    assert [r.mD for r in ra] == [r.mD for r in rb]
    assert np.array_equal(ua, ub)
    a.close(); b.close(); p.close()


def test_dcgs2_matches_cgs2():
    p, a = _solver()
    ua, ta, ra, _ = _run(a, p.tend(), 40)
    _, b = _solver(path=F.PATH_DCGS2 | F.PATH_DEFERRED)
    ub, tb, rb, _ = _run(b, p.tend(), 40)
    ma = np.array([r.m for r in ra]); mb = np.array([r.m for r in rb])
    print(f"[paths] dcgs2: t {ta:.6e} vs {tb:.6e}, m differs in {np.sum(ma != mb)} of {len(ma)} steps, "
          f"rel {rel_inf(ub, ua):.2e}", flush=True)
    assert abs(ta - tb) <= 1e-9 * ta
    assert rel_inf(ub, ua) <= 1e-9
    a.close(); b.close(); p.close()


def test_linearised_products():
    p, s = _solver()
    _run(s, p.tend(), 30)          # a state with fronts in it
    u, _ = s.get_state()
    rng = np.random.default_rng(3)
    v = rng.uniform(-1.0, 1.0, p.N)
    # zero perturbation: both quotients are exactly zero (G(u) reproduced bit for bit)
    z = np.zeros(p.N)
    assert not np.any(s.jv(z, 1e-6, lin=False))
    assert not np.any(s.jv(z, 1e-6, lin=True))
    # the solver's FD step for a unit vector
    eps = np.sqrt(np.finfo(float).eps) * (1.0 + np.linalg.norm(u)) / np.linalg.norm(v)
    jf = s.jv(v, eps, lin=False)
    jl = s.jv(v, eps, lin=True)
    d = np.linalg.norm(jl - jf) / np.linalg.norm(jf)
    print(f"[paths] linearised FD product: rel diff {d:.3e} at eps {eps:.3e}", flush=True)
    assert d <= 1e-5
    # linearisation error || G~(u + h v) - G(u + h v) || = h || jl(h) - jf(h) || is O(h^2)
    errs = []
    for h in (1e-2, 1e-3, 1e-4):
        errs.append(h * np.linalg.norm(s.jv(v, h, lin=True) - s.jv(v, h, lin=False)))
    print(f"[paths] linearisation error at h = 1e-2, 1e-3, 1e-4: {errs}", flush=True)
    assert errs[1] < 0.03 * errs[0]
    assert errs[2] < 0.03 * errs[1]
    s.close(); p.close()


@pytest.fixture(scope="module")
def oracle_end():
    """The oracle's solution of the small workload at t_end (about 10 s on 8 cores)."""
    from oracle.oracle import Oracle
    import os
    p = T.Problem.config(F.CFG_PLAIN_SMALL)
    o = Oracle(p.view, p.ud)
    ud = p.ud.contents
    ud.NUM_THREADS = min(16, os.cpu_count() or 1)
    opt = p.options()
    u = p.initial()
    st = F.exprem_stats_struct()
    t, n = 0.0, 0
    while t < p.tend():
        rc, t, _ = o.exprem_step(ud, u, t, p.tend(), opt, st)
        assert rc == 0
        n += 1
    o.close()
    print(f"[paths] oracle: {n} steps, {st.FuncEvals} RHS evaluations to t={t}", flush=True)
    return u, n, st.FuncEvals


@pytest.mark.parametrize("path", [F.PATH_PLAIN, F.PATH_DCGS2 | F.PATH_DEFERRED, F.PATH_BATCHED,
                                  F.PATH_BATCHED | F.PATH_F32BASIS])
def test_path_to_tend_matches_oracle(oracle_end, path):
    u_ref, n_ref, fe_ref = oracle_end
    p, s = _solver(path=path)
    u, t, recs, st = _run(s, p.tend(), 100000)
    err = rel_inf(u, u_ref)
    print(f"[paths] path {path}: {st.steps} steps ({n_ref} oracle), {st.fevals} RHS evaluations "
          f"({fe_ref} oracle), rel err at t_end {err:.3e}", flush=True)
    assert t == p.tend()
    assert err <= 1e-6
    s.close(); p.close()


def test_f32_basis_needs_dcgs2_and_tracks_the_f64_basis():
    """TSR_PATH_F32BASIS: rejected without DCGS2; with it, 40 steps stay within FP32
    rounding of the same path on the FP64 basis."""
    p, a = _solver(path=F.PATH_BATCHED)
    assert T.lib().tsr_solver_set_path(a.h, F.PATH_F32BASIS | F.PATH_DEFERRED) == 1
    ua, ta, ra, _ = _run(a, p.tend(), 40)
    _, b = _solver(path=F.PATH_BATCHED | F.PATH_F32BASIS)
    ub, tb, rb, _ = _run(b, p.tend(), 40)
    ma = np.array([r.m for r in ra]); mb = np.array([r.m for r in rb])
    d = rel_inf(ub, ua)
    print(f"[paths] f32 basis: t {ta:.6e} vs {tb:.6e}, m differs in {np.sum(ma != mb)} of {len(ma)} steps, "
          f"rel {d:.2e}", flush=True)
    assert abs(ta - tb) <= 1e-6 * ta
    assert d <= 1e-6
    a.close(); b.close(); p.close()
