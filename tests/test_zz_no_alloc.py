"""The allocation-free FD product (tsr_solver_jv_noalloc) against the allocating one
(tsr_solver_jv), through the C-ABI, on two calls in a row: a four-scenario batch, then a
smaller single-scenario problem, so the second call finds the shared device scratch
holding the first call's (more numerous) FD steps.  Both entry points run the same
kernels on the same FD steps, so the products agree bit for bit, with the full RHS and
with the linearised one, before and after the micro-cell kernel is tuned."""
import ctypes as C

import numpy as np
import pytest

import twoscalerich_amd as T
from twoscalerich_amd import _ffi as F

pytestmark = pytest.mark.gpu

_dp = C.POINTER(C.c_double)


def _p(a):
    return a.ctypes.data_as(_dp)


def _products(p, lin, seed):
    L = T.lib()
    s = T.Solver(p)
    if lin:
        # the record is allocated here, before the allocation-free calls
        s.set_path(F.PATH_BATCHED)
    rng = np.random.default_rng(seed)
    s.set_state(p.initial() + 0.3 * rng.standard_normal(p.N), 0.0)
    v = rng.standard_normal(p.N)
    eps = 1e-7
    first = np.full(p.N, np.nan)
    ref = np.full(p.N, np.nan)
    again = np.full(p.N, np.nan)
    # untuned: the allocation-free call uses a fixed launch shape
    assert L.tsr_solver_jv_noalloc(s.h, _p(v), eps, int(lin), _p(first)) == 0, T.last_error()
    assert L.tsr_solver_jv(s.h, _p(v), eps, int(lin), _p(ref)) == 0, T.last_error()
    # tuned now
    assert L.tsr_solver_jv_noalloc(s.h, _p(v), eps, int(lin), _p(again)) == 0, T.last_error()
    s.close()
    return first, ref, again


@pytest.mark.parametrize("lin", [False, True])
def test_noalloc_jv_matches_jv_on_two_problems(lin):
    big = T.Problem.config(F.CFG_SCENARIOS, 9, 6, 4)
    small = T.Problem.config(F.CFG_PLAIN_SMALL, 7, 5)
    assert small.view.contents.n_scen < big.view.contents.n_scen and small.N < big.N
    for p, seed in ((big, 1), (small, 2)):
        first, ref, again = _products(p, lin, seed)
        print(f"[no-alloc] N={p.N} scenarios={p.view.contents.n_scen} lin={lin}: "
              f"max|ref|={np.max(np.abs(ref)):.3e}", flush=True)
        assert np.all(np.isfinite(ref)) and np.max(np.abs(ref)) > 0.0
        assert np.array_equal(first, ref)
        assert np.array_equal(again, ref)
    big.close()
    small.close()


def test_noalloc_jv_rejects_what_it_cannot_do_without_allocating():
    L = T.lib()
    p = T.Problem.config(F.CFG_PLAIN_SMALL, 7, 5)
    s = T.Solver(p)
    v = np.ones(p.N)
    out = np.zeros(p.N)
    # the linearisation record is not allocated on the plain path
    assert L.tsr_solver_jv_noalloc(s.h, _p(v), 1e-7, 1, _p(out)) == 1
    assert "linearisation record" in T.last_error()
    assert L.tsr_solver_jv_noalloc(s.h, _p(v), 0.0, 0, _p(out)) == 1
    s.close()
    p.close()
