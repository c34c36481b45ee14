"""Failure paths of one EXPREM step on the GPU (exprem_c), forced from the host: an RHS on
the host (the oracle's G wrapped so that it returns a nonzero flag where the test says)
drives the device Krylov arithmetic; the CPU oracle's EXPREM runs with the same wrapped
RHS and must react the same way:
  * a failed RHS at the trial point U2 rejects the attempt and halves the step
    (check_flag_Gfunc, include/exprem.h:30); the step then completes as the oracle's;
  * a failed RHS at the step start returns -1 (src/main.cpp:319-322);
  * an RHS that keeps failing at U2 halves the step below MinStep: -1;
  * MaxNumSteps reached: -1 before any work;
  * a non-finite state on the device path (Gfunc_hip): -1 with the reason.
No test makes a kernel fault: the failures are flags and non-finite values."""
import ctypes as C

import numpy as np
import pytest

import twoscalerich_amd as T
from twoscalerich_amd import _ffi as F
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

_dp = C.POINTER(C.c_double)


class FaultyRHS:
    """The oracle's Gfunc_parallel; the first call of a step (the step start) records the
    state; a later call whose state is farther than `far` from it (max norm) is the trial
    point U2 (finite-difference states differ by ~eps only).  fail_start / fail_u2 say how
    many such calls return flag 1."""

    def __init__(self, o, n, fail_start=0, fail_u2=0, far=1e-4):
        self.g = o.lib.Gfunc_parallel
        self.n, self.fail_start, self.fail_u2, self.far = n, fail_start, fail_u2, far
        self.start = None
        self.calls = 0
        self.u2_calls = 0
        self.fn = F.GFUNC_T(self._call)

    def new_step(self):
        self.start = None

    def _call(self, gp, up, ud):
        self.calls += 1
        rc = self.g(gp, up, ud)
        u = np.ctypeslib.as_array(up, shape=(self.n,))
        if self.start is None:
            self.start = u.copy()
            if self.fail_start > 0:
                self.fail_start -= 1
                return 1
            return rc
        if np.max(np.abs(u - self.start)) > self.far:
            self.u2_calls += 1
            if self.fail_u2 > 0:
                self.fail_u2 -= 1
                return 1
        return rc

    def ptr(self):
        return C.cast(self.fn, C.c_void_p)


def _problem():
    p = T.Problem.config(F.CFG_PLAIN_SMALL, 7, 5)
    return p, Oracle(p.view, p.ud)


def _gpu_step(p, rhs, u, t, opt, st):
    rhs.new_step()
    tt = C.c_double(t)
    rc = T.lib().exprem_c(rhs.ptr(), p.N, C.byref(tt), p.tend(), None, None, u.ctypes.data_as(_dp), p.ud.contents,
                          opt, C.byref(st))
    return rc, tt.value


def _cpu_step(p, o, rhs, u, t, opt, st):
    rhs.new_step()
    tt = C.c_double(t)
    rc = o.lib.oracle_exprem(rhs.ptr(), p.N, C.byref(tt), p.tend(), None, None, u.ctypes.data_as(_dp),
                             p.ud.contents, opt, C.byref(st))
    return rc, tt.value


def test_rhs_failure_at_trial_point_halves_the_step():
    p, o = _problem()
    opt = p.options()
    results = []
    for side in ("gpu", "cpu"):
        u = p.initial()
        st = F.exprem_stats_struct()
        clean = FaultyRHS(o, p.N)
        step = _gpu_step if side == "gpu" else (lambda p_, r, u_, t_, op, s_: _cpu_step(p_, o, r, u_, t_, op, s_))
        rc, t1 = step(p, clean, u, 0.0, opt, st)                 # one clean step
        assert rc == 0
        tau_prop = st.CurrentStep
        faulty = FaultyRHS(o, p.N, fail_u2=1)
        rc, t2 = step(p, faulty, u, t1, opt, st)                 # U2 fails once
        assert rc == 0 and faulty.fail_u2 == 0 and faulty.u2_calls >= 2
        results.append((t1, t2, tau_prop, u.copy(), st.FuncEvals))
        print(f"[failure] {side}: step 2 took {t2 - t1:.6e} after proposal {tau_prop:.6e}, "
              f"{faulty.u2_calls} trial-point evaluations", flush=True)
    (g1, g2, gp, gu, gfe), (c1, c2, cp, cu, cfe) = results
    # the retried attempt used half the proposal (or less, after its own error test)
    assert g2 - g1 <= 0.5 * gp * (1 + 1e-12)
    assert abs((g2 - g1) - (c2 - c1)) <= 1e-8 * (c2 - c1) and gfe == cfe
    assert np.max(np.abs(gu - cu)) <= 1e-9 * np.max(np.abs(cu))
    o.close()
    p.close()


def test_rhs_failure_at_step_start_returns_minus_one():
    p, o = _problem()
    opt = p.options()
    for side in ("gpu", "cpu"):
        u = p.initial()
        st = F.exprem_stats_struct()
        r = FaultyRHS(o, p.N, fail_start=1)
        rc, t = (_gpu_step(p, r, u, 0.0, opt, st) if side == "gpu" else _cpu_step(p, o, r, u, 0.0, opt, st))
        assert rc == -1 and t == 0.0, side
    assert "step start" in T.last_error()
    o.close()
    p.close()


def test_rhs_that_keeps_failing_ends_below_min_step():
    p, o = _problem()
    opt = p.options()
    opt.MinStep = 1e-6
    counts = []
    for side in ("gpu", "cpu"):
        u = p.initial()
        u0 = u.copy()
        st = F.exprem_stats_struct()
        r = FaultyRHS(o, p.N, fail_u2=10 ** 6)
        rc, t = (_gpu_step(p, r, u, 0.0, opt, st) if side == "gpu" else _cpu_step(p, o, r, u, 0.0, opt, st))
        # the first attempt's step halves until below MinStep: one trial point per attempt,
        # at most log2(InitStep / MinStep) + 1 of them
        print(f"[failure] {side}: rc {rc}, {r.u2_calls} trial points", flush=True)
        assert rc == -1 and t == 0.0 and np.array_equal(u, u0), side
        assert 1 <= r.u2_calls <= int(np.floor(np.log2(opt.InitStep / opt.MinStep))) + 1
        counts.append(r.u2_calls)
    assert counts[0] == counts[1]
    assert "MinStep" in T.last_error()
    o.close()
    p.close()


def test_max_num_steps():
    p, o = _problem()
    opt = p.options()
    opt.MaxNumSteps = 2
    for side in ("gpu", "cpu"):
        u = p.initial()
        st = F.exprem_stats_struct()
        r = FaultyRHS(o, p.N)
        t = 0.0
        for k in range(2):
            rc, t = (_gpu_step(p, r, u, t, opt, st) if side == "gpu" else _cpu_step(p, o, r, u, t, opt, st))
            assert rc == 0
        calls = r.calls
        rc, t3 = (_gpu_step(p, r, u, t, opt, st) if side == "gpu" else _cpu_step(p, o, r, u, t, opt, st))
        assert rc == -1 and t3 == t and r.calls == calls and st.NumSteps == 2, side
    assert "MaxNumSteps" in T.last_error()
    o.close()
    p.close()


def test_non_finite_state_on_the_device_path():
    p = T.Problem.config(F.CFG_PLAIN_SMALL, 7, 5)
    L = T.lib()
    u = p.initial()
    u[p.N // 2] = np.nan
    st = F.exprem_stats_struct()
    t = C.c_double(0.0)
    gf = C.cast(L.Gfunc_hip, C.c_void_p)
    rc = L.exprem_c(gf, p.N, C.byref(t), p.tend(), None, None, u.ctypes.data_as(_dp), p.ud.contents, p.options(),
                    C.byref(st))
    assert rc == -1 and t.value == 0.0 and "not finite" in T.last_error()
    g = np.empty(p.N)
    assert L.Gfunc_hip(g.ctypes.data_as(_dp), u.ctypes.data_as(_dp), p.ud.contents) == 1
    p.close()
