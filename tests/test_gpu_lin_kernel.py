"""The uniform-grid linearised FD-product kernel (k_micro_lin) against k_micro's own
linearised mode on the same problem and inputs: the two must agree bit for bit (same
per-node expressions in the same order, same exchange-sum order), for every compiled
grid size (11, 16, 21, 31), the grid size read per cell (13, and the ragged workload
of mixed sizes), several scenarios in one launch, per-cell inclusion discs, and every
workgroup shape the tuner may pick.  TSR_MICRO_LIN_OFF, read when a problem is set up,
makes that problem use k_micro."""
import os

import numpy as np
import pytest

import twoscalerich_amd as T
from twoscalerich_amd import _ffi as F

pytestmark = pytest.mark.gpu

CASES = [
    ("plain r11", F.CFG_PLAIN, 9, 11, 0),
    ("plain r16", F.CFG_PLAIN, 9, 16, 0),
    ("plain r21", F.CFG_PLAIN, 7, 21, 0),
    ("plain r31", F.CFG_PLAIN, 5, 31, 0),
    ("plain r13 (runtime size)", F.CFG_PLAIN, 7, 13, 0),
    ("ragged", F.CFG_RAGGED, 9, 0, 0),
    ("4 scenarios r16", F.CFG_SCENARIOS, 5, 16, 4),
    ("large discs r31", F.CFG_LARGE, 5, 31, 0),
]


def _problem(cfg, macro_n, micro_r, n_scen, lin_kernel):
    if not lin_kernel:
        os.environ["TSR_MICRO_LIN_OFF"] = "1"
    try:
        return T.Problem.config(cfg, macro_n, micro_r, n_scen, 0)
    finally:
        os.environ.pop("TSR_MICRO_LIN_OFF", None)


@pytest.mark.parametrize("name,cfg,macro_n,micro_r,n_scen", CASES, ids=[c[0] for c in CASES])
def test_lin_kernel_matches_k_micro_bitwise(name, cfg, macro_n, micro_r, n_scen):
    print(f"[lin-kernel] {name}", flush=True)
    pa = _problem(cfg, macro_n, micro_r, n_scen, False)
    pb = _problem(cfg, macro_n, micro_r, n_scen, True)
    assert pa.N == pb.N
    rng = np.random.default_rng(11)
    u0 = pa.initial()
    u = u0 + 0.8 * rng.standard_normal(pa.N)      # heads around the initial state, some wet
    v = rng.uniform(-1.0, 1.0, pa.N)
    eps = np.sqrt(np.finfo(float).eps) * (1.0 + np.linalg.norm(u)) / np.linalg.norm(v)
    sa, sb = T.Solver(pa), T.Solver(pb)
    for s in (sa, sb):
        s.set_state(u, 0.0)
    ja = sa.jv(v, eps, lin=True)
    jb = sb.jv(v, eps, lin=True)
    assert np.all(np.isfinite(ja))
    diff = int(np.sum(ja != jb))
    print(f"[lin-kernel] {name}: N={pa.N}, entries differing: {diff}", flush=True)
    assert diff == 0
    # a second product on the same record (the solver's Arnoldi iterations) and a zero one
    w = rng.uniform(-1.0, 1.0, pa.N)
    assert np.array_equal(sa.jv(w, eps, lin=True), sb.jv(w, eps, lin=True))
    assert not np.any(sb.jv(np.zeros(pa.N), eps, lin=True))
    sa.close(); sb.close(); pa.close(); pb.close()


def test_lin_kernel_every_workgroup_shape():
    """Every k_micro_lin shape compiled for 21 x 21 cells (one, two, four wavefronts per
    cell) gives the same product: TSR_LIN_BS, read at each problem's first tuning, pins
    the shape."""
    outs = []
    for bs in (64, 128, 256):
        os.environ["TSR_LIN_BS"] = str(bs)
        try:
            p = T.Problem.config(F.CFG_PLAIN, 7, 21, 0, 0)
            s = T.Solver(p)
            rng = np.random.default_rng(5)
            u = p.initial() + 0.8 * rng.standard_normal(p.N)
            v = rng.uniform(-1.0, 1.0, p.N)
            s.set_state(u, 0.0)
            outs.append(s.jv(v, 1e-7, lin=True))
            s.close(); p.close()
        finally:
            os.environ.pop("TSR_LIN_BS", None)
    print(f"[lin-kernel] shapes 64/128/256: max |diff| {max(np.max(np.abs(o - outs[0])) for o in outs):.3e}",
          flush=True)
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])


BUDGET_CASES = [("plain r21", F.CFG_PLAIN, 7, 21, 0), ("ragged", F.CFG_RAGGED, 9, 0, 0),
                ("large discs r31", F.CFG_LARGE, 5, 31, 0)]


@pytest.mark.parametrize("name,cfg,macro_n,micro_r,n_scen", BUDGET_CASES, ids=[c[0] for c in BUDGET_CASES])
def test_lin_kernel_every_register_budget(name, cfg, macro_n, micro_r, n_scen):
    """k_micro_lin under both register budgets the tuner times (the compiler's, and 8
    waves per SIMD, whose larger shapes spill a few registers to scratch), for each
    multi-wavefront shape: the products must not change, bit for bit, and must equal
    k_micro's.  TSR_LIN_BS and TSR_MICRO_WPE, read at a problem's first tuning, pin the
    shape and the budget."""
    print(f"[lin-budgets] {name}", flush=True)
    ref = None
    rng = np.random.default_rng(17)
    for bs, wpe, lin_kernel in ((0, 1, False), (128, 1, True), (128, 8, True), (256, 1, True), (256, 8, True)):
        env = {"TSR_MICRO_WPE": str(wpe)}
        if bs:
            env["TSR_LIN_BS"] = str(bs)
        os.environ.update(env)
        try:
            p = _problem(cfg, macro_n, micro_r, n_scen, lin_kernel)
            if ref is None:
                u = p.initial() + 0.8 * rng.standard_normal(p.N)
                v = rng.uniform(-1.0, 1.0, p.N)
            s = T.Solver(p)
            s.set_state(u, 0.0)
            out = s.jv(v, 1e-7, lin=True)
            s.close(); p.close()
        finally:
            for k in env:
                os.environ.pop(k, None)
        if ref is None:
            ref = out
            assert np.all(np.isfinite(ref))
            continue
        nd = int(np.sum(out != ref))
        print(f"[lin-budgets] {name}: bs {bs}, waves/EU {wpe}: {nd} entries differ from k_micro", flush=True)
        assert nd == 0
