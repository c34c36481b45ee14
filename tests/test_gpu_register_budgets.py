"""k_micro under every register budget the tuner may be given (amdgpu_waves_per_eu 1, 6
and 8; 6 and 8 make the compiler spill registers to scratch in some modes): G, the record
launch and the linearised FD product must not depend on the budget, bit for bit, on the
uniform-grid kernel and on the generic-template kernel (TSR_GENERIC_MICRO), for
node-hosted cells with a centred disc and with random per-cell discs.  TSR_MICRO_WPE,
read when a problem is first tuned, pins the budget (the tuner otherwise skips budgets
whose code spills)."""
import os

import numpy as np
import pytest

import twoscalerich_amd as T
from twoscalerich_amd import _ffi as F

pytestmark = pytest.mark.gpu

CASES = [("small r11", F.CFG_PLAIN_SMALL, 0, 0), ("large discs r31", F.CFG_LARGE, 5, 31)]


def _outputs(cfg, macro_n, micro_r, generic, wpe):
    env = {"TSR_MICRO_WPE": str(wpe)}
    if generic:
        env["TSR_GENERIC_MICRO"] = "1"
    os.environ.update(env)
    try:
        p = T.Problem.config(cfg, macro_n, micro_r, 0, 0)
        s = T.Solver(p)
        rng = np.random.default_rng(21)
        u = p.initial() + 0.8 * rng.standard_normal(p.N)
        v = rng.uniform(-1.0, 1.0, p.N)
        g, rc = s.gfunc(u)
        assert rc == 0
        s.set_state(u, 0.0)
        jl = s.jv(v, 1e-7, lin=True)      # record launch, then the linearised product
        jf = s.jv(v, 1e-7, lin=False)     # FD product of the full G
        s.close()
        p.close()
        return g, jl, jf
    finally:
        for k in env:
            os.environ.pop(k, None)


@pytest.mark.parametrize("generic", [False, True], ids=["uniform-grid kernel", "generic kernel"])
@pytest.mark.parametrize("name,cfg,macro_n,micro_r", CASES, ids=[c[0] for c in CASES])
def test_register_budget_does_not_change_results(name, cfg, macro_n, micro_r, generic):
    print(f"[budgets] {name}, generic={generic}", flush=True)
    ref = _outputs(cfg, macro_n, micro_r, generic, 1)
    for wpe in (6, 8):
        out = _outputs(cfg, macro_n, micro_r, generic, wpe)
        for label, a, b in zip(("G", "linearised FD product", "full FD product"), ref, out):
            nd = int(np.sum(a != b))
            print(f"[budgets] {name} generic={generic} waves/EU {wpe}: {label} differs at {nd} entries", flush=True)
            assert nd == 0
