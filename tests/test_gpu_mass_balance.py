"""Global mass balance of an integration on the GPU, stated for the plain and the
learned + batched paths on the small workload over its whole horizon (t in [0, 0.1] d):
    |W(t_end) - W(0) - integral of F dt| / |W(t_end) - W(0)|,
W the water held (two-scale: macro CVs and, for model 3, the cells' interiors weighted
by the area each cell stands for) and F the inflow through the boundary (Dirichlet top
edge); both from the CPU oracle at every accepted GPU state, the integral by the
trapezoid rule over the accepted steps.  The semi-discrete model conserves water
exactly (the oracle's water-balance identity, tests/test_oracle_pins.py), so the number
measures the time integration (and the trapezoid rule's own O(tau^2) error)."""
import numpy as np
import pytest

import twoscalerich_amd as T
from twoscalerich_amd import _ffi as F
from twoscalerich_amd import controller as CT
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def _mass_balance(p, o, path, learned):
    s = T.Solver(p, p.options())
    s.set_path(path)
    if learned:
        CT.attach(s)
    u = p.initial()
    s.reset()
    s.set_state(u, 0.0)
    tend = p.tend()
    W0, F0, _ = o.water_balance(u)
    Fp, t, integral, n = F0, 0.0, 0.0, 0
    while t < tend:
        rc, _ = s.step(tend)
        assert rc == 0, T.last_error()
        u, tn = s.get_state()
        W, Fn, _ = o.water_balance(u)
        integral += 0.5 * (tn - t) * (Fp + Fn)
        Fp, t, n = Fn, tn, n + 1
    s.close()
    dW = W - W0
    return abs(dW - integral) / abs(dW), dW, n


@pytest.mark.parametrize("learned", [False, True], ids=["plain", "learned_batched"])
def test_mass_balance_small_workload(learned):
    p = T.Problem.config(F.CFG_PLAIN_SMALL)
    o = Oracle(p.view)
    path = F.PATH_BATCHED if learned else F.PATH_PLAIN
    err, dW, n = _mass_balance(p, o, path, learned)
    print(f"[mass] {'learned+batched' if learned else 'plain'}: {n} steps, water gained {dW:.6e}, "
          f"global mass-balance error {err:.3e}", flush=True)
    assert dW > 0 and err <= 1e-3
    o.close()
    p.close()
