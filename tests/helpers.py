"""Shared helpers of the test suite: seeded states and the reference fixtures."""
import ctypes as C
import os

import numpy as np

from twoscalerich_amd import _ffi as F

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def seeded_state(problem, seed, lo=-5.0, hi=-0.3):
    """Initial state plus a seeded field in [lo, hi] so every branch of G is exercised."""
    rng = np.random.default_rng(seed)
    return rng.uniform(lo, hi, problem.N)


def rel_inf(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def soil(Ks, tr, ts, al, n):
    s = F.soil_struct()
    s.Ksat, s.thetar, s.thetas, s.alpha, s.n, s.m = Ks, tr, ts, al, n, 1.0 - 1.0 / n
    return s


def reference_problem(macro="macro_structured_10by10.msh", micro="microB_unstructured_10by10.msh",
                      bc=(1, 3, 2, 2, 2), model=3, h0=-2.0, host=F.HOST_ELEMENT, boundary_flux=1.222222e-5 * 3600):
    """The reference's shipped test inputs (tests/Richards_Equation/), copied as data
    fixtures under tests/golden/.  Soils of problem_options.txt:8-17 in m and hours."""
    import twoscalerich_amd as T
    A = soil(0.044, 0.058, 0.41, 7.3, 1.89)
    B = soil(0.044e-3, 0.058, 0.41, 7.3, 1.89)
    return T.Problem.reference(os.path.join(GOLDEN, macro), os.path.join(GOLDEN, micro),
                               os.path.join(GOLDEN, "Keff.txt"), A, B, boundary_flux, list(bc), model, h0, host)
