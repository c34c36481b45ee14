"""Learned step controller: training data from the solver's own step records, training
with the HIP kernels of csrc/mlp.hip, and attachment to a solver.

The reference's step law (src/set_exprem_options.cpp:20-26 options; the EXPREM driver
include/main.h:29-32) is the classical one,
    tau_next = tau * min(MaxStepIncFac, max(MinStepDecFac, SafetyFac * err^(-1/3))),
which only sees the previous step's error.  The learned controller replaces the
proposal (step size and Krylov budget) at the start of a step; EXPREM's own error
estimate still accepts or rejects the attempt, and the classical law is the fallback
whenever the network's output is unusable.

Inputs: the 16 features the solver writes into every tsr_step_record (order in
include/tsr/tsr.h).  Targets, from the same record:
    y0 = log10(tau_first * err_first^(-1/3))   the step at which the first attempt's
                                               error would have been exactly 1
                                               (local error ~ tau^3: exprb32 is of order 2)
    y1 = m                                     the Krylov dimension the step needed
Rejected first attempts (err_first > 1) are training samples like any other: their
target is the smaller step they should have tried.  Features are standardised with
the training set's mean and standard deviation; targets are standardised with a
power-of-two scale, so folding the target scale back into the output layer is exact in
BF16.
"""
import ctypes as C
import os

import numpy as np

from . import _ffi as F
from . import lib, last_error

N_IN, N_H, N_OUT = 16, 64, 2
N_PARAMS = N_H * N_IN + N_H + N_H * N_H + N_H + N_OUT * N_H + N_OUT
O_W3 = N_H * N_IN + N_H + N_H * N_H + N_H
O_B3 = O_W3 + N_OUT * N_H
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "controller_weights.npz")

REC_FIELDS = ("t", "tau", "tau_proposed", "err", "tau_first", "err_first", "m", "mD", "rejects", "fevals",
              "controller", "m_cap")


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def records_to_arrays(recs):
    """tsr_step_record list -> (features [n,16] float32, dict of per-record fields)."""
    n = len(recs)
    X = np.zeros((n, N_IN), np.float32)
    cols = {k: np.zeros(n) for k in REC_FIELDS}
    for i, r in enumerate(recs):
        X[i] = np.frombuffer(r.feat, dtype=np.float32, count=N_IN)
        for k in REC_FIELDS:
            cols[k][i] = getattr(r, k)
    return X, cols


def targets(cols):
    """(targets [n,2] float32, usable mask) from record fields."""
    tf, ef, m = cols["tau_first"], cols["err_first"], cols["m"]
    ok = np.isfinite(ef) & (ef > 0) & (tf > 0) & (m > 0)
    y0 = np.where(ok, np.log10(np.where(ok, tf, 1.0)) - np.log10(np.where(ok, ef, 1.0)) / 3.0, 0.0)
    return np.stack([y0, m], 1).astype(np.float32), ok


def init_params(seed=0):
    """He-uniform weights, zero biases, small output layer (FP32)."""
    rng = np.random.default_rng(seed)
    P = np.zeros(N_PARAMS, np.float32)
    o = 0
    for fan_in, rows in ((N_IN, N_H), (N_H, N_H), (N_H, N_OUT)):
        lim = np.sqrt(6.0 / fan_in)
        P[o:o + rows * fan_in] = rng.uniform(-lim, lim, rows * fan_in)
        o += rows * fan_in + rows
    P[O_W3:O_B3] *= 0.1
    return P


class Mlp:
    """Handle of the device network (tsr_mlp_* entry points)."""

    def __init__(self, params=None):
        self.L = lib()
        assert self.L.tsr_mlp_param_count() == N_PARAMS
        self.h = self.L.tsr_mlp_create()
        if not self.h:
            raise RuntimeError("tsr_mlp_create failed")
        self.set(init_params() if params is None else params)

    def set(self, P):
        P = np.ascontiguousarray(P, np.float32)
        if self.L.tsr_mlp_set_params(self.h, _fp(P)):
            raise RuntimeError("tsr_mlp_set_params failed")

    def get(self):
        P = np.zeros(N_PARAMS, np.float32)
        if self.L.tsr_mlp_get_params(self.h, _fp(P)):
            raise RuntimeError("tsr_mlp_get_params failed")
        return P

    def forward(self, X):
        X = np.ascontiguousarray(X, np.float32)
        out = np.zeros((len(X), 2), np.float32)
        if self.L.tsr_mlp_forward(self.h, _fp(X), len(X), _fp(out)):
            raise RuntimeError("tsr_mlp_forward failed")
        return out

    def train(self, X, T, epochs, batch=256, lr=1e-3, w=(1.0, 1.0), seed=1):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        hist = np.zeros(epochs, np.float32)
        if self.L.tsr_mlp_train(self.h, _fp(X), _fp(T), len(X), epochs, batch, lr, w[0], w[1], seed, _fp(hist)):
            raise RuntimeError("tsr_mlp_train failed")
        return hist

    def close(self):
        if self.h:
            self.L.tsr_mlp_free(self.h)
            self.h = None


def _pow2(x):
    return float(2.0 ** np.round(np.log2(max(float(x), 1e-6))))


def fit(X, T, epochs=60, batch=256, lr=1e-3, seed=1, val_frac=0.1):
    """Train on (X raw features, T raw targets).  Returns a model dict
    {params, feat_mean, feat_std, hist, val} with the target scaling folded into the
    output layer, so the device network maps standardised features to raw targets."""
    X = np.asarray(X, np.float32)
    T = np.asarray(T, np.float32)
    rng = np.random.default_rng(seed)
    idx = rng.permutation(len(X))
    nv = int(len(X) * val_frac)
    va, tr = idx[:nv], idx[nv:]
    mu = X[tr].mean(0)
    sd = np.maximum(X[tr].std(0), 1e-3)
    tmu = T[tr].mean(0)
    tsd = np.array([_pow2(T[tr, 0].std()), _pow2(T[tr, 1].std())], np.float32)
    Xs = (X - mu) / sd
    Ts = (T - tmu) / tsd
    net = Mlp(init_params(seed))
    try:
        hist = net.train(Xs[tr], Ts[tr], epochs, batch, lr, (1.0, 1.0), seed)
        P = net.get()
        # fold the target scaling into W3, b3 (tsd are powers of two: exact in BF16)
        W3 = P[O_W3:O_B3].reshape(N_OUT, N_H)
        W3 *= tsd[:, None]
        P[O_B3:] = P[O_B3:] * tsd + tmu
        net.set(P)
        val = None
        if nv:
            Y = net.forward(Xs[va])
            val = {"mae_log10_tau": float(np.abs(Y[:, 0] - T[va, 0]).mean()),
                   "mae_m": float(np.abs(Y[:, 1] - T[va, 1]).mean()),
                   "n": int(nv)}
    finally:
        net.close()
    return {"params": P, "feat_mean": mu.astype(np.float32), "feat_std": sd.astype(np.float32),
            "hist": hist, "val": val}


def save(model, path=WEIGHTS, **meta):
    np.savez(path, params=model["params"], feat_mean=model["feat_mean"], feat_std=model["feat_std"],
             hist=model["hist"], **{k: np.asarray(v) for k, v in meta.items()})


def load(path=WEIGHTS):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def attach(solver, model=None):
    """Switch a Solver to the learned controller (committed weights by default)."""
    m = load() if model is None else model
    P = np.ascontiguousarray(m["params"], np.float32)
    mu = np.ascontiguousarray(m["feat_mean"], np.float32)
    sd = np.ascontiguousarray(m["feat_std"], np.float32)
    L = lib()
    if L.tsr_solver_set_mlp(solver.h, _fp(P), _fp(mu), _fp(sd)):
        raise RuntimeError("tsr_solver_set_mlp failed: " + last_error())
    if L.tsr_solver_set_controller(solver.h, 1):
        raise RuntimeError("tsr_solver_set_controller failed")


def collect(solver, tend, max_steps, chunk=500, progress=None):
    """Run a solver to tend (or max_steps) and return its step records."""
    S = solver.problem.view.contents.n_scen
    out = []
    n = 0
    while n < max_steps:
        k = min(chunk, max_steps - n)
        rc, st, recs = solver.run(tend, k, max_records=k * S)
        out.extend(recs[:st.steps * S])
        n += st.steps
        if progress:
            progress(st, n)
        if rc:
            raise RuntimeError("integration failed: " + last_error())
        if st.steps == 0 or st.t >= tend:
            break
    return out
