"""Learned step controller: training data from the solver's own step records, training
with the HIP kernels of csrc/mlp.hip, and attachment to a solver.

The reference's step law (src/set_exprem_options.cpp:20-26 options; the EXPREM driver
include/main.h:29-32) is the classical one,
    tau_next = tau * min(MaxStepIncFac, max(MinStepDecFac, SafetyFac * err^(-1/3))),
which only sees the previous step's error.  On these problems the error estimate is
rarely what limits the step: the Krylov stopping rule is.  The classical law grows the
step by MaxStepIncFac until the Arnoldi process reaches its cap, halves it, and grows
it again, and the Krylov dimension needed rises steeply along that saw-tooth (on the
plain workload from about 6 to 30 while the step grows by less than 2x), so most of the
work goes into the largest steps.

The learned controller proposes, at the start of each step, the step size with the
lowest expected cost per unit of simulated time, and the Krylov dimension it expects
that step to need (the window of the deferred stopping-rule checks).  EXPREM's own error
estimate still accepts or rejects each attempt, the Krylov rule still halves a step
whose basis reaches the cap, and the classical law is the fallback whenever the
network's output is unusable.

Training data: step records (tsr_step_record: the 16 features at the step start, the
first attempt's step and error, the Krylov dimensions) from runs driven by the
exploration controller (controller mode 2), which random-walks the step size so that
nearby steps sample the cost curve.  Labels, per step k, from the records of the same
scenario within +-`window` steps (the state changes little over that span):
    for every candidate step size tau_c among them, the Krylov dimensions m(tau_c) and
    mD(tau_c) are the medians over records with |log10 tau - log10 tau_c| <= 0.05; the
    cost of a step is (in units of one linearised RHS evaluation)
        cost(m, mD) = c0 + m + mD + r (m^2 + 8 m + mD^2 + 5 mD)
    (RHS evaluations, plus r per basis column read in the Gram-Schmidt passes and the
    combinations; c0 = 4.7, r = 0.037 from the measured kernel times); the
    label is the tau_c with the largest tau_c / cost, among candidates whose first
    attempt's error was at most `err_max`:
        y0 = log10 tau_c,  y1 = m(tau_c).
Features are standardised with the training set's mean and standard deviation;
targets with a power-of-two scale, so folding the target scale back into the output
layer is exact in BF16.
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
              "controller", "m_cap", "wall_ms")


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


# cost model of one step on the batched path, in units of one linearised RHS evaluation
# (plain workload on MI355X: 94 us per linearised G with its macro nodes, 7 us per
# basis-column read in the DCGS2 passes, 0.78 ms per step for G(u) with the expansion
# record, G(U2), the step combinations, the controller and the host round trips)
COST_C0 = 8.3      # per-step work outside the Arnoldi iterations
COST_R = 0.075     # one basis-column read in a Gram-Schmidt pass / one RHS evaluation


def step_cost(m, mD, c0=COST_C0, r=COST_R):
    return c0 + m + mD + r * (m * m + 8.0 * m + mD * mD + 5.0 * mD)


def label_sequence(cols, window=20, err_max=0.7, c0=COST_C0, r=COST_R):
    """Labels for one scenario's consecutive records (dict of arrays, in step order).
    Returns (targets [n,2] float32, usable mask)."""
    tf, ef = cols["tau_first"], cols["err_first"]
    m, mD = cols["m"], cols["mD"]
    n = len(tf)
    ok = np.isfinite(ef) & (tf > 0) & (m > 0)
    lt = np.where(ok, np.log10(np.where(tf > 0, tf, 1.0)), 0.0)
    Y = np.zeros((n, 2), np.float32)
    good = np.zeros(n, bool)
    for k in range(n):
        if not ok[k]:
            continue
        lo, hi = max(0, k - window), min(n, k + window + 1)
        idx = np.arange(lo, hi)
        idx = idx[ok[idx]]
        if len(idx) < 5:
            continue
        best, by = -np.inf, None
        for c in idx:
            if ef[c] > err_max:
                continue
            near = idx[np.abs(lt[idx] - lt[c]) <= 0.05]
            mm = float(np.median(m[near]))
            mdd = float(np.median(mD[near]))
            thr = tf[c] / step_cost(mm, mdd, c0, r)
            if thr > best:
                best, by = thr, (lt[c], mm)
        if by is not None:
            Y[k] = by
            good[k] = True
    return Y, good


def records_by_scenario(recs, S):
    """Split a step-major record list (S records per solver step) into per-scenario
    sequences, dropping the all-zero records of finished scenarios."""
    seqs = []
    for q in range(S):
        rq = [r for r in recs[q::S] if r.tau > 0]
        if rq:
            seqs.append(rq)
    return seqs


def dataset(seqs, **kw):
    """(X [n,16], Y [n,2]) over per-scenario record sequences."""
    Xs, Ys = [], []
    for rq in seqs:
        X, cols = records_to_arrays(rq)
        Y, good = label_sequence(cols, **kw)
        Xs.append(X[good])
        Ys.append(Y[good])
    return np.concatenate(Xs), np.concatenate(Ys)


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


# ------------------------------------------------------------------ training pipeline
# (workload, seed, max solver steps): the plain workload from t = 0 to t_end, the ragged
# workload from t = 0, and three batches of 64 scenarios with sampled soils (192 soil
# draws).
PLAN = ((F.CFG_PLAIN, 0, 40000), (F.CFG_RAGGED, 0, 10000),
        (F.CFG_SCENARIOS, 11, 2500), (F.CFG_SCENARIOS, 12, 2500), (F.CFG_SCENARIOS, 13, 2500))


def generate(plan=PLAN, path=F.PATH_BATCHED, log=print):
    """Run the exploration controller over the plan; returns arrays
    {X [n,16], fields [n,len(REC_FIELDS)], seq [n]} (seq = scenario sequence id)."""
    from . import Problem, Solver
    Xs, Fs, Ss = [], [], []
    sid = 0
    for cfg, seed, nmax in plan:
        p = Problem.config(cfg, seed=seed)
        s = Solver(p, p.options())
        s.set_path(path)
        if s.set_controller(2):
            raise RuntimeError("exploration controller unavailable")
        lib().tsr_solver_set_seed(s.h, 7919 * (seed + 1) + cfg)
        s.reset()
        s.set_state(p.initial(), 0.0)
        S = p.view.contents.n_scen

        def prog(st, n):
            log(f"[controller] workload {cfg} seed {seed}: {n} steps, t = {st.t:.4e}")

        recs = collect(s, p.tend(), nmax, progress=prog)
        for rq in records_by_scenario(recs, S):
            X, cols = records_to_arrays(rq)
            Xs.append(X)
            Fs.append(np.stack([cols[k] for k in REC_FIELDS], 1))
            Ss.append(np.full(len(rq), sid))
            sid += 1
        s.close()
        p.close()
    return {"X": np.concatenate(Xs), "fields": np.concatenate(Fs), "seq": np.concatenate(Ss)}


def dataset_from_arrays(d, **kw):
    """Labelled (X, Y) from generate()'s arrays."""
    X, Fd, seq = d["X"], d["fields"], d["seq"]
    Xs, Ys = [], []
    for q in np.unique(seq):
        sel = seq == q
        # files written before later fields were added hold a prefix of REC_FIELDS
        cols = {k: Fd[sel, i] for i, k in enumerate(REC_FIELDS[:Fd.shape[1]])}
        Y, good = label_sequence(cols, **kw)
        Xs.append(X[sel][good])
        Ys.append(Y[good])
    return np.concatenate(Xs), np.concatenate(Ys)


def summarize(workload, controller, path, tend=0.0, max_steps=10 ** 7, log=print):
    """Run one workload to t_end with a controller ("classical", "learned", "explore") and
    a solver path; return a dict of step statistics (the controller's view of the run)."""
    import time
    from . import Problem, Solver
    p = Problem.config(workload)
    s = Solver(p, p.options())
    s.set_path(path)
    if controller == "learned":
        attach(s)
    elif controller == "explore":
        s.set_controller(2)
        lib().tsr_solver_set_seed(s.h, 5)
    tend = tend if tend > 0 else p.tend()
    s.reset()
    s.set_state(p.initial(), 0.0)
    s.run(min(tend, 1e-9), 2)                      # kernel tuning outside the timed run
    s.reset()
    s.set_state(p.initial(), 0.0)
    lib().tsr_device_sync()
    t0 = time.perf_counter()
    recs = collect(s, tend, max_steps, progress=lambda st, n: log(f"[controller] {n} steps, t = {st.t:.4e}"))
    lib().tsr_device_sync()
    wall = time.perf_counter() - t0
    u, t = s.get_state()
    s.close()
    X, c = records_to_arrays(recs)
    m = c["m"].astype(int)
    prev = c["tau"][:-1]
    ratio = c["tau_proposed"][1:] / np.where(prev > 0, prev, 1.0)
    out = {"workload": workload, "controller": controller, "path": path, "t": t, "wall_s": wall,
           "steps": len(recs), "rejects": int(c["rejects"].sum()), "fevals": int(c["fevals"].sum()),
           "fevals_per_step": float(c["fevals"].mean()), "mean_m": float(m.mean()), "mean_mD": float(c["mD"].mean()),
           "steps_at_krylov_cap": int(np.sum(m >= c["m_cap"])),
           "krylov_halvings": int(np.sum((c["tau_first"] < c["tau_proposed"] * 0.999))),
           "proposals_capped_1.1": float(np.mean(np.abs(ratio - 1.1) < 1e-9)) if len(ratio) else 0.0,
           "err_quantiles_10_50_90": [float(q) for q in np.quantile(c["err"], [0.1, 0.5, 0.9])],
           "median_tau": float(np.median(c["tau"])), "m_histogram": np.bincount(m).tolist()}
    p.close()
    return out, u


def main(argv=None):
    import argparse
    import json
    import time
    ap = argparse.ArgumentParser(description="learned step controller: data generation and training")
    ap.add_argument("--data", default="", help="records file (.npz) to write (generate) or read (train)")
    ap.add_argument("--generate", action="store_true")
    ap.add_argument("--train", action="store_true")
    ap.add_argument("--summarize", action="store_true", help="run a workload and print step statistics")
    ap.add_argument("--workload", type=int, default=F.CFG_PLAIN)
    ap.add_argument("--controller", default="classical", choices=["classical", "learned", "explore"])
    ap.add_argument("--path", type=int, default=F.PATH_PLAIN)
    ap.add_argument("--tend", type=float, default=0.0)
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--max-samples", type=int, default=60000)
    a = ap.parse_args(argv)
    if a.summarize:
        out, _ = summarize(a.workload, a.controller, a.path, a.tend)
        print(json.dumps(out), flush=True)
    if a.generate:
        t0 = time.time()
        d = generate()
        np.savez_compressed(a.data, **d)
        print(f"[controller] {len(d['X'])} records in {time.time() - t0:.1f} s -> {a.data}", flush=True)
    if a.train:
        with np.load(a.data, allow_pickle=False) as z:
            d = {k: z[k] for k in z.files}
        X, Y = dataset_from_arrays(d)
        rng = np.random.default_rng(0)
        if len(X) > a.max_samples:
            keep = rng.choice(len(X), a.max_samples, replace=False)
            X, Y = X[keep], Y[keep]
        t0 = time.time()
        model = fit(X, Y, epochs=a.epochs, lr=a.lr)
        print(f"[controller] trained on {len(X)} samples in {time.time() - t0:.1f} s; loss "
              f"{model['hist'][0]:.4f} -> {model['hist'][-1]:.4f}; validation {model['val']}", flush=True)
        save(model, a.out, n_samples=len(X), epochs=a.epochs, lr=a.lr)


if __name__ == "__main__":
    main()
