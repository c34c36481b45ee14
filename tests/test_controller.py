"""Learned step controller: labelling of the solver's step records (CPU), training on the
device and the controller's gain over the classical law (GPU)."""
import numpy as np
import pytest

from twoscalerich_amd import controller as CT


def _cols(tau, m, mD=None, err=None):
    n = len(tau)
    c = {k: np.zeros(n) for k in CT.REC_FIELDS}
    c["tau_first"] = np.asarray(tau, float)
    c["m"] = np.asarray(m, float)
    c["mD"] = np.zeros(n) + 2.0 if mD is None else np.asarray(mD, float)
    c["err_first"] = np.full(n, 1e-3) if err is None else np.asarray(err, float)
    return c


def test_cost_model_counts_rhs_and_basis_reads():
    # no Krylov work: only the fixed part; each Arnoldi iteration costs one RHS
    # evaluation plus the Gram-Schmidt column reads
    assert CT.step_cost(0, 0) == CT.COST_C0
    assert CT.step_cost(1, 0) - CT.step_cost(0, 0) == pytest.approx(1 + CT.COST_R * 9)
    assert CT.step_cost(10, 0, c0=0, r=0) == 10


def test_label_prefers_cheap_steps_over_krylov_wall():
    # saw-tooth as the classical law produces it: the step grows while the Krylov
    # dimension climbs from 6 to 29, then the step is halved (steps 25 % apart here, so
    # the 0.05-decade smoothing band holds one step size at a time)
    tau, m = [], []
    for cyc in range(6):
        t = 1.0
        for mk in (6, 13, 15, 19, 22, 29):
            tau.append(t)
            m.append(mk)
            t *= 1.25
    Y, good = CT.label_sequence(_cols(tau, m), window=20)
    assert good.all()
    # tau / cost is largest at the bottom of the saw-tooth (m = 6, tau = 1)
    assert np.allclose(Y[:, 0], 0.0)
    assert np.allclose(Y[:, 1], 6.0)


def test_label_respects_error_limit():
    # the cheapest step per unit time has a large error estimate: the next one wins
    tau = [1.0, 2.0, 4.0] * 5
    m = [5, 6, 7] * 5
    err = [1e-3, 1e-3, 0.95] * 5
    Y, good = CT.label_sequence(_cols(tau, m, err=err), window=20, err_max=0.7)
    assert good.all()
    assert np.allclose(Y[:, 0], np.log10(2.0))
    assert np.allclose(Y[:, 1], 6.0)


def test_label_uses_local_window_only():
    # two regimes 100 steps apart: each step is labelled from its own neighbourhood
    tau = [1.0, 1.2] * 25 + [10.0, 12.0] * 25
    m = [4, 40] * 25 + [40, 4] * 25
    Y, good = CT.label_sequence(_cols(tau, m), window=10)
    assert np.allclose(Y[:20, 0], 0.0)
    assert np.allclose(Y[-20:, 0], np.log10(12.0))


def test_records_by_scenario_drops_finished():
    class R:
        def __init__(self, tau):
            self.tau = tau
    recs = [R(1.0), R(2.0), R(3.0), R(0.0), R(5.0), R(0.0)]   # 2 scenarios x 3 steps
    seqs = CT.records_by_scenario(recs, 2)
    assert [[r.tau for r in q] for q in seqs] == [[1.0, 3.0, 5.0], [2.0]]


def test_dataset_from_arrays_reads_record_files_with_fewer_fields():
    # a records file written before the last field existed holds a prefix of the fields;
    # it labels the same as the full layout (the labels use tau_first, err_first, m, mD)
    tau, m = [], []
    for cyc in range(4):
        t = 1.0
        for mk in (6, 13, 19, 29):
            tau.append(t)
            m.append(mk)
            t *= 1.25
    cols = _cols(tau, m)
    n = len(tau)
    full = np.stack([cols[k] for k in CT.REC_FIELDS], 1)
    X = np.arange(n * CT.N_IN, dtype=np.float32).reshape(n, CT.N_IN)
    seq = np.zeros(n, np.int64)
    Xa, Ya = CT.dataset_from_arrays({"X": X, "fields": full, "seq": seq}, window=20)
    Xb, Yb = CT.dataset_from_arrays({"X": X, "fields": full[:, :-1], "seq": seq}, window=20)
    assert np.array_equal(Xa, Xb) and np.array_equal(Ya, Yb)
    assert np.allclose(Ya[:, 1], 6.0)


# ------------------------------------------------------------------ GPU

@pytest.mark.gpu
def test_fit_learns_a_smooth_map():
    rng = np.random.default_rng(4)
    X = rng.standard_normal((8000, 16)).astype(np.float32) * 3 + 1
    Y = np.stack([-4.0 + 0.3 * np.tanh(X[:, 0]) + 0.1 * X[:, 1], 8.0 + 2.0 * np.sin(X[:, 2])], 1).astype(np.float32)
    # the oracle's own Adam on this data reaches 0.022 / 0.15 after 40 epochs at this rate
    model = CT.fit(X, Y, epochs=40, lr=3e-3)
    print(f"[controller] fit: {model['val']}", flush=True)
    assert model["val"]["mae_log10_tau"] < 0.05
    assert model["val"]["mae_m"] < 0.5


@pytest.mark.gpu
def test_committed_weights_load_and_predict():
    import twoscalerich_amd as T
    from twoscalerich_amd import _ffi as F
    m = CT.load()
    assert m["params"].shape == (CT.N_PARAMS,)
    net = CT.Mlp(m["params"])
    X = (np.zeros((3, 16), np.float32) - m["feat_mean"]) / m["feat_std"]
    Y = net.forward(X)
    net.close()
    assert np.all(np.isfinite(Y))
    # attached to a solver, every proposal comes from the network
    p = T.Problem.config(F.CFG_PLAIN_SMALL)
    s = T.Solver(p, p.options())
    CT.attach(s)
    s.reset()
    s.set_state(p.initial(), 0.0)
    rc, st, recs = s.run(p.tend(), 20, max_records=20)
    assert rc == 0
    assert all(r.controller == 1 for r in recs[1:st.steps])
    s.close()
    p.close()
