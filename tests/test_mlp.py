"""The learned controller's network: the CPU oracle pinned (CPU), then the HIP kernels
(forward, backward, Adam, training loop) against it (GPU)."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle_mlp as OM


def _data(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, OM.N_IN)).astype(np.float32)
    T = rng.standard_normal((n, 2)).astype(np.float32)
    return X, T


# ------------------------------------------------------------------ oracle pins (CPU)

def test_bf16_round_matches_torch_cast():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(1)
    a = np.concatenate([rng.standard_normal(20000).astype(np.float32) * 10.0 ** rng.integers(-6, 6, 20000),
                        np.array([0.0, -0.0, 1.0, 1.00390625, 1.01171875, 3.4e38, -3.4e38, 1e-40, np.inf, -np.inf],
                                 dtype=np.float32)]).astype(np.float32)
    ref = torch.from_numpy(a).to(torch.bfloat16).to(torch.float32).numpy()
    got = OM.bf16_round(a)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    # ties go to even: 1 + 2^-8 lies halfway between 1 and 1 + 2^-7
    assert OM.bf16_round(np.float32(1.0 + 2.0 ** -8)) == np.float32(1.0)
    assert OM.bf16_round(np.float32(1.0 + 3 * 2.0 ** -8)) == np.float32(1.0 + 2.0 ** -6)


def test_gelu_closed_form_and_derivative():
    assert OM.gelu(np.array(0.0)) == 0.0
    # GELU(z) - GELU(-z) = z
    z = np.linspace(-6, 6, 101)
    assert np.allclose(OM.gelu(z) - OM.gelu(-z), z, atol=1e-14)
    # Phi(1) = 0.841344746068543 (standard normal CDF)
    assert abs(OM.gelu(np.array(1.0)) - 0.841344746068543) < 1e-14
    h = 1e-6
    fd = (OM.gelu(z + h) - OM.gelu(z - h)) / (2 * h)
    assert np.allclose(OM.gelu_d(z), fd, atol=1e-8)


def test_oracle_gradient_matches_central_differences():
    P = OM.init(3).astype(np.float64)
    X, T = _data(7, 4)
    g, L = OM.grad(P, X, T, 1.0, 0.3, 0.5, rounding=False)
    assert abs(L - OM.loss(P, X, T, 1.0, 0.3, 0.5, rounding=False)) < 1e-12 * abs(L)
    rng = np.random.default_rng(5)
    idx = np.concatenate([rng.choice(OM.N_PARAMS, 60, replace=False), np.arange(OM.O_B3, OM.N_PARAMS)])
    for i in idx:
        h = 1e-6
        Pp = P.copy(); Pp[i] += h
        Pm = P.copy(); Pm[i] -= h
        # forward in FP64 with the perturbed FP64 parameters (no FP32 cast)
        fd = (_loss64(Pp, X, T) - _loss64(Pm, X, T)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def _loss64(P, X, T):
    W1 = P[OM.O_W1:OM.O_B1].reshape(OM.N_H, OM.N_IN); b1 = P[OM.O_B1:OM.O_W2]
    W2 = P[OM.O_W2:OM.O_B2].reshape(OM.N_H, OM.N_H); b2 = P[OM.O_B2:OM.O_W3]
    W3 = P[OM.O_W3:OM.O_B3].reshape(OM.N_OUT, OM.N_H); b3 = P[OM.O_B3:]
    Y = OM.gelu(OM.gelu(X @ W1.T + b1) @ W2.T + b2) @ W3.T + b3
    E = Y - T
    return 0.5 * float(np.sum(1.0 * E[:, 0] ** 2 + 0.3 * E[:, 1] ** 2))


def test_oracle_adam_first_step_is_signed_lr():
    # bias-corrected Adam: the first step moves every coordinate by lr * g/(|g| + eps')
    P = np.zeros(5)
    g = np.array([1.0, -2.0, 1e-3, -1e-4, 3.0])
    P1, m, v = OM.adam(P, g, np.zeros(5), np.zeros(5), 1, lr=0.01, eps=1e-12)
    assert np.allclose(P1, -0.01 * np.sign(g), rtol=1e-6)


def test_oracle_fits_a_linear_map():
    # training with the oracle's own gradient reduces the loss by orders of magnitude
    rng = np.random.default_rng(7)
    X = rng.standard_normal((256, OM.N_IN))
    A = rng.standard_normal((OM.N_IN, 2)) * 0.3
    T = X @ A
    P = OM.init(8).astype(np.float64)
    m = np.zeros_like(P); v = np.zeros_like(P)
    L0 = OM.loss(P, X, T, scale=1 / 256, rounding=False)
    for t in range(1, 401):
        g, _ = OM.grad(P, X, T, scale=1 / 256, rounding=False)
        P, m, v = OM.adam(P, g, m, v, t, lr=3e-3)
    assert OM.loss(P, X, T, scale=1 / 256, rounding=False) < 0.02 * L0


# ------------------------------------------------------------------ HIP kernels (GPU)

def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


@pytest.fixture(scope="module")
def mlp():
    import twoscalerich_amd as T
    L = T.lib()
    assert L.tsr_mlp_param_count() == OM.N_PARAMS
    h = L.tsr_mlp_create()
    assert h
    yield L, h
    L.tsr_mlp_free(h)


@pytest.mark.gpu
def test_gpu_forward_matches_oracle(mlp):
    L, h = mlp
    P = OM.init(11)
    assert L.tsr_mlp_set_params(h, _fp(P)) == 0
    for n in (1, 3, 64, 1000):
        X, _ = _data(n, n)
        out = np.zeros((n, 2), np.float32)
        assert L.tsr_mlp_forward(h, _fp(X), n, _fp(out)) == 0
        ref = OM.forward(P, X)
        assert np.allclose(out, ref, rtol=2e-5, atol=2e-5 * np.abs(ref).max()), (n, np.abs(out - ref).max())


@pytest.mark.gpu
def test_gpu_gradient_matches_oracle(mlp):
    L, h = mlp
    P = OM.init(12)
    assert L.tsr_mlp_set_params(h, _fp(P)) == 0
    for n in (1, 5, 300, 4096):
        X, T = _data(n, 100 + n)
        g = np.zeros(OM.N_PARAMS, np.float32)
        loss = np.zeros(1, np.float32)
        assert L.tsr_mlp_grad(h, _fp(X), _fp(T), n, 1.0, 0.25, 1.0 / n, _fp(g), _fp(loss)) == 0
        gr, Lr = OM.grad(P, X, T, 1.0, 0.25, 1.0 / n)
        assert abs(loss[0] - Lr) <= 1e-5 * abs(Lr)
        scale = np.abs(gr).max()
        assert np.abs(g - gr).max() <= 2e-5 * scale, (n, np.abs(g - gr).max(), scale)
        # rerun: the reduction is deterministic
        g2 = np.zeros_like(g)
        assert L.tsr_mlp_grad(h, _fp(X), _fp(T), n, 1.0, 0.25, 1.0 / n, _fp(g2), None) == 0
        assert np.array_equal(g, g2)


@pytest.mark.gpu
def test_gpu_adam_matches_oracle(mlp):
    L, h = mlp
    P = OM.init(13)
    assert L.tsr_mlp_set_params(h, _fp(P)) == 0
    X, T = _data(200, 21)
    Pr = P.astype(np.float64)
    m = np.zeros(OM.N_PARAMS); v = np.zeros(OM.N_PARAMS)
    for t in range(1, 4):
        g = np.zeros(OM.N_PARAMS, np.float32)
        assert L.tsr_mlp_grad(h, _fp(X), _fp(T), 200, 1.0, 1.0, 1 / 200, _fp(g), None) == 0
        assert L.tsr_mlp_adam(h, 1e-3) == 0
        # the oracle takes the device's gradient, so only the update rule is compared
        Pr, m, v = OM.adam(Pr, g.astype(np.float64), m, v, t, lr=1e-3)
        got = np.zeros(OM.N_PARAMS, np.float32)
        assert L.tsr_mlp_get_params(h, _fp(got)) == 0
        assert np.abs(got - Pr).max() <= 1e-6, np.abs(got - Pr).max()
        Pr = got.astype(np.float64)


@pytest.mark.gpu
def test_gpu_training_reduces_loss(mlp):
    L, h = mlp
    rng = np.random.default_rng(9)
    X = rng.standard_normal((4096, OM.N_IN)).astype(np.float32)
    A = (rng.standard_normal((OM.N_IN, 2)) * 0.3).astype(np.float32)
    T = (np.tanh(X @ A) + 0.1 * X[:, :2] ** 2).astype(np.float32)
    P = OM.init(14)
    assert L.tsr_mlp_set_params(h, _fp(P)) == 0
    hist = np.zeros(30, np.float32)
    assert L.tsr_mlp_train(h, _fp(X), _fp(T), 4096, 30, 256, 1e-3, 1.0, 1.0, 5, _fp(hist)) == 0
    assert hist[-1] < 0.05 * hist[0], hist
    got = np.zeros(OM.N_PARAMS, np.float32)
    L.tsr_mlp_get_params(h, _fp(got))
    # the trained network's loss, recomputed by the oracle, matches the device's last epoch
    assert OM.loss(got, X, T, scale=1 / 4096) < 1.5 * hist[-1]
