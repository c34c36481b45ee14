"""CPU oracle of the learned step controller's network.  TEST INFRASTRUCTURE ONLY:
imported by tests/ as the checker of the HIP kernels in twoscalerich_amd/csrc/mlp.hip;
the product never calls it.

The network (the learned controller of the two-scale Richards EXPREM driver; the
reference has no such component, its step law is the classical one of
src/set_exprem_options.cpp:20-26):
    x (16) -> z1 = W1 x + b1 -> a1 = GELU(z1) (64) -> z2 = W2 a1 + b2 -> a2 = GELU(z2) (64)
           -> y = W3 a2 + b3 (2)
with GELU(z) = z Phi(z) (exact, erf form).  The weights used in the products are the
bfloat16 round-to-nearest-even of FP32 master weights; biases stay FP32.  This oracle
computes in FP64 from those values, so the device's FP32 accumulation must agree to
FP32 rounding (about 1e-6 relative).

Parameter vector layout (FP32, row-major): W1[64][16], b1[64], W2[64][64], b2[64],
W3[2][64], b3[2].

Loss of a batch:  L = scale * sum_n (w0 (y_n0 - t_n0)^2 + w1 (y_n1 - t_n1)^2).
Gradients with respect to the masters are those with respect to the rounded weights
(straight-through).  Adam is Kingma & Ba (2015), Algorithm 1, bias-corrected.

Pins: tests/test_mlp.py checks bf16_round against torch's float32 -> bfloat16 cast,
gelu against the closed form at special points and its derivative against central
differences, and the analytic gradient against central differences of the loss in
FP64 (with the rounding switched off, so the difference quotient is smooth).
"""
import numpy as np
from scipy.special import erf

N_IN, N_H, N_OUT = 16, 64, 2
O_W1 = 0
O_B1 = O_W1 + N_H * N_IN
O_W2 = O_B1 + N_H
O_B2 = O_W2 + N_H * N_H
O_W3 = O_B2 + N_H
O_B3 = O_W3 + N_OUT * N_H
N_PARAMS = O_B3 + N_OUT


def bf16_round(a):
    """float32 -> nearest bfloat16 (ties to even), returned as float32; inf/nan unchanged."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    special = (u & 0x7F800000) == 0x7F800000
    r = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    r = np.where(special, u, r).astype(np.uint32)
    return r.view(np.float32)


def unpack(P, rounding=True):
    P = np.asarray(P, dtype=np.float32)
    rw = bf16_round if rounding else (lambda a: a)
    W1 = rw(P[O_W1:O_B1]).astype(np.float64).reshape(N_H, N_IN)
    b1 = P[O_B1:O_W2].astype(np.float64)
    W2 = rw(P[O_W2:O_B2]).astype(np.float64).reshape(N_H, N_H)
    b2 = P[O_B2:O_W3].astype(np.float64)
    W3 = rw(P[O_W3:O_B3]).astype(np.float64).reshape(N_OUT, N_H)
    b3 = P[O_B3:].astype(np.float64)
    return W1, b1, W2, b2, W3, b3


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def forward(P, X, rounding=True):
    """Y[n][2] for inputs X[n][16] (already standardised)."""
    W1, b1, W2, b2, W3, b3 = unpack(P, rounding)
    X = np.asarray(X, dtype=np.float64)
    z1 = X @ W1.T + b1
    a1 = gelu(z1)
    z2 = a1 @ W2.T + b2
    a2 = gelu(z2)
    return a2 @ W3.T + b3


def loss(P, X, T, w0=1.0, w1=1.0, scale=1.0, rounding=True):
    Y = forward(P, X, rounding)
    E = Y - np.asarray(T, dtype=np.float64)
    return scale * float(np.sum(w0 * E[:, 0] ** 2 + w1 * E[:, 1] ** 2))


def grad(P, X, T, w0=1.0, w1=1.0, scale=1.0, rounding=True):
    """(gradient [N_PARAMS] in FP64, loss) by reverse-mode differentiation."""
    W1, b1, W2, b2, W3, b3 = unpack(P, rounding)
    X = np.asarray(X, dtype=np.float64)
    T = np.asarray(T, dtype=np.float64)
    z1 = X @ W1.T + b1
    a1 = gelu(z1)
    z2 = a1 @ W2.T + b2
    a2 = gelu(z2)
    Y = a2 @ W3.T + b3
    E = Y - T
    L = scale * float(np.sum(w0 * E[:, 0] ** 2 + w1 * E[:, 1] ** 2))
    dY = 2.0 * scale * E * np.array([w0, w1])
    g = np.zeros(N_PARAMS)
    g[O_W3:O_B3] = (dY.T @ a2).ravel()
    g[O_B3:] = dY.sum(0)
    dz2 = (dY @ W3) * gelu_d(z2)
    g[O_W2:O_B2] = (dz2.T @ a1).ravel()
    g[O_B2:O_W3] = dz2.sum(0)
    dz1 = (dz2 @ W2) * gelu_d(z1)
    g[O_W1:O_B1] = (dz1.T @ X).ravel()
    g[O_B1:O_W2] = dz1.sum(0)
    return g, L


def adam(P, g, m, v, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (step t >= 1); returns new (P, m, v) in FP64."""
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    mh = m / (1.0 - b1 ** t)
    vh = v / (1.0 - b2 ** t)
    return P - lr * mh / (np.sqrt(vh) + eps), m, v


def init(seed=0):
    """He-uniform weights, zero biases (FP32)."""
    rng = np.random.default_rng(seed)
    P = np.zeros(N_PARAMS, dtype=np.float32)
    for o, fan_in, rows in ((O_W1, N_IN, N_H), (O_W2, N_H, N_H), (O_W3, N_H, N_OUT)):
        lim = np.sqrt(6.0 / fan_in)
        P[o:o + rows * fan_in] = rng.uniform(-lim, lim, rows * fan_in).astype(np.float32)
    P[O_W3:O_B3] *= 0.1
    return P
