"""Float64 CPU restatement of the learned plan selector's MLP -- test
infrastructure only (imported by ``tests/`` as the checker of the HIP
inference and training kernels and of their host replay; the library never
calls it).

The reference has no learned selector (it ranks candidates with its
analytic model, ``src/plan.cpp:721-789``), so this restates the network the
library defines: inputs x (50 normalised features), two ReLU layers of
width 64, one linear output,

    z1 = W1 x + b1,  a1 = relu(z1)
    z2 = W2 a1 + b2, a2 = relu(z2)
    out = w3 . a2 + b3

with the loss ``inv_m * sum_s w_s (out_s - y_s)^2`` (w = 1 when unweighted) and the Adam update
(Kingma & Ba, 2015, Algorithm 1) with bias corrections ``c1 = 1/(1 -
beta1^t)``, ``c2 = 1/(1 - beta2^t)`` applied as
``p -= lr * (m c1) / (sqrt(v c2) + eps)``.

Parameter vector layout: W1[64][50], b1[64], W2[64][64], b2[64], w3[64], b3.

Pinned in ``tests/test_selector_cpu.py`` by: a hand-worked two-unit network
(forward), central finite differences of the loss (backward), and Adam's
closed-form first step ``-lr * g / (|g| + eps)`` (update).
"""
import numpy as np

F = 50
H = 64
P = H * F + H + H * H + H + H + 1


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + H * F].reshape(H, F); o += H * F
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    w3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def pack(W1, b1, W2, b2, w3, b3):
    return np.concatenate([np.ravel(W1), b1, np.ravel(W2), b2, w3, [b3]])


def forward(p, x):
    """Outputs for x of shape [n, F]."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    x = np.asarray(x, dtype=np.float64)
    a1 = np.maximum(x @ W1.T + b1, 0.0)
    a2 = np.maximum(a1 @ W2.T + b2, 0.0)
    return a2 @ w3 + b3


def loss(p, x, y, inv_m, w=None):
    out = forward(p, x)
    w = 1.0 if w is None else np.asarray(w, dtype=np.float64)
    return inv_m * np.sum(w * (out - np.asarray(y, dtype=np.float64)) ** 2)


def grad(p, x, y, inv_m, w=None):
    """(gradient of loss(p, x, y, inv_m, w) w.r.t. p, weighted sum of squared errors)."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    z1 = x @ W1.T + b1
    a1 = np.maximum(z1, 0.0)
    z2 = a1 @ W2.T + b2
    a2 = np.maximum(z2, 0.0)
    out = a2 @ w3 + b3
    err = out - y
    w = np.ones_like(err) if w is None else np.asarray(w, dtype=np.float64)
    d = 2.0 * w * err * inv_m                   # d loss / d out
    gw3 = a2.T @ d
    gb3 = d.sum()
    dz2 = np.outer(d, w3) * (z2 > 0)
    gW2 = dz2.T @ a1
    gb2 = dz2.sum(axis=0)
    dz1 = (dz2 @ W2) * (z1 > 0)
    gW1 = dz1.T @ x
    gb1 = dz1.sum(axis=0)
    return pack(gW1, gb1, gW2, gb2, gw3, gb3), float(np.sum(w * err * err))


def adam(p, m, v, g, lr, beta1, beta2, t, eps):
    """One Adam step at step number t (1-based); returns (p, m, v)."""
    p = np.asarray(p, dtype=np.float64)
    m = beta1 * np.asarray(m, dtype=np.float64) + (1 - beta1) * g
    v = beta2 * np.asarray(v, dtype=np.float64) + (1 - beta2) * g * g
    c1 = 1.0 / (1.0 - beta1 ** t)
    c2 = 1.0 / (1.0 - beta2 ** t)
    return p - lr * (m * c1) / (np.sqrt(v * c2) + eps), m, v
