"""CPU restatement of the radial-surrogate network (TEST INFRASTRUCTURE ONLY).

The reference has no learned radial function; this module is the float64
numpy statement of the network the library trains on the GPU, so the HIP
forward, tangent, loss gradient, Adam step and output-layer refit can be
checked against it.  Only tests and the benchmark's CPU baseline import it.

Network for one species pair (H hidden units, O outputs, tanh):
    a1 = W1 x + b1,  h1 = tanh(a1)                      x in [-1, 1]
    a2 = W2 h1 + b2, h2 = tanh(a2)
    y  = W3 h2 + b3
and its input tangent dy/dx:
    d1 = (1 - h1^2) W1,  d2 = (1 - h2^2) (W2 d1),  dy = W3 d2.
Loss over a batch of B samples (O outputs per sample):
    L = 1/(O B) sum_s sum_o [ wv_o (y - t)^2 + wd_o (dy - dt)^2 ].
Parameter vector layout: W1[H], b1[H], W2[H][H], b2[H], W3[O][H], b3[O].

The gradient is pinned by central finite differences in tests/test_oracle_mlp.py.
"""
from __future__ import annotations

import numpy as np


def n_params(H: int, O: int) -> int:
    return 3 * H + H * H + O * H + O


def unpack(p: np.ndarray, H: int, O: int):
    o = 0
    W1 = p[o:o + H]; o += H
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + O * H].reshape(O, H); o += O * H
    b3 = p[o:o + O]; o += O
    return W1, b1, W2, b2, W3, b3


def init_params(H: int, O: int, rng: np.random.Generator) -> np.ndarray:
    """Initial weights: first layer spread over the input interval, Xavier-scaled others."""
    W1 = rng.uniform(-4.0, 4.0, H)
    b1 = rng.uniform(-4.0, 4.0, H)
    W2 = rng.normal(0.0, 1.0 / np.sqrt(H), (H, H))
    b2 = np.zeros(H)
    W3 = rng.normal(0.0, 1.0 / np.sqrt(H), (O, H))
    b3 = np.zeros(O)
    return np.concatenate([W1, b1, W2.ravel(), b2, W3.ravel(), b3])


def forward(p: np.ndarray, x: np.ndarray, H: int, O: int):
    """y (B, O), dy/dx (B, O) and the activations the gradient needs."""
    W1, b1, W2, b2, W3, b3 = unpack(p, H, O)
    h1 = np.tanh(x[:, None] * W1[None] + b1[None])
    d1 = (1.0 - h1 * h1) * W1[None]
    a2 = h1 @ W2.T + b2[None]
    t2 = d1 @ W2.T
    h2 = np.tanh(a2)
    d2 = (1.0 - h2 * h2) * t2
    y = h2 @ W3.T + b3[None]
    dy = d2 @ W3.T
    return y, dy, (h1, d1, h2, d2, t2)


def loss_and_grad(p, x, t, dt, wv, wd, H, O):
    """Loss and dL/dp by reverse differentiation of forward() (value and tangent paths)."""
    W1, b1, W2, b2, W3, b3 = unpack(p, H, O)
    B = x.shape[0]
    y, dy, (h1, d1, h2, d2, t2) = forward(p, x, H, O)
    sc = 1.0 / (O * B)
    ry, rd = y - t, dy - dt
    L = sc * float((wv[None] * ry * ry).sum() + (wd[None] * rd * rd).sum())
    ey = 2.0 * sc * wv[None] * ry          # dL/dy
    ed = 2.0 * sc * wd[None] * rd          # dL/d(dy)
    gW3 = ey.T @ h2 + ed.T @ d2
    gb3 = ey.sum(0)
    gh2 = ey @ W3
    gd2 = ed @ W3
    u2 = 1.0 - h2 * h2
    gt2 = u2 * gd2                          # d2 = u2 * t2
    gh2 = gh2 - 2.0 * h2 * t2 * gd2
    ga2 = u2 * gh2
    gW2 = ga2.T @ h1 + gt2.T @ d1
    gb2 = ga2.sum(0)
    gh1 = ga2 @ W2
    gd1 = gt2 @ W2
    u1 = 1.0 - h1 * h1
    gW1 = (u1 * gd1).sum(0)                 # d1 = u1 * W1
    gh1 = gh1 - 2.0 * h1 * W1[None] * gd1
    ga1 = u1 * gh1
    gW1 = gW1 + (ga1 * x[:, None]).sum(0)
    gb1 = ga1.sum(0)
    g = np.concatenate([gW1, gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return L, g


def adam_step(p, g, m, v, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, Algorithm 1) with bias correction; step counts from 1."""
    m = beta1 * m + (1.0 - beta1) * g
    v = beta2 * v + (1.0 - beta2) * g * g
    mh = m / (1.0 - beta1 ** step)
    vh = v / (1.0 - beta2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def refit_output(p, x, t, dt, wd, H, O, ridge=1e-10):
    """Exact least-squares output layer for the fixed hidden layers (value + weighted tangent rows)."""
    _, _, (h1, d1, h2, d2, t2) = forward(p, x, H, O)
    B = x.shape[0]
    Ht = np.concatenate([h2, np.ones((B, 1))], 1)
    Dt = np.concatenate([d2, np.zeros((B, 1))], 1)
    AH, AD = Ht.T @ Ht, Dt.T @ Dt
    RH, RD = Ht.T @ t, Dt.T @ dt
    q = p.copy()
    W1, b1, W2, b2, W3, b3 = unpack(q, H, O)
    for j in range(O):
        A = AH + wd[j] * AD
        A = A + ridge * np.trace(A) / (H + 1) * np.eye(H + 1)
        w = np.linalg.solve(A, RH[:, j] + wd[j] * RD[:, j])
        W3[j] = w[:H]
        b3[j] = w[H]
    return q
