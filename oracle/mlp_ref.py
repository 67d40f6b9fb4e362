"""Oracle of the surrogate MLP -- TEST INFRASTRUCTURE ONLY.

float64 numpy restatement of the network of autoquantizer_amd/csrc/aq_mlp.hip
(in -> 64 tanh -> 64 tanh -> 64 tanh -> out, parameters W1, b1, ..., W4, b4
in one flat array, y = W x + b) and of the gradient of mean((y - t)^2); the
gradient is pinned by central finite differences in tests/test_mlp_oracle.py.
"""
import numpy as np

IN, H, OUT = 24, 64, 4
SHAPES = [(H, IN), (H,), (H, H), (H,), (H, H), (H,), (OUT, H), (OUT,)]
NPARAM = sum(int(np.prod(s)) for s in SHAPES)


def unflatten(p):
    out, o = [], 0
    for s in SHAPES:
        k = int(np.prod(s))
        out.append(np.asarray(p[o:o + k], dtype=np.float64).reshape(s))
        o += k
    return out


def forward(p, x):
    """(y [B, OUT], activations [h1, h2, h3])."""
    W1, b1, W2, b2, W3, b3, W4, b4 = unflatten(p)
    x = np.asarray(x, dtype=np.float64)
    h1 = np.tanh(x @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    h3 = np.tanh(h2 @ W3.T + b3)
    return h3 @ W4.T + b4, (h1, h2, h3)


def loss_and_grad(p, x, t, weight_decay=0.0):
    """mean((y - t)^2) and the gradient of mean((y - t)^2) + weight_decay *
    (sum of squared weight-matrix entries) with respect to the flat
    parameters (the returned loss excludes the decay term)."""
    W1, b1, W2, b2, W3, b3, W4, b4 = unflatten(p)
    x = np.asarray(x, dtype=np.float64)
    y, (h1, h2, h3) = forward(p, x)
    e = y - t
    L = float(np.mean(e * e))
    d4 = 2.0 * e / e.size
    d3 = (d4 @ W4) * (1 - h3 * h3)
    d2 = (d3 @ W3) * (1 - h2 * h2)
    d1 = (d2 @ W2) * (1 - h1 * h1)
    wd2 = 2.0 * weight_decay
    g = [d1.T @ x + wd2 * W1, d1.sum(0), d2.T @ h1 + wd2 * W2, d2.sum(0), d3.T @ h2 + wd2 * W3, d3.sum(0),
         d4.T @ h3 + wd2 * W4, d4.sum(0)]
    return L, np.concatenate([a.ravel() for a in g])


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam step (Kingma & Ba 2015, Algorithm 1)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
