"""The MLP oracle's gradient against central finite differences (CPU)."""
import numpy as np

from oracle import mlp_ref as M


def test_gradient_matches_central_differences():
    rng = np.random.default_rng(0)
    p = rng.standard_normal(M.NPARAM) * 0.2
    x = rng.standard_normal((7, M.IN))
    t = rng.standard_normal((7, M.OUT))
    L, g = M.loss_and_grad(p, x, t)
    for q in rng.choice(M.NPARAM, 40, replace=False):
        e = np.zeros(M.NPARAM)
        e[q] = 1e-6
        fd = (M.loss_and_grad(p + e, x, t)[0] - M.loss_and_grad(p - e, x, t)[0]) / 2e-6
        assert abs(fd - g[q]) <= 1e-7 + 1e-6 * abs(g[q])


def test_parameter_layout():
    assert M.NPARAM == 10180
    p = np.arange(M.NPARAM, dtype=np.float64)
    W1, b1, W2, b2, W3, b3, W4, b4 = M.unflatten(p)
    assert W1[1, 0] == 24 and b1[0] == 64 * 24 and b4[-1] == M.NPARAM - 1


def test_adam_first_step_moves_by_lr():
    p = np.zeros(5)
    g = np.array([1.0, -2.0, 0.5, 3.0, -0.1])
    p1, m, v = M.adam(p, g, np.zeros(5), np.zeros(5), 1, 1e-3)
    np.testing.assert_allclose(p1, -1e-3 * np.sign(g), rtol=1e-6)


def test_weight_decay_gradient_matches_central_differences():
    """The decay term's gradient against differences of the full objective
    mean((y - t)^2) + wd * sum of squared weights (biases excluded)."""
    rng = np.random.default_rng(1)
    p = rng.standard_normal(M.NPARAM) * 0.2
    x = rng.standard_normal((5, M.IN))
    t = rng.standard_normal((5, M.OUT))
    wd = 0.05
    W = M.unflatten(np.ones(M.NPARAM))
    is_weight = np.concatenate([a.ravel() * (1.0 if a.ndim == 2 else 0.0) for a in W])

    def objective(q):
        return M.loss_and_grad(q, x, t)[0] + wd * float((is_weight * q * q).sum())

    _, g = M.loss_and_grad(p, x, t, weight_decay=wd)
    for q in list(rng.choice(M.NPARAM, 30, replace=False)) + [M.NPARAM - 1, 24 * 64]:
        e = np.zeros(M.NPARAM)
        e[q] = 1e-6
        fd = (objective(p + e) - objective(p - e)) / 2e-6
        assert abs(fd - g[q]) <= 1e-7 + 1e-6 * abs(g[q]), q
