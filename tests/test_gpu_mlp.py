"""The surrogate's HIP training and inference kernels (aq_mlp_forward,
aq_mlp_backward, aq_adam_step) against the f64 numpy oracle (oracle/mlp_ref.py,
its gradient pinned by finite differences) and a plain torch fp32 reference."""
import numpy as np
import pytest
import torch

from oracle import mlp_ref as M

pytestmark = pytest.mark.gpu


def model(seed=0):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from autoquantizer_amd.surrogate import Surrogate
    return Surrogate(seed=seed)


def torch_forward(p, x):
    W1, b1, W2, b2, W3, b3, W4, b4 = [torch.as_tensor(a, dtype=torch.float32) for a in M.unflatten(p)]
    x = torch.as_tensor(x, dtype=torch.float32)
    h = torch.tanh(x @ W1.T + b1)
    h = torch.tanh(h @ W2.T + b2)
    h = torch.tanh(h @ W3.T + b3)
    return (h @ W4.T + b4).numpy()


def test_forward_matches_oracle_and_torch():
    m = model(1)
    rng = np.random.default_rng(0)
    x = rng.standard_normal((300, M.IN))
    xt = torch.as_tensor(x, dtype=torch.float32, device="cuda")
    y, act = m._forward(xt)
    p = m.params.cpu().numpy().astype(np.float64)
    yo, (h1, h2, h3) = M.forward(p, x.astype(np.float32).astype(np.float64))
    np.testing.assert_allclose(y.cpu().numpy(), yo, rtol=1e-4, atol=1e-5)
    np.testing.assert_allclose(act.cpu().numpy(), np.hstack([h1, h2, h3]), rtol=1e-4, atol=1e-5)
    np.testing.assert_allclose(y.cpu().numpy(), torch_forward(p, x), rtol=1e-4, atol=1e-5)


@pytest.mark.parametrize("wd", [0.0, 3e-3])
def test_backward_matches_oracle(wd):
    m = model(2)
    rng = np.random.default_rng(3)
    x = rng.standard_normal((257, M.IN)).astype(np.float32)
    t = rng.standard_normal((257, M.OUT)).astype(np.float32)
    L, g = m.loss_and_grad(torch.as_tensor(x, device="cuda"), torch.as_tensor(t, device="cuda"), wd)
    Lo, go = M.loss_and_grad(m.params.cpu().numpy().astype(np.float64), x.astype(np.float64), t.astype(np.float64),
                             weight_decay=wd)
    assert L == pytest.approx(Lo, rel=1e-5)
    g = g.cpu().numpy()
    assert np.abs(g - go).max() <= 1e-4 * np.abs(go).max() + 1e-7


def test_adam_step_matches_oracle():
    m = model(4)
    rng = np.random.default_rng(5)
    p0 = m.params.cpu().numpy().astype(np.float64)
    mm = np.zeros(M.NPARAM)
    vv = np.zeros(M.NPARAM)
    p = p0
    for t in range(1, 4):
        g = rng.standard_normal(M.NPARAM)
        m.adam_step(torch.as_tensor(g, dtype=torch.float32, device="cuda"), 1e-2)
        p, mm, vv = M.adam(p, g.astype(np.float32).astype(np.float64), mm, vv, t, 1e-2)
    np.testing.assert_allclose(m.params.cpu().numpy(), p, rtol=1e-5, atol=1e-6)


def test_training_fits_a_smooth_map():
    """Full-batch training on a smooth 24 -> 4 map lowers the loss 10x and
    generalises to held-out inputs."""
    m = model(6)
    rng = np.random.default_rng(7)
    A = rng.standard_normal((M.IN, M.OUT)) / np.sqrt(M.IN)
    f = lambda x: np.sin(x @ A) + 0.1 * (x @ A) ** 2  # noqa: E731
    X = rng.standard_normal((512, M.IN))
    Xh = rng.standard_normal((64, M.IN))
    hist = m.fit(X, f(X), steps=1500, lr=3e-3)
    assert hist[-1] < 0.1 * hist[0]
    err = np.mean((m.predict(Xh) - f(Xh)) ** 2)
    base = np.mean((f(X).mean(0) - f(Xh)) ** 2)
    assert err < 0.3 * base
