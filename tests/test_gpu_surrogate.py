"""The learned scheme search on small adjoint scenes: the surrogate trained on
adjoint runs predicts the per-field sensitivity coefficients better than the
training mean, the surrogate path is faster than the adjoint path, and with
the adjoint as judge every allocation meets the error bound (falling back to
the adjoint allocation when the prediction misses)."""
import time

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

N, GRID, STEPS, CK = 2048, 64, 24, 8


@pytest.fixture(scope="module")
def data():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from autoquantizer_amd.surrogate import make_dataset
    X, G, R, P = make_dataset(40, seed=1, n=N, n_grid=GRID, nsteps=STEPS, checkpoint_every=CK)
    return X, G, R, P


def test_dataset_is_finite(data):
    X, G, R, P = data
    assert np.all(np.isfinite(X)) and np.all(np.isfinite(G)) and np.all(G >= 0)
    assert X.shape == (40, 24) and G.shape == (40, 12)


def test_surrogate_beats_the_mean_and_the_adjoint_judges(data):
    from autoquantizer_amd.surrogate import Surrogate, field_targets, random_scene, search_scheme_surrogate
    from autoquantizer_amd.search import search_scheme
    from autoquantizer_amd.surrogate import make_solver
    X, G, R, P = data
    Y = np.array([field_targets(g) for g in G])
    model = Surrogate(seed=0)
    model.fit(X[:32], Y[:32], steps=1500)
    pred = model.predict(X[32:])
    err = np.mean((pred - Y[32:]) ** 2)
    base = np.mean((Y[:32].mean(0) - Y[32:]) ** 2)
    print("held-out MSE of log10 sums: surrogate", err, "training mean", base)
    assert err < base

    # held-out scenes: surrogate path vs adjoint path, the adjoint judging
    rng = np.random.default_rng(1)
    for _ in range(32):  # the scenes of the training set
        random_scene(rng, N, GRID)
    t_sur, t_adj = [], []
    for k in range(8):
        st, E, dt = random_scene(rng, N, GRID)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ref = search_scheme(lambda: make_solver(N, GRID, E, dt), st, STEPS, tol_rel=1e-3, ref=1.0,
                            checkpoint_every=CK)
        t_adj.append(time.perf_counter() - t0)
        r = search_scheme_surrogate(model, st, E, dt, GRID, STEPS, ref=1.0, tol_rel=1e-3, judge_g=ref["grads"])
        t_sur.append(r["seconds"])
        # after judging (and falling back when needed) the bound holds under the adjoint's sums
        from autoquantizer_amd.solvers import modelled_error
        G_true = ref["grads"] * r["ranges"] ** 2
        assert modelled_error(G_true, r["bits"] - 1) <= 1e-6 * (1 + 1e-9)
    print("search seconds: adjoint", np.mean(t_adj), "surrogate", np.mean(t_sur))
    assert np.mean(t_sur) < np.mean(t_adj)
