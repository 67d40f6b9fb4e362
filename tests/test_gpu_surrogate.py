"""The learned scheme search: the surrogate trained on adjoint runs predicts
the per-field sensitivity coefficients better than the training mean, the
surrogate path is faster than the adjoint path, and with the adjoint as judge
every allocation meets the error bound (falling back to the adjoint
allocation when the prediction misses)."""
import time

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

N, GRID, STEPS, CK = 2048, 64, 24, 8
TOL = 1e-4  # error bound on the energy loss, relative to the loss


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


def test_training_on_the_shipped_dataset_beats_the_mean():
    """Trained by the HIP kernels on the 512 training scenes of the shipped
    adjoint dataset (65,536 particles, 256^2, 256 substeps each), the network
    predicts the 64 held-out scenes' per-field log10 sums with less than half
    the squared error of the training mean."""
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import os
    from autoquantizer_amd.surrogate import Surrogate, field_targets
    d = np.load(os.path.join(os.path.dirname(__file__), "..", "autoquantizer_amd", "data",
                             "surrogate_elastic2d_dataset.npz"), allow_pickle=False)
    X, Y = d["X"], np.array([field_targets(g) for g in d["G"]])
    model = Surrogate(seed=0)
    hist = model.fit(X[:512], Y[:512], steps=3000)
    assert hist[-1] < 0.2 * hist[0]
    err = np.mean((model.predict(X[512:]) - Y[512:]) ** 2)
    base = np.mean((Y[:512].mean(0) - Y[512:]) ** 2)
    print("held-out MSE of log10 sums: surrogate", err, "training mean", base)
    assert err < 0.5 * base


def test_adjoint_judges_and_falls_back(data):
    """On small held-out scenes every surrogate allocation is checked by a
    dithered quantized run of the scene; it either meets the bound there or
    the search falls back to exactly the adjoint search's allocation.  The
    adjoint's sums judge every allocation.  A forced miss (predictions 10^12
    too small: about 20 fraction bits fewer) fails its run and falls back.
    The surrogate's prediction + range run + solve is faster than the
    adjoint search (the check run is timed at full size by the benchmark)."""
    from autoquantizer_amd.surrogate import Surrogate, field_targets, random_scene, search_scheme_surrogate
    from autoquantizer_amd.search import search_scheme
    from autoquantizer_amd.surrogate import make_solver
    X, G, R, P = data
    Y = np.array([field_targets(g) for g in G])
    model = Surrogate(seed=0)
    model.fit(X[:32], Y[:32], steps=1500)
    rng = np.random.default_rng(1)
    for _ in range(32):  # the scenes of the training set
        random_scene(rng, N, GRID)
    t_sur, t_adj = [], []
    for k in range(8):
        st, E, dt = random_scene(rng, N, GRID)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ref = search_scheme(lambda: make_solver(N, GRID, E, dt), st, STEPS, tol_rel=TOL, checkpoint_every=CK)
        torch.cuda.synchronize()
        t_adj.append(time.perf_counter() - t0)
        r = search_scheme_surrogate(model, st, E, dt, GRID, STEPS, tol_rel=TOL, judge=ref, check_runs=1,
                                    fallback=lambda: ref)
        t_sur.append(r["seconds"])
        assert r["met"] in (True, False)
        assert r["check_passed"] or (r["fell_back"] and np.array_equal(r["bits"], ref["bits"]))
        if k == 0:
            model.margin = np.full(4, -12.0)
            miss = search_scheme_surrogate(model, st, E, dt, GRID, STEPS, tol_rel=TOL, judge=ref, check_runs=1,
                                           fallback=lambda: ref)
            model.margin = np.zeros(4)
            assert not miss["met"] and not miss["check_passed"]
            assert miss["fell_back"]
            np.testing.assert_array_equal(miss["bits"], ref["bits"])
    print("search seconds: adjoint", np.mean(t_adj), "surrogate", np.mean(t_sur))
    assert np.mean(t_sur) < np.mean(t_adj)


def test_trained_surrogate_on_heldout_scenes():
    """The shipped network (trained by autoquantizer_amd.train_surrogate on
    512 adjoint scenes of 65,536 particles, 256^2, 256 substeps) on the 64
    held-out scenes of the shipped dataset: its allocations meet the error
    bound under the adjoint's sums on at least 90 % of them (the adjoint is
    the fallback for the rest), and on a held-out scene the surrogate's
    prediction + range run + solve is at least 3x faster than the adjoint
    search (both run the range pass; the surrogate replaces the checkpointed
    forward and the backward pass, which cost about 3-4 range passes)."""
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import os
    from autoquantizer_amd.surrogate import Surrogate, random_scene, search_scheme_surrogate, make_solver
    from autoquantizer_amd.search import search_scheme
    from autoquantizer_amd.train_surrogate import allocations
    data = os.path.join(os.path.dirname(__file__), "..", "autoquantizer_amd", "data")
    model = Surrogate.load(os.path.join(data, "surrogate_elastic2d.npz"))
    d = np.load(os.path.join(data, "surrogate_elastic2d_dataset.npz"), allow_pickle=False)
    X, G, R = d["X"][512:], d["G"][512:], d["R"][512:]
    met = np.mean([m for _, _, m in allocations(model.predict(X), G, R, 1e-3, model.margin)])
    print("held-out allocations meeting the bound:", met)
    assert met >= 0.9
    rng = np.random.default_rng(2024)
    for _ in range(512):
        random_scene(rng, 65536, 256)
    st, E, dt = random_scene(rng, 65536, 256)
    search_scheme_surrogate(model, st, E, dt, 256, 8, tol_rel=TOL)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ref = search_scheme(lambda: make_solver(65536, 256, E, dt), st, 256, tol_rel=TOL, checkpoint_every=16)
    torch.cuda.synchronize()
    t_adj = time.perf_counter() - t0
    r = search_scheme_surrogate(model, st, E, dt, 256, 256, tol_rel=TOL, judge=ref)
    print("search seconds: adjoint", t_adj, "surrogate", r["seconds"], "met", r["met"])
    assert t_adj >= 3 * r["seconds"]
