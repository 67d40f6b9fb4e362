"""Trains the scheme-search surrogate on adjoint runs and evaluates it.

    python -m autoquantizer_amd.train_surrogate [--train 512] [--test 64]
        [--particles 65536] [--grid 256] [--steps 256] [--out autoquantizer_amd/data]

Scenes: 2D elastic blocks of the adjoint configuration (random position,
size 0.1-0.3, |v0| <= 3, E log-uniform in [5e3, 5e4], seeded).  Each scene
gets a range run and a checkpointed adjoint run (the method's own kernels);
the MLP is trained by the library's HIP kernels on the features -> per-field
log10 sensitivity coefficients.  The held-out scenes measure the prediction
error, how often the predicted allocation meets the error bound under the
adjoint's sums (with a safety margin calibrated on a validation split of the
training scenes), and the search time of both paths.  Writes
surrogate_elastic2d.npz (weights, standardisation, margin) and
surrogate_elastic2d_eval.json.
"""
import argparse
import json
import os
import sys
import time

import numpy as np


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=512)
    ap.add_argument("--test", type=int, default=64)
    ap.add_argument("--particles", type=int, default=65536)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--steps", type=int, default=256)
    ap.add_argument("--checkpoint-every", type=int, default=16)
    ap.add_argument("--tol-rel", type=float, default=1e-3)
    ap.add_argument("--fit-steps", type=int, default=3000)
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "data"))
    a = ap.parse_args(argv)

    import torch
    from .search import search_scheme
    from .solvers import LagrangianEB, modelled_error
    from .surrogate import (Surrogate, component_sums, field_targets, make_dataset, make_solver, random_scene,
                            search_scheme_surrogate)

    log = lambda m: print(m, flush=True)  # noqa: E731
    t0 = time.perf_counter()
    X, G, R, P = make_dataset(a.train + a.test, seed=2024, n=a.particles, n_grid=a.grid, nsteps=a.steps,
                              checkpoint_every=a.checkpoint_every, log=log)
    t_data = time.perf_counter() - t0
    Y = np.array([field_targets(g) for g in G])
    ntr = a.train
    nval = max(ntr // 8, 1)
    # margin: a high quantile of the validation residuals (under-prediction of a sum costs the bound)
    m0 = Surrogate(seed=0)
    m0.fit(X[:ntr - nval], Y[:ntr - nval], steps=a.fit_steps)
    resid = Y[ntr - nval:ntr] - m0.predict(X[ntr - nval:ntr])
    margin = np.quantile(resid, 0.9, axis=0)
    margin = np.maximum(margin, 0.0)
    model = Surrogate(seed=0)
    t1 = time.perf_counter()
    hist = model.fit(X[:ntr], Y[:ntr], steps=a.fit_steps)
    torch.cuda.synchronize()
    t_fit = time.perf_counter() - t1
    pred = model.predict(X[ntr:])
    mse = float(np.mean((pred - Y[ntr:]) ** 2))
    base = float(np.mean((Y[:ntr].mean(0) - Y[ntr:]) ** 2))

    # held-out: allocation from the prediction (+ margin) against the adjoint's sums
    met = met_margin = 0
    for k in range(a.test):
        g = G[ntr + k]
        ranges = R[ntr + k]
        Gt = g * ranges ** 2
        ref = 1.0
        bound = (ref * a.tol_rel) ** 2
        b0, _ = LagrangianEB(component_sums(pred[k]) * ranges ** 2, ref=ref, tol_rel=a.tol_rel)
        b1, _ = LagrangianEB(component_sums(pred[k] + margin) * ranges ** 2, ref=ref, tol_rel=a.tol_rel)
        met += modelled_error(Gt, b0) <= bound
        met_margin += modelled_error(Gt, b1) <= bound
    # timing of both paths on a few held-out scenes
    rng = np.random.default_rng(2024)
    for _ in range(ntr):
        random_scene(rng, a.particles, a.grid)
    t_adj, t_sur = [], []
    model.margin = margin
    for k in range(min(4, a.test)):
        st, E, dt = random_scene(rng, a.particles, a.grid)
        torch.cuda.synchronize()
        s0 = time.perf_counter()
        search_scheme(lambda: make_solver(a.particles, a.grid, E, dt), st, a.steps, tol_rel=a.tol_rel, ref=1.0,
                      checkpoint_every=a.checkpoint_every)
        t_adj.append(time.perf_counter() - s0)
        r = search_scheme_surrogate(model, st, E, dt, a.grid, a.steps, ref=1.0, tol_rel=a.tol_rel)
        t_sur.append(r["seconds"])
    os.makedirs(a.out, exist_ok=True)
    model.save(os.path.join(a.out, "surrogate_elastic2d.npz"))
    ev = {"scenes": {"train": ntr, "test": a.test, "particles": a.particles, "grid": a.grid, "substeps": a.steps,
                     "checkpoint_every": a.checkpoint_every},
          "dataset_seconds": t_data, "fit_seconds": t_fit, "fit_steps": a.fit_steps,
          "loss_first_last": [float(hist[0]), float(hist[-1])],
          "heldout_mse_log10": mse, "heldout_mse_log10_training_mean": base,
          "bound_met_fraction": met / a.test, "bound_met_fraction_with_margin": met_margin / a.test,
          "margin_log10": margin.tolist(), "tol_rel": a.tol_rel,
          "search_seconds_adjoint": float(np.mean(t_adj)), "search_seconds_surrogate": float(np.mean(t_sur)),
          "search_speedup": float(np.mean(t_adj) / np.mean(t_sur))}
    json.dump(ev, open(os.path.join(a.out, "surrogate_elastic2d_eval.json"), "w"), indent=1)
    print(json.dumps(ev), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
