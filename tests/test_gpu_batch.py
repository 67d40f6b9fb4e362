"""Batched scheme search (autoquantizer_amd.batch): simulations of uneven
sizes, dimensions and materials, each with its own candidate scheme, run
from one captured HIP graph over several streams.  Every replay (with eager
work in between) reproduces the one-by-one run of the same simulations to
rounding level (sort order and P2G summation order are not deterministic),
and each quantized run's final positions stay close to its FP64 run."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_graph_replays_match_sequential_runs():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from autoquantizer_amd.batch import Batch, fp64_positions, position_errors, random_specs
    specs = random_specs(6, seed=3, n_min=4096, n_max=40000)
    assert {(s.dim, s.fluid) for s in specs} >= {(3, True)}
    print("[batch] specs", [(s.dim, s.fluid, s.n, s.n_grid) for s in specs], flush=True)
    nsteps = 40
    b = Batch(specs)
    b.autotune(candidates=(4, 8), steps=8)
    assert all(s.sort_every in (4, 8) for s in specs)
    b.reset()
    b.run_sequential(nsteps)
    seq = b.positions()
    b.capture(nsteps, nstreams=3)
    for replay in range(2):
        b.run_graph()
        got = b.positions()
        for sp, a, r in zip(specs, got, seq):
            assert np.all(np.isfinite(a))
            # rounding-level agreement: far below one cell after 40 substeps
            assert np.sqrt(np.mean((a - r) ** 2)) < 1e-3 / sp.n_grid, (replay, sp.n, sp.dim)
        torch.zeros(1000, device="cuda").add_(1.0)  # eager work between replays
        torch.cuda.synchronize()
    ref = fp64_positions(specs, nsteps)
    errs = position_errors(got, ref)
    for sp, e in zip(specs, errs):
        step = 2.0 ** (1 - sp.bits["x"])
        print(sp.dim, sp.fluid, sp.n, sp.bits, e, flush=True)
        assert e["rms"] < 0.05 / sp.n_grid + 4 * step


def test_group_matches_sequential_runs():
    """The grouped launches (aq_group: one fused launch per class of
    simulations, one gather per dimension) reproduce each simulation's own
    aq_advance with the same sort interval to rounding level, eagerly and
    from a captured graph replayed twice."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from autoquantizer_amd.batch import Batch, random_specs
    specs = random_specs(7, seed=5, n_min=4096, n_max=60000)
    kinds = {(s.dim, s.fluid) for s in specs}
    assert len(kinds) >= 2, kinds
    print("[group] specs", [(s.dim, s.fluid, s.n, s.n_grid) for s in specs], flush=True)
    nsteps, k = 37, 8
    b = Batch(specs)
    for sp in specs:
        sp.sort_every = k
    b.run_sequential(nsteps)
    seq = b.positions()
    for wg in (0, 64):
        b.run_group(nsteps, sort_every=k, workgroups=wg)
        got = b.positions()
        for sp, a, r in zip(specs, got, seq):
            assert np.all(np.isfinite(a))
            assert np.sqrt(np.mean((a - r) ** 2)) < 1e-3 / sp.n_grid, (wg, sp.n, sp.dim, sp.fluid)
    b.group_sort_every, b.group_workgroups = k, 0
    b.capture_group(nsteps)
    for replay in range(2):
        b.run_group_graph()
        got = b.positions()
        for sp, a, r in zip(specs, got, seq):
            assert np.sqrt(np.mean((a - r) ** 2)) < 1e-3 / sp.n_grid, (replay, sp.n, sp.dim, sp.fluid)
        torch.zeros(1000, device="cuda").add_(1.0)  # eager work between replays
        torch.cuda.synchronize()
    assert all(sim.step_count == nsteps for sim in b.sims)


def test_full_size_batch_properties():
    """The batched search at its full size: 32 simulations (2D fluid and
    elastic, 3D fluid, 16k-500k particles, grids 64^2-128^3, each its own
    scheme), 300 substeps, the grouped launches from a captured graph
    against the one-by-one runs.  Stated bounds: every run finite; group
    and one-by-one runs within 1/20 of a cell RMS of each other (sort and
    summation orders differ, and 300 substeps amplify rounding); each
    quantized run within 1 % of the domain RMS of its FP64 run (the batch
    holds candidates as coarse as 9 digits for F, whose elastic block
    deforms visibly differently: 0.55 % on a 256^2 grid)."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from autoquantizer_amd.batch import Batch, fp64_positions, position_errors, random_specs
    specs = random_specs(32, seed=0)
    ns = [s.n for s in specs]
    assert len(specs) == 32 and min(ns) >= 16000 and max(ns) <= 500000
    assert {(s.dim, s.fluid) for s in specs} == {(2, False), (2, True), (3, True)}
    print("[batch] full size:", sum(ns), "particles", flush=True)
    nsteps = 300
    b = Batch(specs)
    for sp in specs:
        sp.sort_every = 16
    b.run_sequential(nsteps)
    seq = b.positions()
    b.group_sort_every, b.group_workgroups = 16, 0
    b.capture_group(nsteps)
    b.run_group_graph()
    got = b.positions()
    del b
    torch.cuda.empty_cache()
    ref = fp64_positions(specs, nsteps)
    errs = position_errors(got, ref)
    worst_pair, worst_ref = 0.0, 0.0
    for sp, a, r, e in zip(specs, got, seq, errs):
        assert np.all(np.isfinite(a)), (sp.dim, sp.fluid, sp.n)
        d = float(np.sqrt(np.mean((a - r) ** 2)))
        worst_pair = max(worst_pair, d * sp.n_grid)
        worst_ref = max(worst_ref, e["rms"] * sp.n_grid)
        assert d < 0.05 / sp.n_grid, (sp.dim, sp.fluid, sp.n, d)
        assert e["rms"] < 0.01, (sp.dim, sp.fluid, sp.n, sp.bits, e)
    print(f"largest group vs one-by-one RMS {worst_pair:.3g} cells, vs FP64 {worst_ref:.3g} cells", flush=True)
