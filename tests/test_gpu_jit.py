"""The fused substep kernel specialised to the record layout at run time
(MPMSolver.specialize -> aq_sim_specialize, hipRTC) against the generic
descriptor-driven kernel: from the same words and the same work items, the
records after two substeps are bit-identical, for the benchmark record (3D
fluid, dithered), the 2D elastic record (shared exponents for C and F) and a
record whose fields are out of component order; unspecialize returns to the
generic kernel.  The sparse scenes spread few particles over the whole
domain, so that each workgroup walks many work items (the specialised
kernel loads the next item during the current item's scatter)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def solver(kind, n):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import autoquantizer_amd as aq
    from autoquantizer_amd import layouts
    cfg = aq.MPMConfig()
    rng = np.random.default_rng(5)
    if kind.startswith("elastic_2d"):
        cfg.dim, cfg.use_fluid, cfg.n_grid, cfg.dt, cfg.E, cfg.nu, cfg.sticky = 2, False, 128, 1e-4, 1e4, 0.2, True
        _, t, g, sh = layouts.elastic_2d()
    else:
        cfg.dim, cfg.use_fluid, cfg.n_grid, cfg.dt, cfg.la = 3, True, 64, 2e-4, 400.0
        _, t, g, sh = layouts.fluid_3d()
        if kind == "fluid_3d_reordered":
            g = [("pack", 0, ["J"] + [f"C{a}{b}" for a in range(3) for b in range(3)]
                  + [f"v{a}" for a in range(3)] + [f"x{a}" for a in range(3)])]
    d = cfg.dim
    cfg.p_vol = (1 / cfg.n_grid * 0.5) ** d
    cfg.n_particles = n
    cfg.seed = 3
    cfg.compute_fp = "f32"
    sim = aq.MPMSolver(cfg, types=t, groups=g, shared=sh)
    vals = np.zeros((sim.ncomp, n))
    lo, hi = (0.08, 0.92) if kind.endswith("_sparse") else (0.2, 0.6)
    vals[:d] = rng.uniform(lo, hi, (d, n))
    vals[d:2 * d] = rng.uniform(-1, 1, (d, n))
    vals[2 * d:2 * d + d * d] = rng.uniform(-0.5, 0.5, (d * d, n))
    if kind.startswith("elastic_2d"):
        vals[-4:] = np.eye(2).reshape(-1, 1) + rng.uniform(-0.01, 0.01, (4, n))
    else:
        vals[-1] = rng.uniform(0.99, 1.01, n)
    sim.set_state(vals)
    return sim


@pytest.mark.parametrize("kind,n", [("fluid_3d", 40000), ("elastic_2d", 40000), ("fluid_3d_reordered", 40000),
                                    ("fluid_3d_sparse", 20000), ("elastic_2d_sparse", 4000)])
def test_specialised_kernel_is_bit_identical(kind, n, tmp_path):
    import torch
    print(f"[jit] {kind} n={n}", flush=True)
    sim = solver(kind, n)
    sim.advance(3, sort_every=1)  # sorted: the next substeps run the tiled kernels on these items
    w0, ids0, sc = sim.state.clone(), sim.ids.clone(), sim.step_count
    sim.advance(2, sort_every=0)
    generic = sim.state.clone()
    sim.state.copy_(w0)
    sim.ids.copy_(ids0)
    sim.step_count = sc
    sim.specialize(cache_dir=str(tmp_path))
    sim.advance(2, sort_every=0)
    assert torch.equal(sim.state, generic)
    sim.unspecialize()
    sim.state.copy_(w0)
    sim.ids.copy_(ids0)
    sim.step_count = sc
    sim.advance(2, sort_every=0)
    assert torch.equal(sim.state, generic)


def test_specialised_long_run_matches_generic_statistically(tmp_path):
    """100 substeps with re-sorting: specialised and generic runs agree to
    the statistical level (sort order is not deterministic, so the tiles
    differ in summation order)."""
    a = solver("fluid_3d", 30000)
    b = solver("fluid_3d", 30000)
    b.specialize(cache_dir=str(tmp_path))
    a.advance(100, sort_every=8)
    b.advance(100, sort_every=8)
    sa, sb = a.get_state(), b.get_state()
    for sl in (slice(0, 3), slice(3, 6), slice(15, 16)):
        ra = np.sqrt(np.mean(sa[sl] ** 2))
        assert np.sqrt(np.mean(sb[sl] ** 2)) == pytest.approx(ra, rel=1e-2)
