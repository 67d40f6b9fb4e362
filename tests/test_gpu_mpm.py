"""GPU parity of the MPM substep kernels with the C oracle.

- P2G grid (mass, momentum) within f32 accumulation-order tolerance, for the
  global-atomic and the LDS-tiled kernels;
- grid update exact on the GPU's own P2G grid;
- G2P from the GPU's own grid against the oracle's G2P + encode on that grid:
  within one quantization step, nearly all records identical (the kernels
  sum the stencil axis by axis); the codec fused into G2P bit-exact against
  the oracle's encode of the fp32-storage kernel's results on the same grid;
- whole substeps re-seeded from the GPU state each time: one GPU substep vs
  one oracle substep from identical words; fixed-point codes within one step,
  shared-exponent digits within two (only the P2G summation order differs);
- multi-substep runs (fused G2P->P2G, sorting) against separate substeps,
  statistically (chaotic amplification of rounding differences);
- the plain fp32 path against the reference regression test_mpm88.py:97-104;
- the plain fp64 path against the fp64 oracle;
- sort, limits and energy.
"""
import numpy as np
import pytest
from pytest import approx

from gpu_util import to_oracle

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def scene(orc, kind, n, seed=0, storage="quant"):
    """Random blob scenes for the two benchmark materials (moderate C, F, J so
    a few substeps stay smooth)."""
    _torch()
    import autoquantizer_amd as aq
    from autoquantizer_amd import layouts
    cfg = aq.MPMConfig()
    rng = np.random.default_rng(seed)
    if kind in ("elastic_2d", "hand_2d"):
        cfg.dim, cfg.use_fluid, cfg.n_grid, cfg.dt, cfg.E, cfg.nu, cfg.sticky = 2, False, 128, 1e-4, 1e4, 0.2, True
        names, t, g, sh = layouts.elastic_2d() if kind == "elastic_2d" else layouts.hand_scheme_2d()
        s = orc.mpm_params(2, orc.ELASTIC, 128, 1e-4, E=1e4, nu=0.2, boundary=orc.STICKY, gravity=9.8)
        lo, hi = 0.3, 0.5
    else:
        cfg.dim, cfg.use_fluid, cfg.n_grid, cfg.dt, cfg.la = 3, True, 64, 2e-4, 400.0
        names, t, g, sh = layouts.fluid_3d()
        s = orc.mpm_params(3, orc.FLUID, 64, 2e-4, la=400.0, gravity=9.8)
        lo, hi = 0.1, 0.5
    cfg.p_vol = (1 / cfg.n_grid * 0.5) ** cfg.dim
    cfg.n_particles = n
    cfg.compute_fp = "f32"
    cfg.seed = 11
    if storage == "quant":
        sim = aq.MPMSolver(cfg, types=t, groups=g, shared=sh)
    else:
        cfg.compute_fp = storage
        sim = aq.MPMSolver(cfg)
    d = cfg.dim
    vals = np.zeros((len(names), n))
    vals[:d] = rng.uniform(lo, hi, (d, n))
    vals[d:2 * d] = rng.uniform(-1, 1, (d, n))
    vals[2 * d:2 * d + d * d] = rng.uniform(-0.5, 0.5, (d * d, n))
    if kind in ("elastic_2d", "hand_2d"):
        vals[-4:] = np.eye(2).reshape(-1, 1) + rng.uniform(-0.01, 0.01, (4, n))
    else:
        vals[-1] = rng.uniform(0.99, 1.01, n)
    sim.set_state(vals)
    return sim, s


def oracle_comps(sim):
    return to_oracle(sim.comps)


def step_sizes(sim, dec):
    """Quantization step of every component at every particle: fixed-point
    scale, or one digit of the (shared- or own-) exponent group (the group's
    largest member's power of two over 2^(digit bits - 1)); for unquantized
    f32 members 2^-18 of the component's largest value.  A float member's
    step is at least 2^-20 of its group's largest value over all particles:
    near zero its digit step falls below the f32 rounding of the physics
    (v is a sum of grid velocities of the field's magnitude)."""
    steps = np.zeros_like(dec)
    for k, c in enumerate(sim.comps):
        if c["kind"] == 0:
            steps[k] = c["scale"]
        elif c["kind"] == 3:
            # unquantized f32: operand-order differences of a 9-term sum
            # whose terms exceed the result (C is a difference quotient of
            # grid velocities), relative to the component's largest value
            steps[k] = max(float(np.abs(dec[k]).max()), 1e-30) * 2.0 ** -18
        else:
            grp = [j for j, cc in enumerate(sim.comps) if cc["kind"] == 2 and cc["exp_offset"] == c["exp_offset"]]
            mx = np.abs(dec[grp]).max(axis=0)
            db = c["num_bits"] - c["is_signed"]
            steps[k] = np.exp2(np.floor(np.log2(np.maximum(mx, 1e-30)))) * 2.0 ** (1 - db)
            steps[k] = np.maximum(steps[k], max(float(mx.max()), 1e-30) * 2.0 ** -20)
    return steps


@pytest.mark.parametrize("kind", ["elastic_2d", "fluid_3d"])
@pytest.mark.parametrize("tiled", [False, True])
def test_p2g_grid(orc, kind, tiled):
    print(f"[p2g] {kind} tiled={tiled}", flush=True)
    sim, s = scene(orc, kind, 4000, seed=1)
    if tiled:
        sim.build_pid()
    st = sim.get_state(original_order=False).astype(np.float32)
    sim.p2g_bls()
    g = sim.grid_dense()
    gm, gv = orc.grids(s)
    orc.p2g(s, st, gm, gv)
    d = s.dim
    assert g[d] == approx(gm, rel=2e-5, abs=1e-6 * s.p_mass)
    mom_scale = np.abs(gv).max()
    assert np.abs(g[:d] - gv).max() <= 2e-5 * mom_scale


@pytest.mark.parametrize("kind", ["elastic_2d", "fluid_3d"])
def test_gather_reaches_particles_far_from_the_sort(orc, kind):
    """After a sort, a third of the particles are moved several grid blocks
    (in place, so the sort's work items stay current): their stencils leave
    their tiles, they scatter through the untiled path into blocks the
    sort's block list does not hold, and the gather must still add them.
    Twice in a row (the second P2G reaches the same far blocks again, now
    held by the list), against the oracle's P2G of the moved state."""
    torch = _torch()
    print(f"[gather list] {kind}", flush=True)
    sim, s = scene(orc, kind, 4000, seed=23, storage="f32")
    sim.build_pid()
    n = sim.current_idx
    idx = torch.arange(0, n, 3, device="cuda")
    shift = 0.25 if kind == "fluid_3d" else 0.2  # 16 cells: 4 blocks (3D), 3 blocks (2D)
    sim.state[0, idx] += shift
    sim.state[1, idx] -= shift / 2
    st = sim.get_state(original_order=False).astype(np.float32)
    gm, gv = orc.grids(s)
    orc.p2g(s, st, gm, gv)
    d = s.dim
    for rep in range(2):
        sim.p2g_bls()
        g = sim.grid_dense()
        assert g[d] == approx(gm, rel=2e-5, abs=1e-6 * s.p_mass), rep
        assert np.abs(g[:d] - gv).max() <= 2e-5 * np.abs(gv).max(), rep
    # and back: the far blocks, now empty, are cleared by the next gather
    sim.state[0, idx] -= shift
    sim.state[1, idx] += shift / 2
    st = sim.get_state(original_order=False).astype(np.float32)
    gm, gv = orc.grids(s)
    orc.p2g(s, st, gm, gv)
    sim.p2g_bls()
    g = sim.grid_dense()
    assert g[d] == approx(gm, rel=2e-5, abs=1e-6 * s.p_mass)


@pytest.mark.parametrize("kind", ["elastic_2d", "fluid_3d"])
def test_fused_advance_with_far_particles(orc, kind):
    """advance() on a sorted state that was translated by two grid blocks
    after the sort (in place, so the sort's work items stay current): every
    particle's stencil leaves its tile, so every fused kernel scatters
    through the untiled path into blocks the sort's list does not hold; the
    first gather appends them to the list, and each later tiled kernel
    commits the last gather's appends while the flagged-block counters
    alternate between gathers.  Against the same translated state advanced
    untiled (no sort: every gather walks the whole grid): the runs agree
    statistically."""
    torch = _torch()
    print(f"[gather list] advance {kind}", flush=True)
    sim_a, s = scene(orc, kind, 4000, seed=29, storage="f32")
    sim_a.build_pid()
    shift = 0.125  # 8 cells (3D, 64^3) = 2 blocks; 16 cells (2D, 128^2) = 2 blocks
    sim_a.state[0, :sim_a.current_idx] += shift
    sim_a.state[1, :sim_a.current_idx] += shift / 2
    sim_b, _ = scene(orc, kind, 4000, seed=29, storage="f32")
    sim_b.set_state(sim_a.get_state(original_order=True))
    # the random elastic state (E = 1e4, dt 1e-4 on 128^2: CFL number 1.3)
    # diverges after 7 substeps, in the oracle too; the fluid stays calm
    steps = 9 if kind == "fluid_3d" else 5
    sim_a.advance(steps)  # tiled: the sort's items stay valid (state moved in place)
    sim_b.advance(steps)  # untiled
    a, b = sim_a.get_state(), sim_b.get_state()
    assert np.all(np.isfinite(a)) and np.all(np.isfinite(b))
    _close_runs(a, b, sim_a)


@pytest.mark.parametrize("kind", ["elastic_2d", "fluid_3d"])
def test_p2g_grid_edge_particles(orc, kind):
    """Particles at and beyond the domain edge: the stencil base is clamped
    to [0, n_grid - 3], fx leaves [0.5, 1.5] and a weight turns negative;
    the sorted (tiled) P2G scatters them through the float-atomic path and
    agrees with the oracle's P2G."""
    print(f"[p2g edge] {kind}", flush=True)
    sim, s = scene(orc, kind, 4000, seed=2, storage="f32")
    st = sim.get_state(original_order=False)
    d = s.dim
    dx = 1.0 / s.n_grid
    edge = [0.0, 0.05 * dx, 0.3 * dx, 1.0 - 2.6 * dx, 1.0 - 0.5 * dx, 1.0]
    for j, xe in enumerate(edge * 3):
        st[j % d, j] = xe
        st[(j + 1) % d, j] = 0.4
    sim.set_state(st)
    sim.build_pid()
    st = sim.get_state(original_order=False).astype(np.float32)
    sim.p2g_bls()
    g = sim.grid_dense()
    gm, gv = orc.grids(s)
    orc.p2g(s, st, gm, gv)
    assert g[d] == approx(gm, rel=2e-5, abs=1e-6 * s.p_mass)
    mom_scale = np.abs(gv).max()
    assert np.abs(g[:d] - gv).max() <= 2e-5 * mom_scale


def assert_within_steps(orc, sim, wg, wo, frac_equal=0.95):
    """Decoded records of two word arrays within one quantization step
    (fixed point) / two digits (shared exponent) of each other, and nearly
    all records identical."""
    oc = oracle_comps(sim)
    a = orc.unpack(oc, sim.nwords, wg).astype(np.float64)
    b = orc.unpack(oc, sim.nwords, wo).astype(np.float64)
    steps = step_sizes(sim, b)
    rel = np.abs(a - b) / steps
    for k, c in enumerate(sim.comps):
        lim = 1.0 if c["kind"] == 0 else 2.0
        assert rel[k].max() <= lim + 1e-9, (k, rel[k].max())
    assert np.mean((wg == wo).all(axis=0)) >= frac_equal
    return float(rel.max())


@pytest.mark.parametrize("kind", ["elastic_2d", "fluid_3d"])
@pytest.mark.parametrize("tiled", [False, True])
def test_g2p_on_gpu_grid(orc, kind, tiled):
    """G2P from the GPU's own grid against the oracle's G2P (reference
    operand order) + encode on that grid: the kernels sum the stencil axis by
    axis, so records agree within one quantization step and nearly all are
    identical (noise keyed by the particle id: compared in original order)."""
    print(f"[g2p] {kind} tiled={tiled}", flush=True)
    sim, s = scene(orc, kind, 4000, seed=2)
    if tiled:
        sim.build_pid()
    st = sim.get_state(original_order=True).astype(np.float32)
    sim.p2g_bls()
    sim.grid_op()
    g = sim.grid_dense()
    sim.g2p_bls()
    words_gpu = sim.get_words(original_order=True)
    gv = np.ascontiguousarray(g[:s.dim]).astype(np.float32)
    out = orc.g2p(s, st, gv)
    words_orc = orc.pack(oracle_comps(sim), sim.nwords, out, seed=11, step=0)
    assert_within_steps(orc, sim, words_gpu, words_orc, frac_equal=0.97)


@pytest.mark.parametrize("kind", ["elastic_2d", "fluid_3d", "hand_2d"])
@pytest.mark.parametrize("tiled", [False, True])
def test_g2p_codec_bit_exact(orc, kind, tiled):
    """The codec fused into G2P is bit-exact: the quantized kernel's words
    equal the oracle's encode of what the fp32-storage kernel computes from
    the same decoded state on the same grid."""
    print(f"[g2p codec] {kind} tiled={tiled}", flush=True)
    sim_q, s = scene(orc, kind, 4000, seed=9)
    sim_f, _ = scene(orc, kind, 4000, seed=9, storage="f32")
    dec = sim_q.get_state(original_order=True)
    sim_f.set_state(dec)
    if tiled:
        sim_q.build_pid()
        sim_f.build_pid()
    sim_q.p2g_bls()
    sim_q.grid_op()
    sim_f.load_grid(sim_q.grid_velocity())
    sim_q.g2p_bls()
    sim_f.g2p_bls()
    out = sim_f.get_state(original_order=True).astype(np.float32)
    words_orc = orc.pack(oracle_comps(sim_q), sim_q.nwords, out, seed=11, step=0)
    assert np.array_equal(sim_q.get_words(original_order=True), words_orc)


def test_g2p_codec_bit_exact_any_field_order(orc):
    """A record whose fields are not in component order (J first, then C's
    shared group, v, x) goes through the general encoder: still bit-exact."""
    _torch()
    import autoquantizer_amd as aq
    from autoquantizer_amd import layouts
    names, t, g, sh = layouts.fluid_3d()
    order = ["J"] + [f"C{a}{b}" for a in range(3) for b in range(3)] + [f"v{a}" for a in range(3)] + \
            [f"x{a}" for a in range(3)]
    sims = []
    for storage in ("quant", "f32"):
        cfg = aq.MPMConfig()
        cfg.dim, cfg.use_fluid, cfg.n_grid, cfg.dt, cfg.la = 3, True, 64, 2e-4, 400.0
        cfg.p_vol = (1 / 64 * 0.5) ** 3
        cfg.n_particles = 3000
        cfg.seed = 11
        cfg.compute_fp = "f32"
        if storage == "quant":
            sims.append(aq.MPMSolver(cfg, types=t, groups=[("pack", 0, order)], shared=sh))
        else:
            cfg.compute_fp = "f32"
            sims.append(aq.MPMSolver(cfg))
    sim_q, sim_f = sims
    assert sim_q.comps[15]["bit_offset"] < sim_q.comps[0]["bit_offset"]
    rng = np.random.default_rng(12)
    vals = np.zeros((16, 3000))
    vals[:3] = rng.uniform(0.1, 0.5, (3, 3000))
    vals[3:6] = rng.uniform(-1, 1, (3, 3000))
    vals[6:15] = rng.uniform(-0.5, 0.5, (9, 3000))
    vals[15] = rng.uniform(0.99, 1.01, 3000)
    sim_q.set_state(vals)
    sim_f.set_state(sim_q.get_state(original_order=True))
    sim_q.build_pid()
    sim_f.build_pid()
    sim_q.p2g_bls()
    sim_q.grid_op()
    sim_f.load_grid(sim_q.grid_velocity())
    sim_q.g2p_bls()
    sim_f.g2p_bls()
    out = sim_f.get_state(original_order=True).astype(np.float32)
    words_orc = orc.pack(oracle_comps(sim_q), sim_q.nwords, out, seed=11, step=0)
    assert np.array_equal(sim_q.get_words(original_order=True), words_orc)


@pytest.mark.parametrize("kind", ["elastic_2d", "fluid_3d"])
def test_grid_op_exact(orc, kind):
    """Grid update on the GPU's P2G grid equals the oracle's grid_op exactly on
    every node that received mass."""
    sim, s = scene(orc, kind, 3000, seed=3)
    sim.p2g_bls()
    g0 = sim.grid_dense()
    sim.grid_op()
    g1 = sim.grid_dense()
    d = s.dim
    gm = np.ascontiguousarray(g0[d]).astype(np.float32)
    gv = np.ascontiguousarray(g0[:d]).astype(np.float32)
    orc.grid_op(s, gm, gv)
    touched = g0[d] != 0
    assert np.array_equal(g1[:d][:, touched].astype(np.float32), gv[:, touched])


@pytest.mark.parametrize("kind", ["elastic_2d", "fluid_3d", "hand_2d"])
def test_substeps_reseeded(orc, kind):
    """Each GPU substep (sort + tiled P2G + grid + tiled G2P) from state S_t
    against one oracle substep from the same words S_t."""
    print(f"[substeps] {kind}", flush=True)
    sim, s = scene(orc, kind, 3000, seed=4)
    oc = oracle_comps(sim)
    worst = 0.0
    for t in range(5):
        w0 = sim.get_words(original_order=True)
        sim.substep()
        wg = sim.get_words(original_order=True)
        wo = w0.copy()
        orc.run_quant(s, oc, sim.nwords, wo, 1, seed=11, step0=t)
        a = orc.unpack(oc, sim.nwords, wg).astype(np.float64)
        b = orc.unpack(oc, sim.nwords, wo).astype(np.float64)
        steps = step_sizes(sim, b)
        rel = np.abs(a - b) / steps
        for k, c in enumerate(sim.comps):
            lim = 1.0 if c["kind"] == 0 else 2.0
            assert rel[k].max() <= lim + 1e-9, (t, k, rel[k].max())
        # nearly all records are identical (unquantized f32 members keep the operand-order differences)
        if all(c["kind"] != 3 for c in sim.comps):
            assert np.mean((wg != wo).any(axis=0)) < 0.05
        worst = max(worst, float(rel.max()))
    print("max difference in quantization steps:", worst)


def _close_runs(a, b, sim):
    d = sim.dim
    assert np.abs(a[:d] - b[:d]).max() < 1e-4
    for sl in (slice(d, 2 * d), slice(2 * d, 2 * d + d * d), slice(2 * d + d * d, None)):
        ra = np.sqrt(np.mean((a[sl] - b[sl]) ** 2))
        assert ra <= 1e-2 * np.sqrt(np.mean(b[sl] ** 2)) + 1e-6


def test_fused_advance_matches_substeps(orc):
    """advance(n) (G2P fused with the next P2G, tiled) vs n separate substeps."""
    sim_a, s = scene(orc, "fluid_3d", 5000, seed=5)
    sim_b, _ = scene(orc, "fluid_3d", 5000, seed=5)
    for _ in range(8):
        sim_a.substep()
    sim_b.advance(8, sort_every=4)
    _close_runs(sim_b.get_state(), sim_a.get_state(), sim_a)


def test_fused_untiled_matches_tiled(orc):
    sim_a, s = scene(orc, "fluid_3d", 5000, seed=6)
    sim_b, _ = scene(orc, "fluid_3d", 5000, seed=6)
    sim_a.advance(6)
    sim_b.advance(6, sort_every=2)
    _close_runs(sim_b.get_state(), sim_a.get_state(), sim_a)


def test_fused_elastic(orc):
    sim_a, s = scene(orc, "elastic_2d", 5000, seed=7)
    sim_b, _ = scene(orc, "elastic_2d", 5000, seed=7)
    for _ in range(6):
        sim_a.substep()
    sim_b.advance(6, sort_every=3)
    _close_runs(sim_b.get_state(), sim_a.get_state(), sim_a)


def test_sort_is_permutation(orc):
    sim, s = scene(orc, "fluid_3d", 7000, seed=7)
    w0 = sim.get_words(original_order=True)
    sim.build_pid()
    ids = sim.ids[:7000].cpu().numpy()
    assert np.array_equal(np.sort(ids), np.arange(7000))
    assert np.array_equal(sim.get_words(original_order=True), w0)
    st = sim.get_state(original_order=False)
    cb = np.clip(np.floor(st[:3] * 64 - 0.5), 0, 61).astype(int)
    base = cb // 4
    key = (base[0] * 16 + base[1]) * 16 + base[2]
    assert np.all(np.diff(key) >= 0)
    # inside a block, layer by layer: the r-th particle of every cell holding
    # more than r, in cell order, then the (r+1)-th ...
    cell = ((cb[0] & 3) << 4) | ((cb[1] & 3) << 2) | (cb[2] & 3)
    for k in np.unique(key):
        seq = cell[key == k]
        cnt = np.bincount(seq, minlength=64)
        expect = [c for r in range(cnt.max()) for c in range(64) if cnt[c] > r]
        assert seq.tolist() == expect


def test_mpm88_regression_gpu():
    """test_mpm88.py:9-104 on the plain fp32 GPU path (fused, sorted, tiled)."""
    _torch()
    import autoquantizer_amd as aq
    cfg = aq.MPMConfig()
    cfg.dim, cfg.use_fluid, cfg.n_grid, cfg.dt, cfg.la = 2, True, 128, 2e-4, 400.0
    cfg.p_vol = (1 / 128 * 0.5) ** 2
    N = 64
    n = N * N
    cfg.n_particles = n
    cfg.compute_fp = "f32"
    sim = aq.MPMSolver(cfg)
    i = np.arange(n)
    vals = np.zeros((9, n))
    vals[0] = (i % N) / N * 0.4 + 0.2
    vals[1] = i / N / N * 0.4 + 0.05
    vals[3] = -3.0
    vals[8] = 1.0
    sim.set_state(vals)
    sim.advance(500, sort_every=8)
    pos = sim.get_state()[:2].T.copy()
    pos[:, 1] *= 2
    regression = [0.31722742, 0.15826741, 0.10224003, 0.07810827]
    for k in range(4):
        assert (pos ** (k + 1)).mean() == approx(regression[k], rel=1e-2)


def test_plain_f64_matches_oracle(orc):
    """FP64 plain storage: 20 substeps agree with the oracle to 1e-9 relative."""
    sim, s = scene(orc, "fluid_3d", 3000, seed=8, storage="f64")
    st = sim.get_state()
    for _ in range(20):
        sim.substep()
    ref = st.copy()
    orc.run_plain(s, ref, 20)
    got = sim.get_state()
    scale = np.abs(ref).max(axis=1, keepdims=True) + 1e-30
    assert (np.abs(got - ref) / scale).max() < 1e-9


def test_limits_and_energy(orc):
    sim, s = scene(orc, "fluid_3d", 9000, seed=9)
    st = sim.get_state()
    sim.detect_bound_kernel()
    lim = sim.detect_bound()
    assert lim == approx(np.maximum(np.abs(st.max(axis=1)), np.abs(st.min(axis=1))), rel=1e-6)
    e = sim.compute_ek()
    assert e == approx(orc.energy(s, st.astype(np.float32)), rel=1e-6)


@pytest.mark.parametrize("storage", ["f32", "quant"])
def test_range_run_in_one_call(orc, storage):
    """advance_detect (aq_advance_limits: the limits after every substep
    inside one advance) against one substep per call with detect_bound_kernel
    after each: the same per-component ranges (to the rounding level at which
    the two schedules' tiled sums differ)."""
    a, _ = scene(orc, "fluid_3d", 6000, seed=13, storage=storage)
    b, _ = scene(orc, "fluid_3d", 6000, seed=13, storage=storage)
    a.init_limits()
    a.advance_detect(40, sort_every=8)
    b.init_limits()
    for t in range(40):
        b.advance(1, sort_every=8 if t % 8 == 0 else 0)
        b.detect_bound_kernel()
    la, lb = a.get_limits(), b.get_limits()
    assert np.all(np.isfinite(la))
    np.testing.assert_allclose(la, lb, rtol=1e-3)


def test_wrap_counter(orc):
    """Range-violation counter (aq_sim_count_wraps): a fluid record whose v
    range (0.25) is below the particles' speeds.  After one G2P the count per
    component equals the number of stores whose rounded value left the
    member's range, recomputed from the fp32-storage kernel's G2P output on
    the same grid with the same dithering noise; the records are still the
    oracle's encode of that output (counting changes nothing stored)."""
    _torch()
    import autoquantizer_amd as aq
    from autoquantizer_amd import layouts
    n = 3000
    names, t, g, sh = layouts.fluid_3d(v_range=0.25)
    sims = []
    for storage in ("quant", "f32"):
        cfg = aq.MPMConfig()
        cfg.dim, cfg.use_fluid, cfg.n_grid, cfg.dt, cfg.la = 3, True, 64, 2e-4, 400.0
        cfg.p_vol = (1 / 64 * 0.5) ** 3
        cfg.n_particles = n
        cfg.seed = 21
        cfg.compute_fp = "f32"
        sims.append(aq.MPMSolver(cfg, types=t, groups=g, shared=sh) if storage == "quant" else aq.MPMSolver(cfg))
    sim_q, sim_f = sims
    rng = np.random.default_rng(31)
    vals = np.zeros((16, n))
    vals[:3] = rng.uniform(0.2, 0.6, (3, n))
    vals[3:6] = rng.uniform(-0.24, 0.24, (3, n))
    vals[6:15] = rng.uniform(-30, 30, (9, n))  # a strong affine field: grid speeds well above 0.25
    vals[15] = rng.uniform(0.99, 1.01, n)
    sim_q.set_state(vals)
    sim_f.set_state(sim_q.get_state(original_order=True))
    sim_q.build_pid()
    sim_f.build_pid()
    sim_q.p2g_bls()
    sim_q.grid_op()
    sim_f.load_grid(sim_q.grid_velocity())
    step = sim_q.step_count
    sim_q.count_wraps()
    sim_q.g2p_bls()
    sim_f.g2p_bls()
    got = sim_q.wrap_counts()
    out = sim_f.get_state(original_order=True).astype(np.float32)
    words_orc = orc.pack(oracle_comps(sim_q), sim_q.nwords, out, seed=21, step=step)
    assert np.array_equal(sim_q.get_words(original_order=True), words_orc)
    want = np.zeros(16, dtype=np.int64)
    f32 = np.float32
    for c, cc in enumerate(sim_q.comps):
        if cc["kind"] != 0:
            continue
        u = np.array([orc.dither_u(21, step, p, c) for p in range(n)], dtype=f32) if cc["use_dithering"] else 0
        y = (out[c] - f32(cc["offset"])) * f32(1.0 / cc["scale"])
        if cc["use_dithering"]:
            y = (y + u) - f32(0.5)
        y = y + np.copysign(f32(0.5), y)
        q = np.trunc(y.astype(np.float64))
        nb = cc["num_bits"]
        lo, hi = (-2.0 ** (nb - 1), 2.0 ** (nb - 1) - 1) if cc["is_signed"] else (0.0, 2.0 ** nb - 1)
        want[c] = int(np.sum(np.isnan(y) | (q < lo) | (q > hi)))
    print("wrapped stores per component:", got.tolist(), flush=True)
    assert want[3:6].sum() > 0  # the scene does overflow v
    np.testing.assert_array_equal(got, want)
    sim_q.count_wraps(False)


def test_nan_inf_injection_flagged():
    """Failure detection from the host: NaN and +Inf injected into two grid
    nodes that hold mass reach the particles of their stencils in the next
    G2P.  The range-violation counter of the quantized simulation counts, per
    fixed-point member, exactly the stores whose value was not finite (the
    same G2P on fp32 storage with the same grid shows which), and the limits
    detector of the fp32 simulation reports an infinite v range."""
    _torch()
    import autoquantizer_amd as aq
    from autoquantizer_amd import layouts
    n = 3000
    names, t, g, sh = layouts.fluid_3d(v_range=8.0)
    sims = []
    for storage in ("quant", "f32"):
        cfg = aq.MPMConfig()
        cfg.dim, cfg.use_fluid, cfg.n_grid, cfg.dt, cfg.la = 3, True, 64, 2e-4, 400.0
        cfg.p_vol = (1 / 64 * 0.5) ** 3
        cfg.n_particles = n
        cfg.seed = 23
        cfg.compute_fp = "f32"
        sims.append(aq.MPMSolver(cfg, types=t, groups=g, shared=sh) if storage == "quant" else aq.MPMSolver(cfg))
    sim_q, sim_f = sims
    rng = np.random.default_rng(41)
    vals = np.zeros((16, n))
    vals[:3] = rng.uniform(0.3, 0.5, (3, n))
    vals[3:6] = rng.uniform(-0.5, 0.5, (3, n))
    vals[15] = 1.0
    sim_q.set_state(vals)
    sim_f.set_state(sim_q.get_state(original_order=True))
    sim_q.build_pid()
    sim_f.build_pid()
    sim_q.p2g_bls()
    sim_q.grid_op()
    grid = sim_q.grid_velocity()
    heavy = np.flatnonzero(grid[:, 3] > 0)
    assert heavy.size > 100
    grid[heavy[heavy.size // 3], 0] = np.nan
    grid[heavy[2 * heavy.size // 3], 1] = np.inf
    sim_q.load_grid(grid)
    sim_f.load_grid(grid)
    print(f"[inject] NaN at node {heavy[heavy.size // 3]}, +Inf at node {heavy[2 * heavy.size // 3]}", flush=True)
    sim_q.count_wraps()
    sim_q.g2p_bls()
    sim_f.g2p_bls()
    got = sim_q.wrap_counts()
    sim_q.count_wraps(False)
    out = sim_f.get_state(original_order=True)
    want = np.zeros(16, dtype=np.int64)
    for c, cc in enumerate(sim_q.comps):
        if cc["kind"] == 0:  # fixed point: x, v, J (the scene keeps every finite value in range)
            want[c] = int(np.sum(~np.isfinite(out[c])))
    print("non-finite stores per component:", got.tolist(), flush=True)
    assert want[3:6].sum() > 0 and want[:3].sum() > 0
    np.testing.assert_array_equal(got, want)
    sim_f.init_limits()
    sim_f.detect_bound_kernel()
    lim = sim_f.detect_bound()
    assert np.isinf(lim[4]), lim  # +Inf reached v_y
