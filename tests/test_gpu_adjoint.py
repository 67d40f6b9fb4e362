"""GPU adjoint of the plain substep (aq_adj_substep, autoquantizer_amd.adjoint)
against the f64 oracle adjoint (itself pinned by finite differences) and
against central finite differences of the GPU forward."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def make(orc, kind, n, storage, seed=0):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import autoquantizer_amd as aq
    rng = np.random.default_rng(seed)
    cfg = aq.MPMConfig()
    if kind == "elastic_2d":
        cfg.dim, cfg.use_fluid, cfg.n_grid, cfg.dt, cfg.E, cfg.nu, cfg.sticky = 2, False, 32, 1e-4, 1e4, 0.2, True
        s = orc.mpm_params(2, orc.ELASTIC, 32, 1e-4, E=1e4, nu=0.2, boundary=orc.STICKY)
    elif kind == "fluid_2d":
        cfg.dim, cfg.use_fluid, cfg.n_grid, cfg.dt, cfg.la = 2, True, 32, 2e-4, 400.0
        s = orc.mpm_params(2, orc.FLUID, 32, 2e-4, la=400.0)
    else:
        cfg.dim, cfg.use_fluid, cfg.n_grid, cfg.dt, cfg.la = 3, True, 16, 2e-4, 400.0
        s = orc.mpm_params(3, orc.FLUID, 16, 2e-4, la=400.0)
    d = cfg.dim
    cfg.p_vol = (1 / cfg.n_grid * 0.5) ** d
    cfg.n_particles = n
    cfg.compute_fp = storage
    sim = aq.MPMSolver(cfg)
    nc = sim.ncomp
    st = np.zeros((nc, n))
    st[:d] = rng.uniform(0.3, 0.6, (d, n))
    st[d:2 * d] = rng.uniform(-1, 1, (d, n))
    st[2 * d:2 * d + d * d] = rng.uniform(-2, 2, (d * d, n))
    if kind == "elastic_2d":
        st[-4:] = np.eye(2).reshape(-1, 1) + rng.uniform(-0.05, 0.05, (4, n))
    else:
        st[-1] = rng.uniform(0.95, 1.05, n)
    sim.set_state(st)
    return sim, s, st


@pytest.mark.parametrize("kind", ["fluid_2d", "elastic_2d", "fluid_3d"])
@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_adjoint_substep_vs_oracle(orc, kind, storage):
    from autoquantizer_amd.adjoint import SubstepAdjoint
    print(f"[adj step] {kind} {storage}", flush=True)
    sim, s, st = make(orc, kind, 300, storage, seed=1)
    rng = np.random.default_rng(2)
    adj_next = rng.standard_normal(st.shape)
    ref = orc.adj_substep(s, sim.get_state(), adj_next)
    A = SubstepAdjoint(sim)
    an = torch.zeros_like(sim.state)
    an[:, :300] = torch.as_tensor(adj_next, dtype=sim.state.dtype, device="cuda")
    out = torch.zeros_like(sim.state)
    gsum = torch.zeros(sim.ncomp, dtype=torch.float64, device="cuda")
    A.step(sim.state, an, out, gsum)
    got = out[:, :300].double().cpu().numpy()
    scale = np.abs(ref).max(axis=1, keepdims=True) + 1e-30
    tol = 1e-9 if storage == "f64" else 2e-4
    assert (np.abs(got - ref) / scale).max() < tol
    np.testing.assert_allclose(gsum.cpu().numpy(), (got ** 2).sum(axis=1), rtol=1e-6 if storage == "f32" else 1e-12)


@pytest.mark.parametrize("kind", ["fluid_2d", "elastic_2d"])
def test_error_contributions_vs_oracle(orc, kind):
    """Checkpointed run (segments of 4) in f64 against the oracle's full run."""
    from autoquantizer_amd.adjoint import error_contributions
    sim, s, st = make(orc, kind, 200, "f64", seed=3)
    L, grad0, g = error_contributions(sim, 10, checkpoint_every=4)
    L_ref, adj_ref, g_ref = orc.adjoint_run(s, sim.get_state(), 10)
    assert L == pytest.approx(L_ref, rel=1e-10)
    np.testing.assert_allclose(g, g_ref, rtol=1e-7)
    scale = np.abs(adj_ref).max(axis=1, keepdims=True) + 1e-30
    assert (np.abs(grad0 - adj_ref) / scale).max() < 1e-8


@pytest.mark.parametrize("kind", ["fluid_2d", "elastic_2d", "fluid_3d"])
def test_adjoint_vs_finite_differences(orc, kind):
    """Directional derivative of the loss of 8 substeps: central differences
    of the GPU f64 forward (h = 1e-6 of each component's scale) against the
    f64 adjoint (1e-6 relative) and the fp32 adjoint (1e-3 relative)."""
    from autoquantizer_amd.adjoint import error_contributions, energy_loss, _advance_plain
    print(f"[adj fd] {kind}", flush=True)
    n, T = 400, 8
    sim64, s, st = make(orc, kind, n, "f64", seed=5)
    sim32, _, _ = make(orc, kind, n, "f32", seed=5)
    _, g64, _ = error_contributions(sim64, T, checkpoint_every=3)
    _, g32, _ = error_contributions(sim32, T, checkpoint_every=3)
    rng = np.random.default_rng(6)
    delta = rng.standard_normal(st.shape) * (np.abs(st).max(axis=1, keepdims=True) + 1e-2)
    base = sim64.state.clone()

    def loss_at(eps):
        x = base.clone()
        x[:, :n] += torch.as_tensor(eps * delta, dtype=x.dtype, device="cuda")
        _advance_plain(sim64, x, T)
        return float(energy_loss(sim64, x))

    h = 1e-6
    fd = (loss_at(h) - loss_at(-h)) / (2 * h)
    assert float((g64 * delta).sum()) == pytest.approx(fd, rel=1e-6)
    assert float((g32 * delta).sum()) == pytest.approx(fd, rel=1e-3)


def test_rollout_autograd_linear_loss(orc):
    """Rollout as a torch op: gradient of a linear functional of s_T against
    the oracle adjoint of the same functional (f64)."""
    from autoquantizer_amd.adjoint import Rollout
    sim, s, st = make(orc, "fluid_2d", 150, "f64", seed=7)
    rng = np.random.default_rng(8)
    w = rng.standard_normal(st.shape)
    W = torch.zeros_like(sim.state)
    W[:, :150] = torch.as_tensor(w, device="cuda")
    s0 = sim.state.clone().requires_grad_(True)
    sT = Rollout.apply(s0, sim, 5, 2)
    (W * sT).sum().backward()
    _, adj_ref, _ = orc.adjoint_run(s, sim.get_state(), 5, loss="linear", direction=w)
    got = s0.grad[:, :150].cpu().numpy()
    scale = np.abs(adj_ref).max(axis=1, keepdims=True) + 1e-30
    assert (np.abs(got - adj_ref) / scale).max() < 1e-8


# ---- the adjoint error-gradient run: d |x_T - x_ref|^2 / d sigma -----------

def _sigma(st):
    return np.abs(st).max(axis=1) * 2.0 ** -10 + 1e-6


@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_noise_add_and_dot_vs_oracle(orc, storage):
    from autoquantizer_amd.adjoint import Noise
    print(f"[noise add/dot] {storage}", flush=True)
    n = 1000
    sim, s, st = make(orc, "elastic_2d", n, storage, seed=21)
    sigma = _sigma(st)
    nz = Noise(sigma, seed=77)
    x = sim.state.clone()
    nz.add(sim, x, 5)
    xi = orc.noise_xi(77, 5, sim.ncomp, n)
    expect = sim.get_state() + sigma[:, None] * xi
    got = x[:, :n].double().cpu().numpy()
    if storage == "f64":
        assert np.array_equal(got, expect)
    else:
        np.testing.assert_allclose(got, expect, rtol=2e-7, atol=1e-7 * sigma.max())
    rng = np.random.default_rng(22)
    a = rng.standard_normal((sim.ncomp, n))
    A = torch.zeros_like(sim.state)
    A[:, :n] = torch.as_tensor(a, dtype=A.dtype, device="cuda")
    out = torch.zeros(sim.ncomp, dtype=torch.float64, device="cuda")
    nz.dot(sim, A, 5, out)
    a_used = A[:, :n].double().cpu().numpy()
    np.testing.assert_allclose(out.cpu().numpy(), (a_used * xi).sum(axis=1), rtol=1e-11, atol=1e-12)
    # keyed by particle id: a permuted state with its ids draws each particle's own noise
    perm = np.random.default_rng(24).permutation(n)
    ids = torch.arange(sim.capacity, dtype=torch.int32, device="cuda")
    ids[:n] = torch.as_tensor(perm, dtype=torch.int32, device="cuda")
    xp = torch.zeros_like(sim.state)
    xp[:, :n] = sim.state[:, :n][:, torch.as_tensor(perm, device="cuda")]
    nz.add(sim, xp, 5, ids)
    got_p = xp[:, :n].double().cpu().numpy()
    if storage == "f64":
        assert np.array_equal(got_p, expect[:, perm])
    else:
        np.testing.assert_allclose(got_p, expect[:, perm], rtol=2e-7, atol=1e-7 * sigma.max())
    out_p = torch.zeros(sim.ncomp, dtype=torch.float64, device="cuda")
    Ap = torch.zeros_like(sim.state)
    Ap[:, :n] = A[:, :n][:, torch.as_tensor(perm, device="cuda")]
    nz.dot(sim, Ap, 5, out_p, ids)
    np.testing.assert_allclose(out_p.cpu().numpy(), out.cpu().numpy(), rtol=1e-11, atol=1e-12)


@pytest.mark.parametrize("kind", ["elastic_2d", "fluid_2d", "fluid_3d"])
def test_noise_scale_gradients_vs_oracle(orc, kind):
    """Checkpointed f64 run (segments of 3) against the oracle's stored-state run."""
    from autoquantizer_amd.adjoint import noise_scale_gradients, _advance_plain
    print(f"[noise grad] {kind}", flush=True)
    n, T = 200, 10
    sim, s, st = make(orc, kind, n, "f64", seed=23)
    x = sim.state.clone()
    _advance_plain(sim, x, T)
    xref = x[:sim.dim, :n].cpu().numpy()
    sigma = _sigma(st)
    L, dsig = noise_scale_gradients(sim, T, sigma, xref, seed=3, checkpoint_every=3)
    L_ref, dsig_ref = orc.noise_scale_run(s, sim.get_state(), T, sigma, 3, xref)
    assert L == pytest.approx(L_ref, rel=1e-8)
    np.testing.assert_allclose(dsig, dsig_ref, rtol=1e-7, atol=1e-9 * np.abs(dsig_ref).max())


@pytest.mark.parametrize("kind", ["elastic_2d", "fluid_3d"])
def test_noise_scale_fp32_adjoint_vs_finite_differences(orc, kind):
    """Per-field derivative of |x_T - x_T^FP64|^2 w.r.t. the field's noise
    scale: the fp32 adjoint kernels (f64 reduction) around the f64 forward,
    against central differences (h = 1e-3) of that f64 forward, within
    1e-3 relative for every field."""
    from autoquantizer_amd.adjoint import (Noise, _Schedule, _advance_plain, field_gradients,
                                           final_position_error, noise_scale_gradients)
    print(f"[noise fd] {kind}", flush=True)
    n, T = 500, 16
    sim64, s, st = make(orc, kind, n, "f64", seed=25)
    sim32, _, _ = make(orc, kind, n, "f32", seed=25)
    d = sim64.dim
    x = sim64.state.clone()
    _advance_plain(sim64, x, T)
    xref = x[:d, :n].clone()
    sigma = _sigma(st)
    _, dsig = noise_scale_gradients(sim64, T, sigma, xref.cpu().numpy(), seed=4, checkpoint_every=4,
                                    adjoint_solver=sim32)
    fields = [("x", slice(0, d)), ("v", slice(d, 2 * d)), ("C", slice(2 * d, 2 * d + d * d)),
              ("J|F", slice(2 * d + d * d, sim64.ncomp))]
    ad = field_gradients(dsig, sigma, fields)
    base = sim64.state.clone()

    def loss(scale_f):
        sched = _Schedule(sim64, T, 4, Noise(sigma * scale_f, seed=4))
        return float(final_position_error(sim64, sched.forward(base), xref))

    h = 1e-3
    for name, sl in fields:
        up = np.ones(sim64.ncomp)
        dn = np.ones(sim64.ncomp)
        up[sl] += h
        dn[sl] -= h
        fd = (loss(up) - loss(dn)) / (2 * h)
        print(name, "adjoint", ad[name], "fd", fd, flush=True)
        assert ad[name] == pytest.approx(fd, rel=1e-3), name


def test_noise_scale_adjoint_full_size_scene():
    """The adjoint error-gradient run at its full size (2D elastic jelly,
    65,536 particles, 256^2, 256 substeps, checkpoints every 16; the
    benchmark's adjoint_check): per field, with that field's noise alone,
    the fp32 adjoint around the f64 forward agrees with central differences
    (h = 1e-3) of the f64 forward within 1e-3 relative, and so does the f64
    adjoint."""
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import bench
    print("[noise fd] full-size adjoint scene", flush=True)
    r = bench.adjoint_check()
    for name, v in r["per_field"].items():
        print(name, v["adjoint_fp32"], v["adjoint_f64"], v["finite_difference"], flush=True)
        assert np.isfinite(v["finite_difference"]) and v["finite_difference"] != 0.0, name
        assert v["relative_difference_fp32"] <= 1e-3, (name, v)
        assert v["relative_difference_f64"] <= 1e-3, (name, v)
