"""The f64 oracle adjoint of the plain substep (oracle/adj_ref.c), pinned by
central finite differences of its own forward (the reference's Taichi
autodiff cannot run here) and by the accumulation range of the
error-contribution sums (mpm_solver.py:1167-1188: states 1..T-1)."""
import numpy as np
import pytest


def small_scene(orc, kind, n, seed=0):
    rng = np.random.default_rng(seed)
    if kind == "fluid_2d":
        s = orc.mpm_params(2, orc.FLUID, 32, 2e-4, la=400.0)
        d = 2
    elif kind == "elastic_2d":
        s = orc.mpm_params(2, orc.ELASTIC, 32, 1e-4, E=1e4, nu=0.2, boundary=orc.STICKY)
        d = 2
    elif kind == "elastic_3d":
        s = orc.mpm_params(3, orc.ELASTIC, 16, 1e-4, E=5e3, nu=0.3)
        d = 3
    else:
        s = orc.mpm_params(3, orc.FLUID, 16, 2e-4, la=400.0)
        d = 3
    nc = orc.num_comps(s)
    st = np.zeros((nc, n))
    st[:d] = rng.uniform(0.3, 0.6, (d, n))
    st[d:2 * d] = rng.uniform(-1, 1, (d, n))
    st[2 * d:2 * d + d * d] = rng.uniform(-2, 2, (d * d, n))
    if kind == "elastic_2d":
        st[-4:] = np.eye(2).reshape(-1, 1) + rng.uniform(-0.05, 0.05, (4, n))
    elif kind == "elastic_3d":
        st[-9:] = np.eye(3).reshape(-1, 1) + rng.uniform(-0.05, 0.05, (9, n))
    else:
        st[-1] = rng.uniform(0.95, 1.05, n)
    return s, st


def loss_of(orc, s, st, nsteps, loss, direction=None):
    x = st.copy()
    orc.run_plain(s, x, nsteps)
    if loss == "energy":
        return orc.energy(s, x)
    return float((direction * x).sum())


@pytest.mark.parametrize("kind", ["fluid_2d", "elastic_2d", "fluid_3d", "elastic_3d"])
@pytest.mark.parametrize("loss", ["energy", "linear"])
def test_adjoint_matches_central_differences(orc, kind, loss):
    s, st = small_scene(orc, kind, 150, seed=1)
    rng = np.random.default_rng(2)
    nsteps = 6
    direction = rng.standard_normal(st.shape) if loss == "linear" else None
    L, adj0, _ = orc.adjoint_run(s, st, nsteps, loss=loss, direction=direction)
    assert L == pytest.approx(loss_of(orc, s, st, nsteps, loss, direction), rel=1e-12)
    # perturbation of every component, scaled to the component's spread
    delta = rng.standard_normal(st.shape) * (np.abs(st).max(axis=1, keepdims=True) + 1e-3)
    h = 1e-6
    fd = (loss_of(orc, s, st + h * delta, nsteps, loss, direction)
          - loss_of(orc, s, st - h * delta, nsteps, loss, direction)) / (2 * h)
    ad = float((adj0 * delta).sum())
    assert ad == pytest.approx(fd, rel=1e-6, abs=1e-9 * abs(L) + 1e-12)


@pytest.mark.parametrize("kind", ["fluid_2d", "elastic_2d"])
def test_adjoint_per_component(orc, kind):
    """Each component's adjoint alone against finite differences."""
    s, st = small_scene(orc, kind, 80, seed=3)
    nsteps = 4
    _, adj0, _ = orc.adjoint_run(s, st, nsteps)
    rng = np.random.default_rng(4)
    for c in range(st.shape[0]):
        delta = np.zeros_like(st)
        delta[c] = rng.standard_normal(st.shape[1])
        h = 1e-5 * (np.abs(st[c]).max() + 1e-2)
        fd = (loss_of(orc, s, st + h * delta, nsteps, "energy")
              - loss_of(orc, s, st - h * delta, nsteps, "energy")) / (2 * h)
        ad = float((adj0 * delta).sum())
        assert ad == pytest.approx(fd, rel=1e-5, abs=1e-10), c


def test_error_sums_cover_intermediate_states(orc):
    """g_c sums (dL/ds_t)^2 over t = 1..T-1: the adjoint at state t is the
    adjoint of the run started from s_t."""
    s, st = small_scene(orc, "fluid_2d", 60, seed=5)
    T = 5
    _, _, gsum = orc.adjoint_run(s, st, T)
    expect = np.zeros(st.shape[0])
    x = st.copy()
    for t in range(1, T):
        orc.run_plain(s, x, 1)
        _, adj_t, _ = orc.adjoint_run(s, x, T - t)
        expect += (adj_t ** 2).sum(axis=1)
    np.testing.assert_allclose(gsum, expect, rtol=1e-10)


# ---- the adjoint error-gradient run: d |x_T - x_ref|^2 / d sigma -----------

def test_noise_xi_is_centred_uniform(orc):
    xi = orc.noise_xi(7, 3, 4, 20000)
    assert xi.min() >= -0.5 and xi.max() < 0.5
    assert abs(xi.mean()) < 0.01 and abs(xi.var() - 1 / 12) < 0.002
    assert xi[2, 11] == orc.noise_u(7, 3, 11, 2) - 0.5
    # components and steps draw different noise
    assert abs(np.corrcoef(xi[0], xi[1])[0, 1]) < 0.03
    assert abs(np.corrcoef(xi[0], orc.noise_xi(7, 4, 1, 20000)[0])[0, 1]) < 0.03


def _noise_scene(orc, kind, n, T, seed):
    s, st = small_scene(orc, kind, n, seed=seed)
    ref = st.copy()
    orc.run_plain(s, ref, T)
    d = 3 if kind == "fluid_3d" else 2
    sigma = np.abs(st).max(axis=1) * 2.0 ** -10 + 1e-6
    return s, st, ref[:d], sigma


@pytest.mark.parametrize("kind", ["elastic_2d", "fluid_2d", "fluid_3d"])
def test_noise_scale_gradient_matches_central_differences(orc, kind):
    """Every component's dL/dsigma against central differences with
    h = 1e-3 of sigma."""
    T = 6
    s, st, xref, sigma = _noise_scene(orc, kind, 120, T, seed=11)
    L, dsig = orc.noise_scale_run(s, st, T, sigma, 5, xref)
    assert L == pytest.approx(orc.noise_scale_loss(s, st, T, sigma, 5, xref), rel=1e-12)
    for c in range(len(sigma)):
        h = 1e-3 * sigma[c]
        e = np.zeros_like(sigma)
        e[c] = h
        fd = (orc.noise_scale_loss(s, st, T, sigma + e, 5, xref)
              - orc.noise_scale_loss(s, st, T, sigma - e, 5, xref)) / (2 * h)
        assert dsig[c] == pytest.approx(fd, rel=1e-6, abs=1e-9 * L / sigma[c]), c


def test_noise_scale_gradient_is_homogeneous(orc):
    """Against the noise-free run, L is quadratic in sigma to leading order,
    so sum_c sigma_c dL/dsigma_c = 2 L (Euler) up to the curvature of the
    dynamics; with sigma = 0, L and its gradient vanish."""
    T = 8
    s, st, xref, sigma = _noise_scene(orc, "elastic_2d", 150, T, seed=12)
    L, dsig = orc.noise_scale_run(s, st, T, sigma * 1e-3, 9, xref)
    assert (sigma * 1e-3) @ dsig == pytest.approx(2 * L, rel=1e-4)
    L0, d0 = orc.noise_scale_run(s, st, T, np.zeros_like(sigma), 9, xref)
    assert L0 == 0.0 and np.all(d0 == 0.0)


def _collider_substep(orc, s, col, st):
    """One plain f64 substep with the collider after the grid update
    (substep_difftaichi, mpm_solver.py:792-798)."""
    gm, gv = orc.grids(s, f64=True)
    orc.p2g(s, st, gm, gv)
    orc.grid_op(s, gm, gv)
    orc.lib().orc_collide_f64(orc.ctypes.byref(s), orc.ctypes.byref(col), orc._ptr(gv))
    return orc.g2p(s, st, gv)


def test_collider_adjoint_matches_central_differences(orc):
    """The adjoint of a substep with the SDF collider (I - n n^T on nodes
    whose velocity after the walls approaches the collider) and per-axis
    walls, against central differences of the oracle's forward, for a linear
    functional of s_{t+1}; the collider's term is not vacuous (the adjoint
    without it differs)."""
    rng = np.random.default_rng(7)
    n = 400
    s = orc.mpm_params(3, orc.FLUID, 16, 2e-4, la=400.0)
    orc.set_walls(s, [3, 3, 3], [13, 11, 13])
    st = np.zeros((16, n))
    st[:3] = rng.uniform(0.25, 0.7, (3, n))
    st[3:6] = rng.uniform(-3, 3, (3, n))
    st[6:15] = rng.uniform(-2, 2, (9, n))
    st[15] = rng.uniform(0.95, 1.05, n)
    sdf = np.zeros((16, 16, 16), np.float32)
    I = np.indices(sdf.shape).reshape(3, -1).T
    inside = np.sum(((I - np.array([8, 7, 8])) / np.array([5.0, 3.0, 4.0])) ** 2, axis=1) < 1.0
    sdf.reshape(-1)[inside] = rng.uniform(-0.05, -2e-5, inside.sum()).astype(np.float32)
    nrm = rng.normal(size=(16, 16, 16, 3)).astype(np.float32)
    nrm /= np.linalg.norm(nrm, axis=-1, keepdims=True)
    col = orc.collider(sdf, nrm, 0.3, 1, 20.0, (8, 8, 8))
    direction = rng.standard_normal(st.shape)
    ad = orc.adj_substep(s, st, direction, col)
    ad_plain = orc.adj_substep(s, st, direction)
    assert np.abs(ad - ad_plain).max() > 1e-2 * np.abs(ad).max()
    delta = rng.standard_normal(st.shape) * (np.abs(st).max(axis=1, keepdims=True) + 1e-3)
    h = 1e-7
    fd = (float((direction * _collider_substep(orc, s, col, st + h * delta)).sum())
          - float((direction * _collider_substep(orc, s, col, st - h * delta)).sum())) / (2 * h)
    assert float((ad * delta).sum()) == pytest.approx(fd, rel=1e-6)
    # per component group, so that no group's error hides behind another's
    for sl in (slice(0, 3), slice(3, 6), slice(6, 15), slice(15, 16)):
        dd = np.zeros_like(delta)
        dd[sl] = delta[sl]
        fd = (float((direction * _collider_substep(orc, s, col, st + h * dd)).sum())
              - float((direction * _collider_substep(orc, s, col, st - h * dd)).sum())) / (2 * h)
        assert float((ad * dd).sum()) == pytest.approx(fd, rel=1e-5, abs=1e-9), sl
