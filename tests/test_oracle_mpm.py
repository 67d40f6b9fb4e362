"""Pins the oracle MPM substep.

- the reference's own plain-MPM regression (taichi_submission/tests/python/test_mpm88.py:97-104);
- closed forms of the quadratic B-spline transfer used by
  mpm_exps/engine/mpm_solver.py:590-758: partition of unity, P2G mass and
  momentum conservation, G2P reproduction of an affine velocity field;
- grid boundary conditions (mpm_solver.py:659-664) and the sticky variant,
  per-axis walls (MPMConfig.lb / ub) and the bound clip (:696-698);
- rest state and free fall closed forms;
- the SDF collider (sdf_collision, rot_scene, :1444-1545): a static
  half-space removes only the velocity into it, a quarter turn maps nodes
  to voxels by the rotated offsets, and a node inside a spinning collider
  with no velocity of its own is carried at omega x r.
"""
import numpy as np
import pytest
from pytest import approx


def mpm88_state(orc):
    """Initial state of test_mpm88.py:85-88: 64x64 lattice, v=(0,-3), J=1."""
    N = 64
    n = N * N
    i = np.arange(n)
    st = np.zeros((orc.mpm_params(2, orc.FLUID, 128, 2e-4).dim * 2 + 4 + 1, n), dtype=np.float32)
    st[0] = (i % N) / N * 0.4 + 0.2
    st[1] = i / N / N * 0.4 + 0.05
    st[3] = -3.0
    st[8] = 1.0
    return st


def test_mpm88_regression(orc):
    """test_mpm88.py:9-104: mean(pos^k) after 500 substeps, rel 1e-2."""
    s = orc.mpm_params(2, orc.FLUID, 128, 2e-4, la=400.0, gravity=9.8, bound=3)
    st = mpm88_state(orc)
    orc.run_plain(s, st, 500)
    pos = st[:2].T.astype(np.float64).copy()
    pos[:, 1] *= 2
    regression = [0.31722742, 0.15826741, 0.10224003, 0.07810827]
    for k in range(4):
        assert (pos ** (k + 1)).mean() == approx(regression[k], rel=1e-2)


def random_state(orc, s, n, seed=0, f64=False):
    rng = np.random.default_rng(seed)
    d = s.dim
    nc = orc.num_comps(s)
    st = np.zeros((nc, n), dtype=np.float64 if f64 else np.float32)
    st[:d] = rng.uniform(0.3, 0.6, size=(d, n))
    st[d:2 * d] = rng.uniform(-1, 1, size=(d, n))
    st[2 * d:2 * d + d * d] = rng.uniform(-5, 5, size=(d * d, n))
    if s.material == orc.FLUID:
        st[-1] = rng.uniform(0.9, 1.1, size=n)
    else:
        F = np.eye(d).reshape(-1, 1) + rng.uniform(-0.05, 0.05, size=(d * d, n))
        st[-d * d:] = F
    return st


@pytest.mark.parametrize("dim,material", [(2, 0), (3, 0), (2, 1)])
def test_p2g_conservation(orc, dim, material):
    """Sum of node masses = n * p_mass; sum of node momenta = sum m v
    (sum_i w_ip (x_i - x_p) = 0 for quadratic B-splines, so the affine term cancels)."""
    s = orc.mpm_params(dim, material, 32, 1e-4)
    n = 500
    st = random_state(orc, s, n, seed=dim, f64=True)
    gm, gv = orc.grids(s, f64=True)
    orc.p2g(s, st, gm, gv)
    assert gm.sum() == approx(n * s.p_mass, rel=1e-12)
    mv = s.p_mass * st[dim:2 * dim].sum(axis=1)
    assert gv.sum(axis=1) == approx(mv, rel=1e-9, abs=1e-12 * s.p_mass)
    # partition of unity per particle: one particle, mass sums to p_mass
    gm1, gv1 = orc.grids(s, f64=True)
    orc.p2g(s, np.ascontiguousarray(st[:, :1]), gm1, gv1)
    assert gm1.sum() == approx(s.p_mass, rel=1e-14)
    assert np.count_nonzero(gm1) == 3 ** dim


@pytest.mark.parametrize("dim", [2, 3])
def test_g2p_affine_reproduction(orc, dim):
    """A linear grid velocity v(x) = a + B x is gathered exactly: v_p = a + B x_p,
    C_p = B (APIC with quadratic B-splines, mpm_solver.py:682-694)."""
    s = orc.mpm_params(dim, orc.FLUID, 32, 1e-4)
    n = 200
    st = random_state(orc, s, n, seed=3, f64=True)
    rng = np.random.default_rng(7)
    a = rng.uniform(-1, 1, size=dim)
    B = rng.uniform(-1, 1, size=(dim, dim))
    ng = s.n_grid
    idx = np.indices((ng,) * dim).reshape(dim, -1).astype(np.float64) * s.dx
    gv = (a[:, None] + B @ idx).copy()
    out = orc.g2p(s, st, gv)
    xp = st[:dim]
    assert out[dim:2 * dim] == approx(a[:, None] + B @ xp, rel=1e-10, abs=1e-12)
    C = out[2 * dim:2 * dim + dim * dim].reshape(dim, dim, n)
    assert C == approx(np.broadcast_to(B[:, :, None], C.shape), rel=1e-9, abs=1e-10)
    assert out[:dim] == approx(xp + s.dt * out[dim:2 * dim], rel=1e-15)


def test_grid_op_boundaries(orc):
    """Separating walls zero only the inward normal component (mpm_solver.py:659-664);
    sticky walls zero all components; gravity is subtracted from v_y first."""
    for boundary in (orc.SEPARATING, orc.STICKY):
        s = orc.mpm_params(2, orc.FLUID, 16, 1e-3, boundary=boundary, bound=3, gravity=9.8)
        gm, gv = orc.grids(s, f64=True)
        gm[:] = 2.0
        vel = np.array([[-1.0, 1.0], [1.0, -1.0], [0.5, 0.5]])
        nodes = [(1, 8), (14, 8), (8, 1), (8, 14), (8, 8)]
        for (i, j) in nodes:
            for v in vel:
                gv[:, i * 16 + j] = 2.0 * v
                orc.grid_op(s, gm.copy(), gv_ := gv.copy())
                got = gv_[:, i * 16 + j]
                exp = v.copy()
                exp[1] -= 1e-3 * 9.8
                scale = 2.0 / (2.0 + 1e-10)
                exp = np.array([v[0] * scale, v[1] * scale - 1e-3 * 9.8])
                I = (i, j)
                for k in range(2):
                    wall = I[k] < 3 or I[k] > 13
                    if boundary == orc.STICKY:
                        if I[0] < 3 or I[0] > 13 or I[1] < 3 or I[1] > 13:
                            exp[k] = 0.0
                    elif wall:
                        if (I[k] < 3 and exp[k] < 0) or (I[k] > 13 and exp[k] > 0):
                            exp[k] = 0.0
                assert got == approx(exp, rel=1e-14, abs=1e-300)


def test_rest_state(orc):
    """No gravity, v = 0, C = 0, F = I: the elastic body stays exactly at rest."""
    s = orc.mpm_params(2, orc.ELASTIC, 64, 1e-4, gravity=0.0, E=1e4, nu=0.2)
    n = 300
    st = random_state(orc, s, n, seed=5)
    st[2:8] = 0.0
    st[8:] = np.eye(2).reshape(-1, 1)
    x0 = st[:2].copy()
    orc.run_plain(s, st, 20)
    assert np.array_equal(st[:2], x0)
    assert np.all(st[2:8] == 0)


def test_free_fall(orc):
    """Far from walls and with J = 1 a fluid blob falls freely: after k substeps
    v_y = -k dt g (grid velocities are exact to rounding), C = 0.  Unit particle
    mass keeps the 1e-10 guard of inv_m = 1/(m + 1e-10) below rounding."""
    s = orc.mpm_params(3, orc.FLUID, 32, 1e-4, gravity=9.8, p_vol=1.0)
    n = 200
    st = random_state(orc, s, n, seed=9, f64=True)
    st[3:] = 0.0
    st[-1] = 1.0
    orc.run_plain(s, st, 10)
    assert st[4] == approx(np.full(n, -10 * 1e-4 * 9.8), rel=1e-7)
    # C = 4 inv_dx sum w v (x_i - x_p): rounding of the nearly uniform v only
    assert np.abs(st[6:15]).max() < 1e-6


def test_energy(orc):
    """compute_ek (mpm_solver.py:1043-1052): 3D 1/2 m |v|^2 + g m y; 2D sum |v|^2."""
    s3 = orc.mpm_params(3, orc.FLUID, 32, 1e-4)
    st = random_state(orc, s3, 50, seed=1, f64=True)
    e = 0.5 * s3.p_mass * (st[3:6] ** 2).sum() + 9.8 * s3.p_mass * st[1].sum()
    assert orc.energy(s3, st) == approx(e, rel=1e-12)
    s2 = orc.mpm_params(2, orc.ELASTIC, 32, 1e-4)
    st2 = random_state(orc, s2, 50, seed=2, f64=True)
    assert orc.energy(s2, st2) == approx((st2[2:4] ** 2).sum(), rel=1e-12)


def test_quant_run_matches_plain_then_quantize(orc):
    """orc_run_quant is decode -> substep -> encode: one quantized substep equals a
    plain substep on the decoded state followed by pack."""
    s = orc.mpm_params(2, orc.ELASTIC, 64, 1e-4, E=1e4, boundary=orc.STICKY)
    n = 256
    st = random_state(orc, s, n, seed=11)
    comps = []
    off = 0
    for c in range(12):
        comps.append(dict(kind=orc.FIXED, nbits=21, is_signed=1, bit_offset=off,
                          scale=orc.fixed_scale(21, 8.0, True)))
        off += 21
    nw = (off + 63) // 64
    w = orc.pack(comps, nw, st)
    dec = orc.unpack(comps, nw, w)
    w2 = w.copy()
    orc.run_quant(s, comps, nw, w2, 1, seed=0)
    ref = dec.copy()
    orc.run_plain(s, ref, 1)
    assert np.array_equal(w2, orc.pack(comps, nw, ref))


@pytest.mark.parametrize("case", ["rotation", "stretch", "rotated_stretch", "compression"])
def test_fixed_corotated_stress_closed_form(orc, case):
    """The oracle's 2D fixed-corotated stress (mpm_solver.py:1012-1031 with
    polar_decompose2d, TS/python/taichi/lang/linalg_impl.py:9-26), read back
    through its P2G: one particle at rest (v = 0, C = 0) scatters momentum
    w_i A (x_i - x_p) with A = -dt p_vol 4 inv_dx^2 sigma, and for quadratic
    B-splines sum_i w_i (x_i - x_p)(x_i - x_p)^T = dx^2/4 I, so
    sigma = (4 / dx^2) sum_i mom_i (x_i - x_p)^T / (-dt p_vol 4 inv_dx^2).
    Closed forms for F = R(theta) diag(s, t): R(theta) is F's rotation, so
    sigma = 2 mu R diag((s-1) s, (t-1) t) R^T + la J (J - 1) I, J = s t;
    a pure rotation is stress free."""
    E, nu, dt, ng = 1e4, 0.2, 1e-4, 64
    mu, la = E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))
    s = orc.mpm_params(2, orc.ELASTIC, ng, dt, gravity=0.0, E=E, nu=nu)
    theta, sx, ty = {"rotation": (0.7, 1.0, 1.0), "stretch": (0.0, 1.2, 1.0),
                     "rotated_stretch": (-0.4, 1.15, 0.9), "compression": (1.1, 0.8, 0.85)}[case]
    c, sn = np.cos(theta), np.sin(theta)
    R = np.array([[c, -sn], [sn, c]])
    F = R @ np.diag([sx, ty])
    st = np.zeros((12, 1), dtype=np.float64)
    xp = np.array([0.5 + 0.37 / ng, 0.5 + 0.61 / ng])
    st[:2, 0] = xp
    st[8:, 0] = F.reshape(-1)
    gm, gv = orc.grids(s, f64=True)
    orc.p2g(s, st, gm, gv)
    dx = 1.0 / ng
    nodes = np.nonzero(gm)[0]
    assert len(nodes) == 9
    A = np.zeros((2, 2))
    for nd in nodes:
        d = np.array([nd // ng, nd % ng]) * dx - xp
        A += np.outer(gv[:, nd], d)
    A *= 4.0 / dx ** 2
    sigma = A / (-dt * s.p_vol * 4 / dx ** 2)
    J = sx * ty
    want = 2 * mu * R @ np.diag([(sx - 1) * sx, (ty - 1) * ty]) @ R.T + la * J * (J - 1) * np.eye(2)
    scale = max(np.abs(want).max(), mu)
    assert np.abs(sigma - want).max() <= 1e-9 * scale, (sigma, want)


def _grid3(orc, n=16, bound=3):
    s = orc.mpm_params(3, orc.FLUID, n, 1e-4, gravity=0.0, bound=bound)
    gm, gv = orc.grids(s)
    gm[:] = 1.0
    return s, gm, gv


def test_per_axis_walls(orc):
    """Separating walls at lb / ub per axis (mpm_solver.py:97-105, 659-664):
    velocities into a wall vanish beyond it, everything else is kept."""
    s, gm, gv = _grid3(orc)
    orc.set_walls(s, [2, 3, 4], [13, 6, 11])
    rng = np.random.default_rng(0)
    gv[:] = rng.uniform(-1, 1, gv.shape).astype(np.float32)
    v0 = gv.copy()
    orc.grid_op(s, gm, gv)
    I = np.indices((16, 16, 16)).reshape(3, -1)
    lb, ub = np.array([2, 3, 4])[:, None], np.array([13, 6, 11])[:, None]
    want = np.where(((I < lb) & (v0 < 0)) | ((I > ub) & (v0 > 0)), 0.0, v0 / np.float32(1 + 1e-10))
    np.testing.assert_allclose(gv, want, rtol=1e-6, atol=0)


def test_bound_clip(orc):
    """use_bound_clip: G2P positions are clamped to [lb dx, ub dx] per axis;
    velocities are not."""
    s = orc.mpm_params(2, orc.FLUID, 32, 1e-2, gravity=0.0)
    orc.set_walls(s, [3, 5], [29, 20], clip=True)
    n = 200
    rng = np.random.default_rng(1)
    st = np.zeros((9, n), dtype=np.float32)
    st[:2] = rng.uniform(4 / 32, 19 / 32, (2, n))
    st[8] = 1.0
    gm, gv = orc.grids(s)
    gv[0] = -100.0  # a uniform grid velocity that carries every particle through the walls
    gv[1] = 100.0
    out = orc.g2p(s, st, gv)
    assert np.all(out[0] == np.float32(3 / 32)) and np.all(out[1] == np.float32(20 / 32))
    assert np.allclose(out[2], -100.0) and np.allclose(out[3], 100.0)
    s.bound_clip = 0
    free = orc.g2p(s, st, gv)
    assert np.all(free[0] < 3 / 32) and np.all(free[1] > 20 / 32)


def test_collider_static_half_space(orc):
    """Every voxel inside, normal +y, no rotation or spin: v_y < 0 becomes 0,
    every other velocity is kept (v_rel - min(v_rel . n, 0) n)."""
    s, gm, gv = _grid3(orc)
    rng = np.random.default_rng(2)
    gv[:] = rng.uniform(-1, 1, gv.shape).astype(np.float32)
    orc.grid_op(s, gm, gv)
    v0 = gv.copy()
    sdf = np.full((16, 16, 16), -1.0, np.float32)
    nrm = np.zeros((16, 16, 16, 3), np.float32)
    nrm[..., 1] = 1.0
    orc.collide(s, gv, sdf, nrm, 0.0, 0, 0.0, (8, 8, 8))
    np.testing.assert_array_equal(gv[0], v0[0])
    np.testing.assert_array_equal(gv[2], v0[2])
    np.testing.assert_array_equal(gv[1], np.maximum(v0[1], 0))


def test_collider_quarter_turn_maps_nodes(orc):
    """theta = pi/2: node (C0 + a, y, C2 + b) reads voxel (C0 - b, y - shift,
    C2 + a) (the offset rotated by -theta, 0.5 added and truncated), and the
    stored normal (1, 0, 0) acts as (cos theta, 0, -sin theta) = (0, 0, -1)."""
    s, gm, gv = _grid3(orc)
    C = (8, 8, 8)
    gv[2] = 1.0  # moving along +z, into a collider whose normal is -z
    orc.grid_op(s, gm, gv)
    v0 = gv.copy()
    sdf = np.zeros((16, 16, 16), np.float32)
    nrm = np.zeros((16, 16, 16, 3), np.float32)
    vox = (6, 7, 10)
    sdf[vox] = -0.5
    nrm[vox] = (1.0, 0.0, 0.0)
    shift = 2
    orc.collide(s, gv, sdf, nrm, np.pi / 2, shift, 0.0, C)
    hit = np.argwhere((gv != v0).any(axis=0).reshape(16, 16, 16))
    # vox = (C0 - b, y - shift, C2 + a) -> a = 2, b = 2, y = 9
    assert hit.tolist() == [[C[0] + 2, 9, C[2] + 2]]
    assert abs(gv[2].reshape(16, 16, 16)[C[0] + 2, 9, C[2] + 2]) < 1e-6


def test_collider_carries_resting_nodes(orc):
    """A node at rest inside a collider spinning at omega about the vertical
    line through C, with the normal along x: the collider's velocity
    omega e_y x r = (omega r_z, 0, -omega r_x) is imposed only along the
    normal, and only where the collider moves into the node (no friction):
    v = (max(omega r_z, 0), 0, 0)."""
    s, gm, gv = _grid3(orc)
    sdf = np.full((16, 16, 16), -1.0, np.float32)
    nrm = np.zeros((16, 16, 16, 3), np.float32)
    nrm[..., 0] = 1.0
    om = 5.0
    orc.grid_op(s, gm, gv)  # all at rest
    orc.collide(s, gv, sdf, nrm, 0.0, 0, om, (8, 8, 8))
    I = np.indices((16, 16, 16)).reshape(3, -1)
    rx, rz = (I[0] - 8) / 16.0, (I[2] - 8) / 16.0
    cvx = om * rz
    np.testing.assert_allclose(gv[0], np.maximum(cvx, 0.0), atol=1e-6)
    np.testing.assert_allclose(gv[2], 0.0, atol=1e-6)
    assert np.all(gv[1] == 0)


def test_collider_schedule(orc):
    """step_collider_difftaichi (lag 1): substep s runs with theta =
    omega dt (s - 1); collider_omega is 0 during substep 0."""
    th, sh, om = orc.collider_state(4e-4, 1 / 64, 0)
    assert th == 0 and sh == 0 and om == 0
    th, sh, om = orc.collider_state(4e-4, 1 / 64, 1)
    assert th == 0 and om == 5.0
    th, sh, om = orc.collider_state(4e-4, 1 / 64, 1001)
    assert th == approx(5.0 * 4e-4 * 1000) and sh == int(0.15 * 4e-4 * 1000 * 64 + 0.5)


@pytest.mark.parametrize("case", ["rotation", "uniform_stretch", "shear"])
def test_elastic_3d_stress_closed_form(orc, case):
    """The oracle's 3D elastic stress (mpm_solver.py:1027-1029:
    mu (F F^T - I) + la log(J), the scalar broadcast to every entry), read
    back through its P2G as in the 2D test (sum_i w_i (x_i - x_p)(x_i - x_p)^T
    = dx^2/4 I for the tensor-product quadratic B-spline).  Closed forms: a
    rotation gives F F^T = I and J = 1, no stress; F = s I gives
    mu (s^2 - 1) I + 3 la log(s) everywhere; a simple shear F = I + g e_x e_y^T
    keeps J = 1, so sigma = mu (F F^T - I)."""
    E, nu, dt, ng = 5e4, 0.3, 1e-4, 32
    mu, la = E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))
    s = orc.mpm_params(3, orc.ELASTIC, ng, dt, gravity=0.0, E=E, nu=nu)
    if case == "rotation":
        a, b = 0.7, -0.4
        Rz = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])
        Rx = np.array([[1, 0, 0], [0, np.cos(b), -np.sin(b)], [0, np.sin(b), np.cos(b)]])
        F = Rz @ Rx
        want = np.zeros((3, 3))
    elif case == "uniform_stretch":
        sc = 1.1
        F = sc * np.eye(3)
        want = mu * (sc * sc - 1) * np.eye(3) + 3 * la * np.log(sc)
    else:
        F = np.eye(3)
        F[0, 1] = 0.3
        want = mu * (F @ F.T - np.eye(3))
    st = np.zeros((24, 1), dtype=np.float64)
    xp = np.array([0.5 + 0.37 / ng, 0.5 + 0.61 / ng, 0.5 + 0.13 / ng])
    st[:3, 0] = xp
    st[15:, 0] = F.reshape(-1)
    gm, gv = orc.grids(s, f64=True)
    orc.p2g(s, st, gm, gv)
    dx = 1.0 / ng
    nodes = np.nonzero(gm)[0]
    assert len(nodes) == 27
    A = np.zeros((3, 3))
    for nd in nodes:
        I = np.array([nd // (ng * ng), (nd // ng) % ng, nd % ng])
        A += np.outer(gv[:, nd], I * dx - xp)
    A *= 4.0 / dx ** 2
    sigma = A / (-dt * s.p_vol * 4 / dx ** 2)
    scale = max(np.abs(want).max(), mu)
    assert np.abs(sigma - want).max() <= 1e-9 * scale, (sigma, want)


def test_quant_run_f64_matches_plain_then_quantize(orc):
    """orc_run_quant_f64 (the reference's 2D quantized runs compute in f64)
    is decode -> substep -> encode in double, dithered with the substep's
    noise.  Its records also differ from an f32 encode of the same values
    (28-bit codes of a range that is not a power of two resolve below f32's
    rounding), so the check would fail if any stage fell back to f32."""
    s = orc.mpm_params(2, orc.ELASTIC, 64, 1e-4, E=1e4, boundary=orc.STICKY)
    n = 256
    st = random_state(orc, s, n, seed=12, f64=True)
    comps = []
    off = 0
    for c in range(12):
        comps.append(dict(kind=orc.FIXED, nbits=28, is_signed=1, bit_offset=off, dither=1,
                          scale=orc.fixed_scale(28, 7.26239, True)))
        off += 28
    nw = (off + 63) // 64
    w = orc.pack(comps, nw, st, seed=3, step=4, f64=True)
    dec = orc.unpack(comps, nw, w, f64=True)
    assert np.abs(dec - st).max() <= comps[0]["scale"] * (1 + 1e-9)  # dithered: within one step
    w2 = w.copy()
    orc.run_quant(s, comps, nw, w2, 1, seed=3, step0=5, f64=True)
    ref = dec.copy()
    orc.run_plain(s, ref, 1)
    assert np.array_equal(w2, orc.pack(comps, nw, ref, seed=3, step=5, f64=True))
    w32 = orc.pack(comps, nw, ref.astype(np.float32), seed=3, step=5)
    assert not np.array_equal(w2, w32)
