"""3D elastic MPM (mpm_solver.py:1027-1029: sigma = mu (F F^T - I) +
la log(J) on every entry) and the letters emission of the 3D elastic
scenes (run_mpm_elastic_3d.py:47-63; add_mesh, mpm_solver.py:1356-1441)
on the GPU, against the oracle (pinned by closed forms and finite
differences on the CPU):

- plain f64 substeps vs the oracle to 1e-9, f32 statistically;
- the f64/f32 adjoint substep vs the oracle adjoint;
- the codec fused into G2P bit-exact on the reference's performance record
  (24 fixed-point components of run_mpm_elastic_3d.py:371-374 in one
  bit_pack, dithered);
- add_mesh seeds count_emit(...) particles, each inside a voxel of the
  mesh, and the run stays finite.
"""
import os

import numpy as np
import pytest

from gpu_util import to_oracle

pytestmark = pytest.mark.gpu

BENCH_BITS = [24, 23, 23, 17, 17, 15, 11, 14, 11, 13, 11, 10, 11, 12, 11, 19, 18, 18, 18, 18, 17, 17, 17, 18]


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def scene(orc, n, storage, seed=0, quant_bits=None):
    _torch()
    import autoquantizer_amd as aq
    cfg = aq.MPMConfig()
    cfg.dim, cfg.use_fluid, cfg.n_grid, cfg.dt, cfg.E, cfg.nu = 3, False, 32, 1e-4, 5e3, 0.3
    cfg.p_vol = (1 / 32 * 0.5) ** 3
    cfg.n_particles = n
    cfg.compute_fp = storage
    cfg.seed = 11
    if quant_bits is not None:
        r = np.load(os.path.join(os.path.dirname(aq.__file__), "data", "ranges_elastic_3d.npy")) * 2.0
        r[:3] = 4.0
        cfg.use_quantaichi = cfg.use_bitpack = cfg.use_dithering = True
        cfg.quant_bits = np.asarray(quant_bits)
        cfg.quant_ranges = r
        cfg.compute_fp = "f32"
    sim = aq.MPMSolver(cfg)
    rng = np.random.default_rng(seed)
    st = np.zeros((24, n))
    st[:3] = rng.uniform(0.3, 0.6, (3, n))
    st[3:6] = rng.uniform(-1, 1, (3, n))
    st[6:15] = rng.uniform(-2, 2, (9, n))
    st[15:] = np.eye(3).reshape(-1, 1) + rng.uniform(-0.05, 0.05, (9, n))
    sim.set_state(st)
    s = orc.mpm_params(3, orc.ELASTIC, 32, 1e-4, E=5e3, nu=0.3)
    return sim, s, st


def test_plain_f64_matches_oracle(orc):
    print("[elastic3d] f64 substeps", flush=True)
    sim, s, _ = scene(orc, 3000, "f64", seed=1)
    st = sim.get_state()
    for _ in range(10):
        sim.substep()
    ref = st.copy()
    orc.run_plain(s, ref, 10)
    got = sim.get_state()
    scale = np.abs(ref).max(axis=1, keepdims=True) + 1e-30
    assert (np.abs(got - ref) / scale).max() < 1e-9


def test_plain_f32_matches_oracle(orc):
    print("[elastic3d] f32 substeps", flush=True)
    sim, s, _ = scene(orc, 3000, "f32", seed=2)
    st = sim.get_state().astype(np.float32)
    for _ in range(5):
        sim.substep()
    ref = st.copy()
    orc.run_plain(s, ref, 5)
    got = sim.get_state()
    scale = np.abs(ref).max(axis=1, keepdims=True) + 1e-30
    err = (np.abs(got - ref) / scale).max(axis=1)
    print("max relative difference per component:", np.round(err, 8).tolist(), flush=True)
    assert err.max() < 1e-4


@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_adjoint_substep_vs_oracle(orc, storage):
    torch = _torch()
    from autoquantizer_amd.adjoint import SubstepAdjoint
    print(f"[elastic3d] adjoint {storage}", flush=True)
    sim, s, st = scene(orc, 300, storage, seed=3)
    rng = np.random.default_rng(4)
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


def test_g2p_codec_bit_exact_benchmark_record(orc):
    """The reference's performance record (24 fixed-point members, 383
    bits in one bit_pack, dithered): the quantized kernel's words equal the
    oracle's encode of the fp32-storage kernel's G2P on the same grid."""
    print("[elastic3d] g2p codec", flush=True)
    sim_q, s, _ = scene(orc, 4000, "f32", seed=5, quant_bits=BENCH_BITS)
    assert sim_q.nwords == 6
    sim_f, _, _ = scene(orc, 4000, "f32", seed=5)
    sim_f.set_state(sim_q.get_state(original_order=True))
    sim_q.build_pid()
    sim_f.build_pid()
    sim_q.p2g_bls()
    sim_q.grid_op()
    sim_f.load_grid(sim_q.grid_velocity())
    sim_q.g2p_bls()
    sim_f.g2p_bls()
    out = sim_f.get_state(original_order=True).astype(np.float32)
    words = orc.pack(to_oracle(sim_q.comps), sim_q.nwords, out, seed=11, step=0)
    assert np.array_equal(sim_q.get_words(original_order=True), words)


def test_letters_emission():
    """add_mesh with the reference's letters mesh (scale (0.1, 0.1, 2),
    offset (0.02, 0.02, 0.5), run_mpm_elastic_3d.py:317) and sample
    density 1: count_emit particles, every one inside a solver cell scaled
    from an inside voxel, velocity (0, -4, 0); 20 substeps stay finite."""
    torch = _torch()
    import autoquantizer_amd as aq
    from autoquantizer_amd.scenes import letters_triangles
    from autoquantizer_amd.voxelizer import Voxelizer, count_emit
    print("[elastic3d] letters", flush=True)
    cfg = aq.MPMConfig()
    cfg.dim, cfg.use_fluid, cfg.n_grid, cfg.dt, cfg.E, cfg.nu = 3, False, 64, 1e-4, 5e5, 0.2
    cfg.size = 4.0
    cfg.p_rho = 1000.0
    cfg.p_vol = (1 / 64 * 0.5) ** 3
    cfg.g = 9.8 * 2.5
    cfg.n_particles = 400000
    cfg.compute_fp = "f32"
    sim = aq.MPMSolver(cfg)
    tri = letters_triangles()
    want = count_emit(Voxelizer((64,) * 3, 1 / 64).voxelize(tri), 1)
    got = sim.add_mesh(tri, sample_density=1, velocity=(0.0, -4.0, 0.0), translation=(1.5, 0.8, 0.1))
    assert got == want and sim.current_idx == want and want > 5000
    st = sim.get_state()
    assert np.allclose(st[4], -4.0) and np.allclose(st[15], 1.0)
    assert st[0].min() >= 1.5 and st[1].min() >= 0.8 and st[2].min() >= 0.1
    sim.advance(20, sort_every=8)
    torch.cuda.synchronize()
    assert np.all(np.isfinite(sim.get_state()))
