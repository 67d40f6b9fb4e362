"""Solver checkpoints on the GPU: the reference's dump_states /
checkpoint_states / load_states files (mpm_exps/engine/mpm_solver.py:
1768-1788, 1868-1895) restore the decoded state, and save_records /
load_records restore the records bit for bit and the run continues where it
stopped."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _fluid(n=20000):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import autoquantizer_amd as aq
    from autoquantizer_amd import layouts
    cfg = aq.MPMConfig()
    cfg.dim, cfg.use_fluid, cfg.n_grid, cfg.dt, cfg.la = 3, True, 64, 2e-4, 400.0
    cfg.p_vol = (1 / 64 * 0.5) ** 3
    cfg.n_particles = n
    cfg.seed = 5
    cfg.compute_fp = "f32"
    _, t, g, sh = layouts.fluid_3d(dither=False)
    sim = aq.MPMSolver(cfg, types=t, groups=g, shared=sh)
    rng = np.random.default_rng(3)
    st = np.zeros((16, n))
    st[:3] = rng.uniform(0.2, 0.5, (3, n))
    st[3:6] = rng.uniform(-1, 1, (3, n))
    st[15] = 1.0
    sim.set_state(st)
    sim.advance(10, sort_every=4)
    return sim


def test_reference_checkpoint_files_restore_the_state(tmp_path):
    from autoquantizer_amd.io import load_particles_only
    print("[io] reference files", flush=True)
    a = _fluid()
    before = a.get_state()
    fn = a.dump_states(str(tmp_path), 12)
    a.checkpoint_states(str(tmp_path))
    assert fn.endswith("data_000012.ply")
    assert np.array_equal(load_particles_only(fn), before[:3].T.astype(np.float32))
    assert np.load(tmp_path / "C.npy").shape == (before.shape[1], 3, 3)
    b = _fluid()
    b.load_states(str(tmp_path), 12)
    after = b.get_state()
    # the files hold the decoded values; re-encoding a decoded value gives back its code
    assert np.array_equal(after, before)


def test_record_checkpoint_resumes_bit_for_bit(tmp_path):
    from autoquantizer_amd.io import load_records, save_records
    print("[io] records", flush=True)
    a = _fluid()
    save_records(a, str(tmp_path / "ck.npz"))
    w0 = a.get_words()
    b = _fluid(20000)
    b.advance(3, sort_every=2)  # a different state, then restored
    load_records(b, str(tmp_path / "ck.npz"))
    assert np.array_equal(b.get_words(), w0)
    assert b.step_count == a.step_count
    a.advance(20, sort_every=8)
    b.advance(20, sort_every=8)
    xa, xb = a.get_state(), b.get_state()
    # the continued runs agree to the rounding of their sums (the sort's
    # order inside a cell is not fixed)
    assert np.sqrt(np.mean((xa[:3] - xb[:3]) ** 2)) < 1e-6
    # a checkpoint of another scheme is refused
    import autoquantizer_amd as aq
    cfg = aq.MPMConfig()
    cfg.dim, cfg.use_fluid, cfg.n_grid, cfg.n_particles, cfg.compute_fp = 3, True, 64, 20000, "f32"
    plain = aq.MPMSolver(cfg)
    with pytest.raises(ValueError):
        load_records(plain, str(tmp_path / "ck.npz"))
