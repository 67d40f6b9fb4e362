"""Quantized records with f64 compute (the reference's 2D quantized runs:
init_quantaichi computes in compute_fp, f64 by default, mpm_solver.py:372-374)
on the GPU against the oracle's f64 codec and f64 quantized substep
(oracle/codec_ref.c orc_pack_f64 / orc_unpack_f64, oracle/mpm_ref.c
orc_run_quant_f64), on the reference's fixed-point record of the 2D jelly
(init_quantaichi: every component signed fixed point in one bit_pack,
ranges of elastic_2d with x at 1):

- aq_sim_pack / aq_sim_unpack: bit-exact with the oracle's f64 encode and
  decode (dithered with the set-up step's noise);
- the codec fused into G2P: the f64-compute quantized kernel's records equal
  the oracle's f64 encode of what the plain f64 kernel computes from the
  same state on the same grid, bit for bit;
- whole substeps re-seeded from identical records against orc_run_quant_f64:
  every member within one quantization step, nearly every record identical
  (the P2G sums of the two differ in order, in f64)."""
import os

import numpy as np
import pytest

from gpu_util import to_oracle

pytestmark = pytest.mark.gpu

BITS = [26, 26, 20, 20, 12, 15, 15, 12, 21, 21, 21, 21]
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def scene(orc, n, seed=0):
    _torch()
    import autoquantizer_amd as aq
    from autoquantizer_amd.search import elastic_2d_ranges, quantized_solver
    cfg = aq.MPMConfig()
    cfg.dim, cfg.use_fluid, cfg.n_grid, cfg.dt, cfg.E, cfg.nu = 2, False, 128, 2e-4, 5e1, 0.2
    cfg.p_vol = (1 / 128 * 0.5) ** 2
    cfg.n_particles = n
    cfg.compute_fp = "f64"
    cfg.bound = [3, 3]
    cfg.seed = 7
    plain = aq.MPMSolver(cfg)
    ranges = elastic_2d_ranges(np.load(os.path.join(GOLDEN, "ranges_elastic_2d_t8193.npy")), 2)
    q = quantized_solver(plain, BITS, ranges, seed=7, compute_fp="f64")
    assert q.compute_f64
    rng = np.random.default_rng(seed)
    st = np.zeros((12, n))
    st[:2] = rng.uniform(0.2, 0.8, (2, n))
    st[2:4] = rng.uniform(-3, 3, (2, n))
    st[4:8] = rng.uniform(-20, 20, (4, n))
    st[8:] = np.eye(2).reshape(-1, 1) + rng.uniform(-0.1, 0.1, (4, n))
    s = orc.mpm_params(2, orc.ELASTIC, 128, 2e-4, E=5e1, nu=0.2)
    return plain, q, s, st


def test_sim_pack_unpack_bit_exact(orc):
    print("[f64 quant] pack / unpack", flush=True)
    _, q, _, st = scene(orc, 5000, seed=1)
    q.set_state(st, dither_step=123)
    words = q.get_words()
    ref = orc.pack(to_oracle(q.comps), q.nwords, st, seed=7, step=123, f64=True)
    assert np.array_equal(words, ref)
    dec = q.get_state()
    assert np.array_equal(dec, orc.unpack(to_oracle(q.comps), q.nwords, ref, f64=True))
    # 26-bit x resolves below f32: the f64 decode is not an f32 value
    assert np.any(dec[0] != dec[0].astype(np.float32))


def test_g2p_codec_bit_exact(orc):
    print("[f64 quant] g2p codec", flush=True)
    plain, q, _, st = scene(orc, 4000, seed=2)
    q.set_state(st)
    plain.set_state(q.get_state(original_order=True))
    q.build_pid()
    plain.build_pid()
    q.p2g_bls()
    q.grid_op()
    plain.load_grid(q.grid_velocity())
    q.g2p_bls()
    plain.g2p_bls()
    out = plain.get_state(original_order=True)
    words = orc.pack(to_oracle(q.comps), q.nwords, out, seed=7, step=0, f64=True)
    got = q.get_words(original_order=True)
    print("records identical:", np.mean((got == words).all(axis=0)), flush=True)
    assert np.array_equal(got, words)


def test_substeps_reseeded_vs_oracle(orc):
    print("[f64 quant] substeps", flush=True)
    _, q, s, st = scene(orc, 3000, seed=3)
    q.set_state(st)
    oc = to_oracle(q.comps)
    scales = np.array([c["scale"] for c in q.comps])[:, None]
    same = []
    for t in range(5):
        w0 = q.get_words(original_order=True)
        q.substep()
        wg = q.get_words(original_order=True)
        wo = w0.copy()
        orc.run_quant(s, oc, q.nwords, wo, 1, seed=7, step0=t, f64=True)
        a = orc.unpack(oc, q.nwords, wg, f64=True)
        b = orc.unpack(oc, q.nwords, wo, f64=True)
        assert np.all(np.abs(a - b) <= scales * (1 + 1e-9)), t
        same.append(float(np.mean((wg == wo).all(axis=0))))
    print("records identical per substep:", same, flush=True)
    assert min(same) > 0.95
