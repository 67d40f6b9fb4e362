"""Quantization-scheme search: ranges, error-contribution sums, bit allocation.

The reference's pipeline per scene (run_mpm_elastic_2d.py:213-242,
run_mpm_fluid_3d.py:149-161): a forward run recording per-component ranges
(run_forward_detect_ranges, mpm_solver.py:933-942), an AD run producing the
error-contribution sums (run_forward_with_ad_nlogn + get_gradients), and the
FP64 analytical solve (LagrangianCR / LagrangianEB) on g_i R_i^2.  Here the
two runs are GPU runs of the plain-storage solver and the solve is host FP64.
"""
import copy
import time

import numpy as np
import torch

from .adjoint import energy_loss, error_contributions
from .solvers import LagrangianCR, LagrangianEB


def detect_ranges(solver, nsteps, sort_every=8):
    """Per-component max |value| over the states after each of nsteps
    substeps (detect_bound_kernel after every substep), in get_limits order
    (x entries set to 2 as the reference does).  Advances the solver."""
    solver.init_limits()
    solver.advance_detect(int(nsteps), sort_every=sort_every)
    return solver.get_limits()


def elastic_2d_ranges(ranges, dim):
    """Range adjustment of run_mpm_elastic_2d.py:229-231."""
    r = np.array(ranges, dtype=np.float64)
    r[:dim] = 1.0
    r[dim:] *= 2.0
    return r


def fluid_3d_ranges(ranges, dim):
    """Range adjustment of run_mpm_fluid_3d.py:150-155."""
    r = np.array(ranges, dtype=np.float64)
    r[2 * dim:2 * dim + dim * dim] *= 24
    r[dim:2 * dim] *= 1.5
    r[-1] *= 1.5
    r[:dim] = 2.0
    return r


def search_scheme(make_solver, init_state, nsteps, rate=0.5, tol_rel=None, ref=None,
                  adjust=elastic_2d_ranges, checkpoint_every=16, loss=energy_loss):
    """Bit allocation for the scene `init_state` ([ncomp, n]) simulated by
    plain-storage solvers from make_solver().  rate: compression-rate target
    (LagrangianCR); or tol_rel: relative error bound on the loss
    (LagrangianEB) with reference magnitude ref (default: the range run's
    final loss, the reference's f64_data).  Returns dict(bits (sign
    included), ranges, grads, loss (the range run's), ad_loss, seconds per
    phase)."""
    times = {}
    t0 = time.perf_counter()
    sim = make_solver()
    sim.set_state(init_state)
    ranges = adjust(detect_ranges(sim, nsteps), sim.dim)
    L_range = sim.compute_ek()
    torch.cuda.synchronize()
    times["ranges"] = time.perf_counter() - t0
    t1 = time.perf_counter()
    sim.set_state(init_state)
    L, _, g = error_contributions(sim, nsteps, loss=loss, checkpoint_every=checkpoint_every)
    times["adjoint"] = time.perf_counter() - t1
    t2 = time.perf_counter()
    G = g * ranges ** 2
    if tol_rel is None:
        bits, _ = LagrangianCR(G, r=rate)
    else:
        bits, _ = LagrangianEB(G, ref=L_range if ref is None else ref, tol_rel=tol_rel)
    times["solve"] = time.perf_counter() - t2
    times["total"] = time.perf_counter() - t0
    return dict(bits=bits + 1, ranges=ranges, grads=g, loss=L_range, ad_loss=L, seconds=times)


# ---- the allocation, instantiated and run (run_mpm_elastic_2d.py:186-208, 227-242) ----

def clamp_bits(bits):
    """set_bits_and_ranges (mpm_solver.py:497-499): every width clamped to [1, 32]."""
    return np.clip(np.rint(np.asarray(bits, dtype=np.float64)), 1, 32).astype(np.int32)


def quantized_solver(plain, bits, ranges, dither=True, use_bitpack=True, seed=0, capacity=None, compute_fp="f32"):
    """The quantized solver of a scheme, as the reference's run(quant=True)
    builds it (run_mpm_elastic_2d.py:186-191 -> init_quantaichi,
    mpm_solver.py:364-455): the plain solver's scene with every component a
    signed fixed-point type of bits[c] (clamped to [1, 32]) and range
    ranges[c] (J with offset 1), in one bit_pack (or the reference's
    bit_struct words), dithered.  compute_fp: "f64" computes in double (the
    reference's 2D runs, 2D only), "f32" in float."""
    from .mpm import MPMSolver
    cfg = copy.copy(plain.param)
    cfg.use_quantaichi = True
    cfg.use_bitpack = bool(use_bitpack)
    cfg.use_dithering = bool(dither)
    cfg.quant_bits = clamp_bits(bits)
    cfg.quant_ranges = np.asarray(ranges, dtype=np.float64)
    cfg.compute_fp = compute_fp
    cfg.seed = int(seed)
    return MPMSolver(cfg, capacity=capacity or plain.capacity, device=plain.device)


def quantized_losses(qsim, init_state, nsteps, runs, sort_every=16, loss="energy", first_run=0):
    """Final losses of `runs` runs of the quantized solver from init_state,
    each with its own dithering noise (run r encodes the initial state and
    every substep with noise steps offset by r 2^24), as the reference's
    run(N) repeats init + run_forward_only with the random stream running
    on.  loss: compute_ek of the final state."""
    out = []
    for r in range(first_run, first_run + runs):
        qsim.set_state(init_state, dither_step=(r << 24) | 0xFFFFFF)
        qsim.step_count = r << 24
        qsim.advance(int(nsteps), sort_every=sort_every)
        out.append(qsim.compute_ek())
    return np.array(out)


def success_fraction(losses, ref_loss, eps):
    """The reference's pass criterion (run_mpm_elastic_2d.py:205-207):
    the fraction of runs with |loss - f64| < 3 eps f64."""
    d = np.abs(np.asarray(losses, dtype=np.float64) - ref_loss)
    return float(np.mean(d < 3.0 * eps * abs(ref_loss)))


def validate_scheme(plain, init_state, nsteps, bits, ranges, ref_loss, eps, runs=20, sort_every=16, seed=0,
                    compute_fp="f32"):
    """The reference's experiment for one allocation: `runs` dithered
    quantized runs of the scene (compute in compute_fp), their losses
    against the f64 loss, and the success fraction.  Returns dict(success,
    losses, rel_error (each run's |loss - f64| / f64), bits (clamped),
    seconds)."""
    t0 = time.perf_counter()
    q = quantized_solver(plain, bits, ranges, seed=seed, compute_fp=compute_fp)
    L = quantized_losses(q, init_state, nsteps, runs, sort_every=sort_every)
    torch.cuda.synchronize()
    return dict(success=success_fraction(L, ref_loss, eps), losses=L, rel_error=np.abs(L - ref_loss) / abs(ref_loss),
                bits=clamp_bits(bits), record_bits=int(clamp_bits(bits).sum()), seconds=time.perf_counter() - t0)


def reference_run(make_solver, init_state, nsteps, sort_every=8):
    """The FP64 run of the scene that the range run is: per-component ranges
    (get_limits order) and the final loss (compute_ek).  make_solver(storage)
    builds the plain solver."""
    sim = make_solver("f64")
    sim.set_state(init_state)
    lim = detect_ranges(sim, nsteps, sort_every=sort_every)
    return sim, lim, sim.compute_ek()
