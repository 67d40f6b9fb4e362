"""Quantization-scheme search: ranges, error-contribution sums, bit allocation.

The reference's pipeline per scene (run_mpm_elastic_2d.py:213-242,
run_mpm_fluid_3d.py:149-161): a forward run recording per-component ranges
(run_forward_detect_ranges, mpm_solver.py:933-942), an AD run producing the
error-contribution sums (run_forward_with_ad_nlogn + get_gradients), and the
FP64 analytical solve (LagrangianCR / LagrangianEB) on g_i R_i^2.  Here the
two runs are GPU runs of the plain-storage solver and the solve is host FP64.
"""
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
    (LagrangianCR); or tol_rel with reference magnitude ref (LagrangianEB).
    Returns dict(bits (sign included), ranges, grads, loss, seconds per phase)."""
    times = {}
    t0 = time.perf_counter()
    sim = make_solver()
    sim.set_state(init_state)
    ranges = adjust(detect_ranges(sim, nsteps), sim.dim)
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
        bits, _ = LagrangianEB(G, ref=L if ref is None else ref, tol_rel=tol_rel)
    times["solve"] = time.perf_counter() - t2
    times["total"] = time.perf_counter() - t0
    return dict(bits=bits + 1, ranges=ranges, grads=g, loss=L, seconds=times)
