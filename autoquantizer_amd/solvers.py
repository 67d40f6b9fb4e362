"""Analytical bit allocation from error-contribution sums (FP64, host side).

The quantization error of a component with range R and b fraction bits is
modelled as uniform of width Delta = R 2^-b, so its contribution to the
squared error of the output is g Delta^2 / 12 with g the summed squared
sensitivity of the output to that component (the adjoint's error-contribution
sum).  With G_i = g_i R_i^2 and x_i = (Delta_i / R_i)^2 = 4^-b_i:

- LagrangianCR minimises sum_i G_i x_i / 12 subject to a memory budget
  sum_i a_i (b_i + 1) = 32 r sum_i a_i (r: compression rate against f32,
  +1 for the sign bit); stationarity gives G_i x_i / a_i = const.
- LagrangianEB meets an error bound sum_i G_i x_i / 12 = (tol_rel z)^2 with
  the fewest bits: G_i x_i / a_i = const again, scaled to the bound.

Both follow solvers/analytical_solver.py:3-84 of the reference (same closed
forms, bits floored for CR and ceiled for EB, returned without the sign bit;
callers add 1, run_mpm_fluid_3d.py:149-161, run_mpm_elastic_2d.py:227-242).
"""
import numpy as np

_LOG2_LN2 = np.log2(np.log(2.0))


def _weights(n, a):
    if a is None:
        return np.ones(n, dtype=np.float64)
    a = np.asarray(a, dtype=np.float64)
    if a.shape != (n,):
        raise ValueError(f"a must have shape ({n},), got {a.shape}")
    return a


def lagrangian_cr_continuous(grads, r=0.5, a=None):
    """Real-valued optimum of the compression-rate problem: (bits, log2 of
    the multiplier).  grads = g_i R_i^2 (> 0)."""
    g = np.asarray(grads, dtype=np.float64)
    a = _weights(len(g), a)
    if np.any(~(g > 0)) or np.any(~np.isfinite(g)):
        raise ValueError("error-contribution sums must be finite and positive")
    budget = r * 32.0 - 1.0  # mean fraction bits (one bit of each component is its sign)
    t = np.log2(a / g)
    log_la = -2.0 * budget - (t @ a) / a.sum() + _LOG2_LN2
    bits = -(t + log_la - _LOG2_LN2) / 2.0
    return bits, log_la


def LagrangianCR(grads, r=0.5, a=None):
    """Fraction bits (int32, sign bit excluded) and the modelled error
    sum_i G_i x_i / 12 at the real-valued optimum, for compression rate r."""
    g = np.asarray(grads, dtype=np.float64)
    a = _weights(len(g), a)
    bits, log_la = lagrangian_cr_continuous(g, r, a)
    x = (a / g) * np.exp2(log_la) / np.log(2.0)  # = 4^-bits at the optimum
    err = float(x @ g / 12.0)
    return np.floor(bits).astype(np.int32), err


def LagrangianEB(grads, ref, tol_rel=1e-3, a=None):
    """Fraction bits (int32, sign bit excluded) meeting the relative error
    bound tol_rel on a quantity of magnitude ref, and the relative error the
    real-valued optimum reaches (= tol_rel)."""
    g = np.asarray(grads, dtype=np.float64)
    a = _weights(len(g), a)
    if np.any(~(g > 0)) or np.any(~np.isfinite(g)):
        raise ValueError("error-contribution sums must be finite and positive")
    tol = (ref * tol_rel) ** 2
    x = 12.0 / a.sum() * tol * (a / g)  # (Delta / R)^2 at the optimum
    bits = np.ceil(-np.log2(x) / 2.0).astype(np.int32)
    rel = float(np.sqrt(x @ g / (12.0 * ref ** 2)))
    return bits, rel


def modelled_error(grads, bits):
    """sum_i G_i 4^-bits_i / 12: the modelled squared output error of a
    bit allocation (fraction bits, sign excluded)."""
    g = np.asarray(grads, dtype=np.float64)
    return float(g @ np.exp2(-2.0 * np.asarray(bits, dtype=np.float64)) / 12.0)
