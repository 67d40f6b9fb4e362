"""Numerical (SLSQP) bit-allocation solvers and the optimality-check harness.

ScipySolver / ScipySolverMemMode follow solvers/scipy_solver.py:25-134 of
the reference: the quantization step x_i of each component (uniform noise of
width x_i, error model g_i x_i^2) is optimised with scipy's SLSQP, then
turned into bits ceil(-log2(sqrt(3) x_i / r_i)).  (The optimisation runs on
the bit widths, see the classes; same problem, same optimum.)
- ScipySolver (error-bound mode) maximises sum_i a_i log2(1 / x_i) (fewest
  bits) subject to sum_i g_i x_i^2 <= eps^2.
- ScipySolverMemMode minimises 1/2 sum_i g_i x_i^2 subject to a memory budget
  sum_i -log2(sqrt(3) x_i / r_i) <= 32 eps n_vars (eps: compression rate).
Both are the numerical twins of the closed forms in solvers.py (LagrangianEB
and LagrangianCR); tests/test_optimality.py checks their solutions against
those closed forms and the KKT conditions.

optimality_check follows run_mpm_optimality_check.py:194-234: starting from
an allocation, take S bits from R randomly chosen components (task 0/1), or
move S1 bits from x to v and S2 from C to F (task 2), then run the
quantized scene N times with dithering and report the success fraction.
"""
import math

import numpy as np


class SolverBase:
    """n_vars components, weights a, tolerance or rate eps, sums grad,
    ranges r (set_range)."""

    def __init__(self, n_vars, a=None, eps=None, grad=None):
        self.n_vars = int(n_vars)
        self.a = np.ones(self.n_vars) if a is None else np.asarray(a, dtype=np.float64)
        self.eps = eps
        self.grad = np.asarray(grad, dtype=np.float64)
        if self.grad.shape != (self.n_vars,) or np.any(~(self.grad > 0)):
            raise ValueError("grad must hold n_vars positive sums")
        self.sol = None
        self.r = np.ones(self.n_vars)

    def set_range(self, _range):
        self.r = np.asarray(_range, dtype=np.float64).copy()

    def run(self):
        raise NotImplementedError


def _step(b, r):
    """Quantization step of b fraction bits over range r: x = r 2^-b / sqrt(3)
    (the inverse of the bits formula ceil(-log2(sqrt(3) x / r)))."""
    return r * np.exp2(-b) / math.sqrt(3)


class ScipySolver(SolverBase):
    """Error-bound mode (scipy_solver.py:25-75): fewest bits
    sum_i a_i b_i with sum_i g_i x_i^2 <= eps^2.

    SLSQP runs on the bit widths b_i = -log2(sqrt(3) x_i / r_i), with the
    constraint as log2(sum_i g_i x_i^2 / eps^2) <= 0: the same optimum as the
    reference's formulation over x_i, which SLSQP cannot follow when the sums
    span many decades (the shipped 2D sums span 1e4..1e18)."""

    def __init__(self, n_vars, a=None, eps=None, grad=None):
        super().__init__(n_vars, a, eps, grad)

    def _err(self, b):
        return np.dot(self.grad, _step(b, self.r) ** 2)

    def run(self):
        from scipy.optimize import minimize
        G = self.grad * self.r ** 2 / 3.0
        # start from the equal-contribution point of the bound
        b0 = 0.5 * np.log2(G * self.n_vars / self.eps ** 2)
        cons = {"type": "ineq", "fun": lambda b: -np.log2(self._err(b) / self.eps ** 2),
                "jac": lambda b: (2.0 * self.grad * _step(b, self.r) ** 2) / self._err(b)}
        res = minimize(lambda b: float(np.dot(self.a, b)), b0, jac=lambda b: self.a, method="SLSQP",
                       constraints=[cons], options={"maxiter": 10000, "ftol": 1e-12})
        self.bits_continuous = np.array(res.x)
        self.sol = _step(self.bits_continuous, self.r)
        return res

    def get_bits(self):
        return np.ceil(-np.log2(np.abs(math.sqrt(3) * self.sol / self.r)))


class ScipySolverMemMode(SolverBase):
    """Memory-budget mode (scipy_solver.py:78-134): least modelled error
    1/2 sum_i g_i x_i^2 with sum_i b_i <= 32 eps n_vars, b_i =
    -log2(sqrt(3) x_i / r_i).  SLSQP runs on the b_i (objective scaled by its
    value at the start), the same optimum as the reference's formulation
    over x_i."""

    def __init__(self, n_vars, a=None, eps=None, grad=None):
        super().__init__(n_vars, a, eps, grad)
        self.lb = 1e-32

    def run(self):
        from scipy.optimize import minimize
        budget = self.eps * self.n_vars * 32
        G = self.grad * self.r ** 2 / 3.0
        b0 = np.full(self.n_vars, budget / self.n_vars)
        f0 = 0.5 * np.dot(G, np.exp2(-2 * b0))
        obj = lambda b: 0.5 * np.dot(G, np.exp2(-2 * b)) / f0  # noqa: E731
        jac = lambda b: -np.log(2.0) * G * np.exp2(-2 * b) / f0  # noqa: E731
        cons = {"type": "ineq", "fun": lambda b: budget - np.sum(b), "jac": lambda b: -np.ones_like(b)}
        res = minimize(obj, b0, jac=jac, method="SLSQP", constraints=[cons], options={"maxiter": 10000, "ftol": 1e-15})
        self.bits_continuous = np.array(res.x)
        self.sol = _step(self.bits_continuous, self.r)
        return res

    def get_bits(self):
        return np.ceil(-np.log2(self.lb + np.abs(math.sqrt(3) * self.sol) / self.r))


def perturb_bits(bits, dim, task, search_bits=1, search_bits_2=1, seed=0):
    """The allocations of run_mpm_optimality_check.py's tasks (:194-234):
    task 0: S bits fewer on all components (R = n), task 1: on R = n/2
    random components, task 2: S1 bits moved from x to v and S2 from C to F
    (2D elastic / 3D elastic component order x, v, C, F)."""
    b = np.array(bits, dtype=np.int64).copy()
    n = len(b)
    if task in (0, 1):
        R = n if task == 0 else n // 2
        a = np.random.default_rng(seed).permutation(n)
        b[a[:R]] -= search_bits
    elif task == 2:
        b[:dim] -= search_bits
        b[dim:2 * dim] += search_bits
        b[2 * dim:2 * dim + dim * dim] -= search_bits_2
        b[2 * dim + dim * dim:2 * dim + 2 * dim * dim] += search_bits_2
    else:
        raise ValueError("task must be 0, 1 or 2")
    return b


def optimality_check(plain, init_state, nsteps, bits, ranges, ref_loss, eps, tasks=((0, 1), (1, 1), (2, 1)),
                     runs=10, seed=0):
    """The allocation and its perturbations, each run `runs` times
    quantized with dithering (search.validate_scheme): success fraction and
    mean relative loss error per allocation."""
    from .search import validate_scheme
    out = []
    base = validate_scheme(plain, init_state, nsteps, bits, ranges, ref_loss, eps, runs=runs, seed=seed)
    out.append(dict(task="base", bits=base["bits"].tolist(), success=base["success"],
                    rel_error=float(np.mean(base["rel_error"]))))
    for task, S in tasks:
        pb = perturb_bits(bits, plain.dim, task, S, S, seed=seed)
        r = validate_scheme(plain, init_state, nsteps, pb, ranges, ref_loss, eps, runs=runs, seed=seed + 1 + task)
        out.append(dict(task=task, search_bits=S, bits=r["bits"].tolist(), success=r["success"],
                        rel_error=float(np.mean(r["rel_error"]))))
    return out
