"""The SLSQP bit-allocation solvers (solvers/scipy_solver.py:25-134 of the
reference, autoquantizer_amd.optimality) against the closed forms and the
KKT conditions of their problems, on the reference's shipped 2D elastic
sums and ranges, and the optimality-check perturbations
(run_mpm_optimality_check.py:194-234).  CPU only."""
import math
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def sums_and_ranges():
    g = np.load(os.path.join(GOLD, "grads_elastic_2d.npy"), allow_pickle=False)
    r = np.load(os.path.join(GOLD, "ranges_elastic_2d_t8193.npy"), allow_pickle=False).copy()
    r[:2] = 1.0  # run_mpm_elastic_2d.py:229-231
    r[2:] *= 2.0
    return g, r


@pytest.mark.parametrize("eps", [0.3, 0.4, 0.5])
def test_mem_mode_matches_the_compression_rate_closed_form(eps):
    """min 1/2 sum g x^2 s.t. sum b <= 32 eps n, b = -log2(sqrt(3) x / r):
    stationarity gives g_i x_i^2 = const, i.e. the LagrangianCR optimum
    (whose widths exclude the sign bit, one bit each below)."""
    from autoquantizer_amd.optimality import ScipySolverMemMode
    from autoquantizer_amd.solvers import lagrangian_cr_continuous
    g, r = sums_and_ranges()
    s = ScipySolverMemMode(12, eps=eps, grad=g)
    s.set_range(r)
    res = s.run()
    assert res.success
    b = -np.log2(math.sqrt(3) * s.sol / r)
    cr, _ = lagrangian_cr_continuous(g * r ** 2, r=eps)
    assert np.abs(b - (cr + 1)).max() < 1e-4
    assert abs(b.sum() - 32 * eps * 12) < 1e-6
    gx2 = g * s.sol ** 2
    assert gx2.max() / gx2.min() == pytest.approx(1.0, abs=1e-4)
    assert np.array_equal(s.get_bits(), np.ceil(b - 1e-9))


def test_error_bound_mode_kkt():
    """min sum a b s.t. sum g x^2 <= eps^2: at the optimum the bound is
    active and g_i x_i^2 / a_i is the same for every component; with a = 1
    every component contributes eps^2 / n."""
    from autoquantizer_amd.optimality import ScipySolver
    g, r = sums_and_ranges()
    eps = 1e-3 * 6893.409
    for a in (np.ones(12), np.random.default_rng(0).uniform(0.5, 2.0, 12)):
        s = ScipySolver(12, a=a, eps=eps, grad=g)
        s.set_range(r)
        res = s.run()
        assert res.success
        err = g * s.sol ** 2
        assert err.sum() == pytest.approx(eps ** 2, rel=1e-6)
        ratio = err / a
        assert ratio.max() / ratio.min() == pytest.approx(1.0, abs=1e-4)
    s = ScipySolver(12, eps=eps, grad=g)
    s.set_range(r)
    s.run()
    assert np.allclose(g * s.sol ** 2, eps ** 2 / 12, rtol=1e-6)


def test_optimality_check_perturbations():
    """Task 0: every component S bits fewer; task 1: half of them (random);
    task 2: S1 bits from x to v and S2 from C to F."""
    from autoquantizer_amd.optimality import perturb_bits
    bits = np.arange(10, 22)
    assert np.array_equal(perturb_bits(bits, 2, 0, 2), bits - 2)
    t1 = perturb_bits(bits, 2, 1, 3, seed=4)
    assert np.sum(t1 != bits) == 6 and np.all((bits - t1)[t1 != bits] == 3)
    t2 = perturb_bits(bits, 2, 2, 1, 2)
    assert np.array_equal(t2 - bits, [-1, -1, 1, 1, -2, -2, -2, -2, 2, 2, 2, 2])
    with pytest.raises(ValueError):
        perturb_bits(bits, 2, 7)
