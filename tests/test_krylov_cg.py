"""CG (tenet_amd.krylov) against the reference's own CG tests and known answers.

The first five tests restate the unit tests of tenet-krylov/src/cg.rs:190-380
(their systems, options and expected solutions / statistics); the rest pin
what the mathematics fixes: CG on an SPD system agrees with a direct solve,
and a zero right-hand side converges before any matvec.
"""
import math

import numpy as np
import pytest

from tenet_amd.krylov import CgBreakdown, CgOptions, HostVector, cg


def _op(f):
    return lambda v: HostVector(f(v.a))


def test_solves_small_real_spd_system():
    """cg.rs:194-205: [[4, 1], [1, 3]] x = [1, 2] -> x = [1/11, 7/11]."""
    res = cg(_op(lambda x: np.array([4 * x[0] + x[1], x[0] + 3 * x[1]])), HostVector([1.0, 2.0]))
    assert res.stats.converged and res.stats.breakdown is None
    assert abs(res.solution.a[0] - 1 / 11) <= 1e-10 and abs(res.solution.a[1] - 7 / 11) <= 1e-10


def test_solves_small_complex_hermitian_spd_system():
    """cg.rs:278-297: A = [[2, i], [-i, 2]], b = [3+4i, 5-3i] -> x = [1+i, 2-i]."""
    A = np.array([[2, 1j], [-1j, 2]])
    res = cg(_op(lambda x: A @ x), HostVector(np.array([3 + 4j, 5 - 3j])))
    assert res.stats.converged
    assert np.allclose(res.solution.a, [1 + 1j, 2 - 1j], atol=1e-10, rtol=0)


def test_damping_changes_diagonal_solution():
    """cg.rs:299-313: diag(2, 4) with damping 1 and b = [6, 10] -> x = [2, 2]."""
    res = cg(_op(lambda x: np.array([2 * x[0], 4 * x[1]])), HostVector([6.0, 10.0]), CgOptions(damping=1.0))
    assert res.stats.converged
    assert np.allclose(res.solution.a, [2.0, 2.0], atol=1e-10, rtol=0)


def test_detects_non_positive_curvature_for_negative_operator():
    """cg.rs:315-330: -I stops before the first step, after one matvec."""
    res = cg(_op(lambda x: -x), HostVector([1.0, 2.0]))
    assert not res.stats.converged
    assert res.stats.breakdown is CgBreakdown.NON_POSITIVE_CURVATURE
    assert res.stats.iterations == 0 and res.stats.matvecs == 1


def test_reports_max_iterations_without_panic():
    """cg.rs:332-350: max_iter 1 with zero tolerances."""
    res = cg(_op(lambda x: np.array([4 * x[0] + x[1], x[0] + 3 * x[1]])), HostVector([1.0, 2.0]),
             CgOptions(max_iter=1, rtol=0.0, atol=0.0))
    assert not res.stats.converged
    assert res.stats.breakdown is CgBreakdown.MAX_ITERATIONS and res.stats.iterations == 1


def test_rejects_invalid_options_without_matvec():
    """cg.rs:352-366: a negative rtol is rejected before any matvec."""
    def boom(_):
        raise AssertionError("invalid options must stop before the matvec")
    res = cg(boom, HostVector([1.0, 2.0]), CgOptions(rtol=-1.0))
    assert not res.stats.converged
    assert res.stats.breakdown is CgBreakdown.INVALID_OPTIONS and res.stats.matvecs == 0


@pytest.mark.parametrize("damping", [0.0, 0.3])
def test_spd_system_matches_direct_solve(damping):
    rng = np.random.default_rng(3)
    M = rng.standard_normal((60, 60))
    A = M @ M.T + 0.5 * np.eye(60)
    b = rng.standard_normal(60)
    res = cg(_op(lambda x: A @ x), HostVector(b), CgOptions(rtol=1e-13, damping=damping))
    want = np.linalg.solve(A + damping * np.eye(60), b)
    assert res.stats.converged and res.stats.matvecs == res.stats.iterations
    assert res.stats.final_residual <= res.stats.tolerance == pytest.approx(1e-13 * np.linalg.norm(b))
    assert np.linalg.norm(res.solution.a - want) <= 1e-9 * np.linalg.norm(want)


def test_zero_rhs_converges_without_matvec():
    res = cg(_op(lambda x: 2 * x), HostVector(np.zeros(4)))
    assert res.stats.converged and res.stats.matvecs == 0 and res.stats.iterations == 0
    assert not res.solution.a.any()


def test_non_finite_rhs_is_reported():
    res = cg(_op(lambda x: 2 * x), HostVector([1.0, math.nan]))
    assert res.stats.breakdown is CgBreakdown.NON_FINITE_RESIDUAL and res.stats.matvecs == 0
