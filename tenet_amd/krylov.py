"""Matrix-free Conjugate Gradient (the reference's Krylov solver) over device vectors.

tenet-krylov has one solver, CG on a ``LinearOperator`` (operator.rs:2-15) over
``KrylovVector`` values (vector.rs:2-48): zero initial guess, stop when
||r|| <= atol + rtol ||b||, optional damping (solves (A + damping I) x = b),
a curvature check, and every exceptional exit reported in the statistics
instead of raised (cg.rs:3-190).  ``cg`` here keeps that contract -- the same
options and defaults, statistics and breakdown reasons, matvec count and
stopping rule -- so a caller of the reference finds the same behaviour.

Vectors are any object with the reference's vector operations:
``zero_like()``, ``clone()``, ``axpy(alpha, x)`` (self += alpha x),
``scale(alpha)``, ``dot_real(rhs)`` and ``norm2()``.  ``DeviceVector`` adapts a
``TensorMap``: the payload stays in HBM, axpy/scale run on the device and
``dot_real`` is the FP64 weighted inner product sum_c dim(c) <x_c, y_c>
computed on the device (typed.rs:699-716, 11210-11262), the inner product of
the reference's TensorMap vectors.  ``HostVector`` wraps a numpy array (real or
complex; dot_real = Re <x, y>) for host operators.
"""
import enum
import math
from dataclasses import dataclass
from typing import Any, Optional

import numpy as np

from .tensor import TensorMap


class CgBreakdown(enum.Enum):
    """Reasons CG stops without converging (cg.rs:48-56)."""
    INVALID_OPTIONS = "InvalidOptions"
    NON_FINITE_RESIDUAL = "NonFiniteResidual"
    NON_POSITIVE_CURVATURE = "NonPositiveCurvature"
    NON_FINITE_CURVATURE = "NonFiniteCurvature"
    MAX_ITERATIONS = "MaxIterations"


@dataclass(frozen=True)
class CgOptions:
    """CgOptions and its defaults (cg.rs:3-28)."""
    rtol: float = 1.0e-12
    atol: float = 0.0
    max_iter: int = 1000
    damping: float = 0.0
    check_curvature: bool = True


@dataclass
class CgStats:
    """CgStats (cg.rs:37-46)."""
    iterations: int = 0
    matvecs: int = 0
    converged: bool = False
    initial_residual: float = 0.0
    final_residual: float = 0.0
    tolerance: float = 0.0
    breakdown: Optional[CgBreakdown] = None


@dataclass
class CgResult:
    solution: Any
    stats: CgStats


def _finite(*xs):
    return all(math.isfinite(x) for x in xs)


def cg(op, b, options: Optional[CgOptions] = None) -> CgResult:
    """Solve op(x) = b, or (op + damping I) x = b, by Conjugate Gradient
    (cg.rs:58-190).  op: callable vector -> vector; b: a vector as described in
    the module docstring.  Returns the solution and its statistics; breakdowns
    (invalid options, non-finite residual or curvature, non-positive
    curvature, iteration limit) are reported in ``stats.breakdown``."""
    opt = options or CgOptions()
    x = b.zero_like()
    r = b.clone()
    p = r.clone()
    b_norm = b.norm2()
    r_norm = r.norm2()
    tol = opt.atol + opt.rtol * b_norm
    st = CgStats(initial_residual=r_norm, final_residual=r_norm, tolerance=tol)

    def stop(reason=None, converged=False):
        st.breakdown, st.converged = reason, converged
        return CgResult(x, st)

    options_ok = _finite(opt.rtol, opt.atol, opt.damping) and min(opt.rtol, opt.atol, opt.damping) >= 0.0
    if not options_ok:
        return stop(CgBreakdown.INVALID_OPTIONS)
    if not _finite(r_norm, b_norm, tol):
        return stop(CgBreakdown.NON_FINITE_RESIDUAL)
    if r_norm <= tol:
        return stop(converged=True)
    rho = r.dot_real(r)
    if not math.isfinite(rho) or rho < 0.0:
        return stop(CgBreakdown.NON_FINITE_RESIDUAL)

    while st.iterations < opt.max_iter:
        ap = op(p)
        st.matvecs += 1
        if opt.damping != 0.0:
            ap.axpy(opt.damping, p)
        curv = p.dot_real(ap)
        if not math.isfinite(curv):
            return stop(CgBreakdown.NON_FINITE_CURVATURE)
        if opt.check_curvature and curv <= 0.0:
            return stop(CgBreakdown.NON_POSITIVE_CURVATURE)
        step = rho / curv
        if not math.isfinite(step):
            return stop(CgBreakdown.NON_FINITE_CURVATURE)
        x.axpy(step, p)
        r.axpy(-step, ap)
        st.iterations += 1
        r_norm = r.norm2()
        st.final_residual = r_norm
        if not math.isfinite(r_norm):
            return stop(CgBreakdown.NON_FINITE_RESIDUAL)
        if r_norm <= tol:
            return stop(converged=True)
        rho_next = r.dot_real(r)
        if not math.isfinite(rho_next) or rho_next < 0.0:
            return stop(CgBreakdown.NON_FINITE_RESIDUAL)
        beta = rho_next / rho
        if not math.isfinite(beta):
            return stop(CgBreakdown.NON_FINITE_RESIDUAL)
        p.scale(beta)
        p.axpy(1.0, r)
        rho = rho_next
    return stop(CgBreakdown.MAX_ITERATIONS)


def _sqrt_or_nan(v):
    return math.sqrt(v) if v >= 0.0 else float("nan")


class DeviceVector:
    """KrylovVector over a device TensorMap (payload in HBM)."""

    def __init__(self, t: TensorMap):
        self.t = t

    def zero_like(self):
        return DeviceVector(TensorMap.zeros(self.t.rt, self.t.space))

    def clone(self):
        return DeviceVector(self.t.copy())

    def axpy(self, alpha, x):
        self.t.axpy_(x.t, alpha)

    def scale(self, alpha):
        self.t.scale_(alpha)

    def dot_real(self, rhs):
        return self.t.inner(rhs.t)

    def norm2(self):
        return _sqrt_or_nan(self.dot_real(self))


class HostVector:
    """KrylovVector over a host numpy array (real or complex)."""

    def __init__(self, a):
        self.a = np.array(a)

    def zero_like(self):
        return HostVector(np.zeros_like(self.a))

    def clone(self):
        return HostVector(self.a.copy())

    def axpy(self, alpha, x):
        self.a += alpha * x.a

    def scale(self, alpha):
        self.a *= alpha

    def dot_real(self, rhs):
        return float(np.real(np.vdot(self.a, rhs.a)))

    def norm2(self):
        return _sqrt_or_nan(self.dot_real(self))
