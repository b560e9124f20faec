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
import ctypes as C
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
    host_syncs: Optional[int] = None  # device loop only: host synchronisations taken


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
    curvature, iteration limit) are reported in ``stats.breakdown``.
    Device vectors take the stream-ordered loop (``cg_device``)."""
    if isinstance(b, DeviceVector):
        return cg_device(op, b, options)
    return cg_generic(op, b, options)


def cg_generic(op, b, options: Optional[CgOptions] = None) -> CgResult:
    """The loop of cg.rs:58-190 over any KrylovVector (one host value per dot)."""
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


class DeviceScalars:
    """A small FP64 buffer in HBM holding the scalars of a device Krylov loop:
    dots are queued into it (tnt_weighted_dot_async), scalar steps run on it
    (tnt_scalars), vector updates read their coefficients from it
    (tnt_axpby_dev), and ``read`` downloads it -- the loop's one host
    synchronisation per iteration."""

    def __init__(self, rt, n=16):
        from . import _native as N
        self.rt, self.N, self.n = rt, N, n
        buf = C.c_void_p()
        rt._check(N.hip().tnt_buf_alloc(rt._ctx, n, N.F64, C.byref(buf)))
        self.buf = buf
        rt._check(N.hip().tnt_buf_zero(rt._ctx, buf))
        self.syncs = 0

    def dot(self, x, y, i):
        x.inner_async(y, self.buf, i)

    def program(self, *ops):
        N = self.N
        arr = (N.ScalarOp * len(ops))(*[N.ScalarOp(op, out, a, b, imm) for op, out, a, b, imm in ops])
        self.rt._check(N.hip().tnt_scalars(self.rt._ctx, self.buf, arr, len(ops)))

    def axpby(self, y, x, ai=-1, af=1.0, bi=-1, bf=1.0, guard=-1):
        """y = (af s[ai]) x + (bf s[bi]) y on the device (a missing index means the factor alone)."""
        self.rt._check(self.N.hip().tnt_axpby_dev(self.rt._ctx, y._buf, x._buf, self.buf, ai, af, bi, bf, guard))

    def read(self):
        out = np.empty(self.n)
        self.rt._check(self.N.hip().tnt_buf_download(self.rt._ctx, self.buf, out.ctypes.data, self.n))
        self.syncs += 1
        return out

    def close(self):
        if self.buf:
            self.N.hip().tnt_buf_free(self.rt._ctx, self.buf)
            self.buf = None


# slots of the device CG loop
_RHO, _CURV, _STEP, _GUARD, _RHO_NEXT, _BETA, _BB, _RR = range(8)


def cg_device(op, b, options: Optional[CgOptions] = None) -> CgResult:
    """cg over DeviceVector with the loop's scalars on the device: each
    iteration queues the matvec, p.Ap, the step and its guard, the x / r
    updates (skipped on the device when the step is rejected), r.r, beta and
    the p update, then reads the scalars once.  The host takes the same
    decisions in the same order as cg.rs:108-185 on the same FP64 values (one
    dot serves both r.norm2() and r.dot_real(r): they are the same
    deterministic sum), so statistics, breakdowns and the returned solution
    are those of the host-synchronous loop.  ``stats.host_syncs`` counts the
    synchronisations (1 + iterations)."""
    opt = options or CgOptions()
    rt = b.t.rt
    S = DeviceScalars(rt)
    try:
        x = b.zero_like()
        r = b.clone()
        p = r.clone()
        S.dot(b.t, b.t, _BB)
        S.dot(r.t, r.t, _RR)
        s = S.read()
        b_norm, r_norm = _sqrt_or_nan(s[_BB]), _sqrt_or_nan(s[_RR])
        tol = opt.atol + opt.rtol * b_norm
        st = CgStats(initial_residual=r_norm, final_residual=r_norm, tolerance=tol)

        def stop(reason=None, converged=False):
            st.breakdown, st.converged = reason, converged
            st.host_syncs = S.syncs
            return CgResult(x, st)

        options_ok = _finite(opt.rtol, opt.atol, opt.damping) and min(opt.rtol, opt.atol, opt.damping) >= 0.0
        if not options_ok:
            return stop(CgBreakdown.INVALID_OPTIONS)
        if not _finite(r_norm, b_norm, tol):
            return stop(CgBreakdown.NON_FINITE_RESIDUAL)
        if r_norm <= tol:
            return stop(converged=True)
        rho = s[_RR]
        if not math.isfinite(rho) or rho < 0.0:
            return stop(CgBreakdown.NON_FINITE_RESIDUAL)
        check = 1.0 if opt.check_curvature else 0.0
        S.program((2, _RHO, _RR, _RR, 0.0))  # rho
        while st.iterations < opt.max_iter:
            ap = op(p)
            st.matvecs += 1
            if opt.damping != 0.0:
                ap.axpy(opt.damping, p)
            S.dot(p.t, ap.t, _CURV)
            S.program((1, _STEP, _RHO, _CURV, 0.0), (5, _GUARD, _CURV, _STEP, check))
            S.axpby(x.t, p.t, ai=_STEP, guard=_GUARD)              # x += step p
            S.axpby(r.t, ap.t, ai=_STEP, af=-1.0, guard=_GUARD)    # r -= step Ap
            S.dot(r.t, r.t, _RHO_NEXT)
            S.program((1, _BETA, _RHO_NEXT, _RHO, 0.0), (2, _RHO, _RHO_NEXT, _RHO_NEXT, 0.0))
            S.axpby(p.t, r.t, bi=_BETA, guard=_GUARD)              # p = r + beta p
            s = S.read()
            curv, step = s[_CURV], s[_STEP]
            if not math.isfinite(curv):
                return stop(CgBreakdown.NON_FINITE_CURVATURE)
            if opt.check_curvature and curv <= 0.0:
                return stop(CgBreakdown.NON_POSITIVE_CURVATURE)
            if not math.isfinite(step):
                return stop(CgBreakdown.NON_FINITE_CURVATURE)
            st.iterations += 1
            rho_next = s[_RHO_NEXT]
            r_norm = _sqrt_or_nan(rho_next)
            st.final_residual = r_norm
            if not math.isfinite(r_norm):
                return stop(CgBreakdown.NON_FINITE_RESIDUAL)
            if r_norm <= tol:
                return stop(converged=True)
            if not math.isfinite(rho_next) or rho_next < 0.0:
                return stop(CgBreakdown.NON_FINITE_RESIDUAL)
            if not math.isfinite(s[_BETA]):
                return stop(CgBreakdown.NON_FINITE_RESIDUAL)
        return stop(CgBreakdown.MAX_ITERATIONS)
    finally:
        S.close()


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
