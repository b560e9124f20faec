"""Two-site effective Hamiltonian on the device and a Lanczos ground-state solver.

The reference's Krylov layer (tenet-krylov) is a matrix-free solver over
``LinearOperator::apply`` and ``KrylovVector::dot_real``
(tenet-krylov/src/operator.rs:2-15, vector.rs:2-48, cg.rs:64-190; restated
over device vectors in tenet_amd.krylov); it has no Lanczos and no H_eff.
Here ``TwoSiteHeff.apply`` is the operator (a fixed
chain of four block-sparse contractions over resident environments) and the
FP64 inner products run on the device (tnt_weighted_dot).

Index conventions (legs of V = bond, P = physical, M = MPO bond):
    theta in V <- P (x) P (x) V          theta[a; s1, s2, c]
    L     in V (x) M <- V                L[a', w; a]
    W     in M (x) P <- P (x) M          W[w, s; s', w']
    R     in V <- V (x) M                R[c; c', w]
    X1 = L . theta                       (a', w ; s1, s2, c)
    X2 = X1 . W1  (w, s1 contracted)     (a', w1 ; s1', s2, c)
    X3 = X2 . W2  (w1, s2 contracted)    (a', w2 ; s1', s2', c)
    H theta = X3 . R (w2, c contracted)  (a' ; s1', s2', c')
"""
import math

import numpy as np

from .tensor import TensorMap
from .workloads import mps_env_spaces


class TwoSiteHeff:
    """Environments of one bond dimension, resident in HBM."""

    def __init__(self, rt, rule_name="U1", D=512, seed=7, precision="f64", mpo="random"):
        """Environments L, R of bond dimension D (seeded normal(0,1)); MPO tensors
        W1 = W2 either seeded normal(0,1) ("random") or the spin-1/2 Heisenberg
        MPO ("heisenberg", U(1) only)."""
        self.rt = rt
        self.rule, V, P, M = mps_env_spaces(rule_name, D)
        self.V, self.P, self.M = V, P, M
        self.D = D
        self.precision = precision
        self.L = TensorMap.rand(rt, [V, M], [V], seed=seed, normal=True)
        self.R = TensorMap.rand(rt, [V], [V, M], seed=seed + 1, normal=True)
        if mpo == "heisenberg":
            if rule_name != "U1":
                raise ValueError("the Heisenberg MPO is built in U(1) form only")
            from .workloads import heisenberg_mpo
            self.W1 = self.W2 = heisenberg_mpo(rt)
        else:
            self.W1 = TensorMap.rand(rt, [M, P], [P, M], seed=seed + 2, normal=True)
            self.W2 = TensorMap.rand(rt, [M, P], [P, M], seed=seed + 3, normal=True)
        self.theta_space = TensorMap.zeros(rt, [V], [P, P, V]).space

    @classmethod
    def from_tensors(cls, rt, L, W1, W2, R, precision="f64"):
        """H_eff of given environments L in Vl (x) M <- Vl, R in Vr <- Vr (x) M and
        MPO tensors W in M (x) P <- P (x) M; theta lives in Vl <- P (x) P (x) Vr."""
        h = object.__new__(cls)
        h.rt, h.precision = rt, precision
        h.L, h.W1, h.W2, h.R = L, W1, W2, R
        P = W1.space.codomain[1]
        h.theta_space = TensorMap.zeros(rt, [L.space.codomain[0]], [P, P, R.space.codomain[0]]).space
        return h

    def gram(self, theta, scratch=None):
        """G theta = L^T (L theta): a real symmetric positive semidefinite operator on
        the theta space built from the same block-sparse contractions (lazy adjoint)."""
        sc = scratch if scratch is not None else {}
        x = self.L.contract([2], theta, [0], [0, 1, 2, 3, 4], 2, out=sc.get("g1"), precision=self.precision)
        sc["g1"] = x
        y = self.L.adjoint().contract([1, 2], x, [0, 1], [0, 1, 2, 3], 1, out=sc.get("g2"),
                                      precision=self.precision)
        sc["g2"] = y
        return y

    def on(self, rt):
        """Views of the environments for another runtime (same device, no copy)."""
        v = object.__new__(TwoSiteHeff)
        v.__dict__.update(self.__dict__)
        v.rt = rt
        for name in ("L", "R", "W1", "W2"):
            t = getattr(self, name)
            setattr(v, name, TensorMap(rt, t.space, _buf=t._buf, _base=t))
        return v

    # The chain (lhs, rhs, lhs axes, rhs axes, output permutation, output
    # codomain rank, result) of H theta; names index the environments, theta
    # and the intermediates.
    CHAIN = (("L", "theta", [2], [0], [0, 1, 2, 3, 4], 2, "x1"),
             ("x1", "W1", [1, 2], [3, 1], [0, 3, 4, 1, 2], 2, "x2"),
             ("x2", "W2", [1, 3], [3, 1], [0, 3, 1, 4, 2], 2, "x3"),
             ("x3", "R", [1, 4], [2, 0], [0, 1, 2, 3], 1, "y"))

    def apply(self, theta, scratch=None, precision=None):
        """H_eff theta (four contractions); scratch: dict of reusable intermediates."""
        prec = precision or self.precision
        sc = scratch if scratch is not None else {}
        env = {"L": self.L, "W1": self.W1, "W2": self.W2, "R": self.R, "theta": theta}
        for lhs, rhs, ax, bx, perm, nout, res in self.CHAIN:
            env[res] = env[lhs].contract(ax, env[rhs], bx, perm, nout, out=sc.get(res), precision=prec)
            sc[res] = env[res]
        if not hasattr(self, "_flops"):
            self._flops = sum(self.rt.contraction(env[l].space, env[r].space, ax, bx, perm, nout).flops
                              for l, r, ax, bx, perm, nout, _ in self.CHAIN)
        return env["y"]

    def launch_list(self, theta, scratch):
        """(artifact, dst, lhs, rhs) device handles of the four tnt_contract calls
        of apply(theta, scratch), for callers that replay the chain on fixed
        buffers (apply must have run once with this theta and scratch)."""
        env = {"L": self.L, "W1": self.W1, "W2": self.W2, "R": self.R, "theta": theta}
        env.update(scratch)
        return [(self.rt.contraction(env[l].space, env[r].space, ax, bx, perm, nout).art, env[res]._buf,
                 env[l]._buf, env[r]._buf) for l, r, ax, bx, perm, nout, res in self.CHAIN]


def lanczos(op, v0, iters=30, tol=1e-10, stats=None):
    """Lowest eigenvalue of a real symmetric operator by Lanczos with full
    reorthogonalisation (two modified Gram-Schmidt passes); vectors, FP64
    inner products and the normalisations stay on the device.

    Each iteration queues the matvec, alpha = q.w, the 2 (k+1) projections
    (each dot into a device scalar, the subtraction reading it there), |w|^2
    and 1/|w|, and reads alpha and |w|^2 back once for the host's tridiagonal
    eigensolve; the normalisation of the next basis vector is queued after
    that decision.  op: callable TensorMap -> TensorMap (same space); v0:
    start vector.  Returns (eigenvalue, ritz residual norm, iterations);
    ``stats`` (a dict, optional) receives the host synchronisations taken.
    """
    from .krylov import DeviceScalars
    rt = v0.rt
    S = DeviceScalars(rt, n=2 * iters + 8)
    A0, B0, C0, C1, INV = 0, iters, 2 * iters, 2 * iters + 1, 2 * iters + 2
    try:
        q = v0.copy()
        S.dot(q, q, C0)
        S.program((3, INV, C0, C0, 0.0))            # 1 / |v0|
        S.axpby(q, q, ai=INV, bf=0.0)              # q = v0 / |v0|
        Q = [q]
        alphas, betas = [], []
        theta_min, res = None, None
        for k in range(iters):
            w = op(Q[-1]).copy()
            S.dot(Q[-1], w, A0 + k)
            for _ in range(2):
                for qi in Q:
                    S.dot(qi, w, C1)
                    S.axpby(w, qi, ai=C1, af=-1.0)  # w -= (q_i . w) q_i
            S.dot(w, w, B0 + k)
            S.program((3, INV, B0 + k, B0 + k, 0.0))
            s = S.read()
            alphas.append(float(s[A0 + k]))
            b = math.sqrt(max(float(s[B0 + k]), 0.0))
            T = np.diag(alphas) + np.diag(betas, 1) + np.diag(betas, -1)
            evals, evecs = np.linalg.eigh(T)
            theta_min = float(evals[0])
            res = abs(b * evecs[-1, 0])
            if res < tol or b < 1e-300:
                return theta_min, res, k + 1
            betas.append(b)
            S.axpby(w, w, ai=INV, bf=0.0)          # w / |w|
            Q.append(w)
        return theta_min, res, iters
    finally:
        if stats is not None:
            stats["host_syncs"] = S.syncs
        S.close()
