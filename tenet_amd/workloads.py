"""Synthetic workloads of the block-sparse contraction path (inputs built on the device).

* ``mps_env``: one left-environment update of a two-site-MPS sweep,
      T1 = L . A,  T2 = T1 . W,  L' = A^dagger . T2
  with bond V (dimension D), physical P (d=2) and MPO bond M (5).
  U(1): V charges -10..10 with Gaussian degeneracies (sigma 4, largest
  remainder), P = {0, 1}, M = {0:3, +1:1, -1:1} (Heisenberg S^z / S^+- channels).
  SU(2): V spins with multiplicities from the same Gaussian rule so that
  sum_j m_j (2j+1) = D, P = spin 1/2, M = {0, 1, 0} (dimension 5).
* ``su2_fmove``: A, B in V (x) V <- V (x) V with spins 2j = 0..8 and
  multiplicity floor(64 / (2j+1)); C = contract(A, [1, 3], B, [2, 0]) -- both
  operand transforms cross the codomain/domain boundary, so inner fusion
  lines recouple (Multi groups with 6j coefficients) before the core GEMM.

Entries are seeded on the device: normal(0,1) (Box-Muller on splitmix64) for
the environment, the reference's uniform [-1, 1) fill for the F-move case.
"""
import math

from .tensor import GradedSpace, Rule, TensorMap


def gaussian_degeneracies(total, labels, sigma, weight=None):
    """Largest-remainder split of `total` over `labels` with Gaussian weights.

    weight(label) multiplies each share's "cost" (e.g. 2j+1 for SU(2)): the
    shares m_l satisfy sum_l m_l * weight(l) == total when possible.
    """
    weight = weight or (lambda l: 1)
    w = [math.exp(-(l ** 2) / (2.0 * sigma ** 2)) for l in labels]
    cost = [weight(l) for l in labels]
    s = sum(wi * ci for wi, ci in zip(w, cost))
    exact = [total * wi / s for wi in w]  # multiplicity (before cost)
    m = [max(1, int(math.floor(x))) for x in exact]
    used = sum(mi * ci for mi, ci in zip(m, cost))
    order = sorted(range(len(labels)), key=lambda i: (exact[i] - math.floor(exact[i])), reverse=True)
    k = 0
    while used < total and k < 10 * len(labels):
        i = order[k % len(order)]
        if used + cost[i] <= total:
            m[i] += 1
            used += cost[i]
        k += 1
    # fill any residue with spin-0 / charge-0 style unit-cost sectors
    if used != total:
        i0 = min(range(len(labels)), key=lambda i: (cost[i], abs(labels[i])))
        while used + cost[i0] <= total:
            m[i0] += 1
            used += cost[i0]
    return dict(zip(labels, m))


def u1_bond(D=4096, qmax=10, sigma=4.0):
    labels = list(range(-qmax, qmax + 1))
    return gaussian_degeneracies(D, labels, sigma)


def su2_bond(D=4096, twoj_max=20, sigma=3.0):
    """Spins 0..twoj_max/2 (integer and half-integer); m_j Gaussian in j; sum m_j(2j+1) = D."""
    labels = list(range(0, twoj_max + 1))  # doubled spins
    deg = gaussian_degeneracies(D, [l / 2.0 for l in labels], sigma, weight=lambda j: int(2 * j + 1))
    return {l: deg[l / 2.0] for l in labels}


def mps_env_spaces(rule_name="U1", D=4096):
    rule = Rule(rule_name)
    if rule_name == "U1":
        V = GradedSpace(rule, sorted(u1_bond(D).items()))
        P = GradedSpace(rule, [(0, 1), (1, 1)])
        M = GradedSpace(rule, [(0, 3), (1, 1), (-1, 1)])
    elif rule_name == "SU2":
        V = GradedSpace(rule, sorted(su2_bond(D).items()))
        P = GradedSpace(rule, [(1, 1)])
        M = GradedSpace(rule, [(0, 2), (2, 1)])
    else:
        raise ValueError(rule_name)
    return rule, V, P, M


class MPSEnvUpdate:
    """Left-environment update L' = A^dagger . ((L . A) . W) on one runtime."""

    def __init__(self, rt, rule_name="U1", D=4096, precision="f64", seed=3):
        self.rt = rt
        self.rule, V, P, M = mps_env_spaces(rule_name, D)
        self.V, self.P, self.M = V, P, M
        self.precision = precision
        self.L = TensorMap.rand(rt, [V, M], [V], seed=seed, normal=True)
        self.A = TensorMap.rand(rt, [V], [P, V], seed=seed + 1, normal=True)
        self.W = TensorMap.rand(rt, [M, P], [P, M], seed=seed + 2, normal=True)
        self.out = None
        self._tmp = {}
        self.run()  # compile every structure (cold) and size the outputs
        c1 = rt.contraction(self.L.space, self.A.space, [2], [0], [0, 1, 2, 3], 2)
        c2 = rt.contraction(self._tmp["T1"].space, self.W.space, [1, 2], [3, 1], [0, 2, 3, 1], 2)
        c3 = rt.contraction(self.A.space, self._tmp["T2"].space, [2, 0], [0, 2], [0, 1, 2], 2, True, False)
        self.contractions = (c1, c2, c3)
        self.flops = sum(c.flops for c in self.contractions)

    def run(self, L=None):
        L = L or self.L
        t1 = L.contract([2], self.A, [0], [0, 1, 2, 3], 2, out=self._tmp.get("T1"),
                        precision=self.precision)
        self._tmp["T1"] = t1
        t2 = t1.contract([1, 2], self.W, [3, 1], [0, 2, 3, 1], 2, out=self._tmp.get("T2"),
                         precision=self.precision)
        self._tmp["T2"] = t2
        self.out = self.A.adjoint().contract([2, 0], t2, [0, 2], [0, 1, 2], 2, out=self.out,
                                             precision=self.precision)
        return self.out


def su2_fmove_space(mult=64, twoj_max=8):
    rule = Rule("SU2")
    V = GradedSpace(rule, [(j, max(1, mult // (j + 1))) for j in range(0, twoj_max + 1)])
    return rule, V


class SU2FMoveContract:
    """C = contract(A, [1, 3], B, [2, 0]) with recoupling on both operands."""

    def __init__(self, rt, mult=64, twoj_max=8, precision="f64"):
        self.rt = rt
        self.rule, V = su2_fmove_space(mult, twoj_max)
        self.V = V
        self.precision = precision
        self.A = TensorMap.rand(rt, [V, V], [V, V], seed=1)
        self.B = TensorMap.rand(rt, [V, V], [V, V], seed=2)
        self.out = None
        self.run()
        self.contraction = rt.contraction(self.A.space, self.B.space, [1, 3], [2, 0], [0, 1, 2, 3], 2)
        self.flops = self.contraction.flops

    def run(self):
        self.out = self.A.contract([1, 3], self.B, [2, 0], [0, 1, 2, 3], 2, out=self.out,
                                   precision=self.precision)
        return self.out
