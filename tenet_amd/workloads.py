"""Synthetic workloads of the block-sparse contraction path (inputs built on the device).

* ``mps_env``: one left-environment update of a two-site-MPS sweep,
      T1 = L . A,  T2 = T1 . W,  L' = A^dagger . T2
  (T2 stays in the core GEMM's output order, so only one recoupling pass
  sits between each pair of GEMMs)
  with bond V (dimension D), physical P (d=2) and MPO bond M (5).
  U(1): V charges -10..10 with Gaussian degeneracies (sigma 4, largest
  remainder), P = {0, 1}, M = {0:3, +1:1, -1:1} (Heisenberg S^z / S^+- channels).
  SU(2): V spins with multiplicities from the same Gaussian rule so that
  sum_j m_j (2j+1) = D (or, counting multiplets, sum_j m_j = D), P = spin 1/2,
  M = {0, 1, 0} (dimension 5).
* ``su2_fmove`` (spins 2j = 0..8, multiplicity floor(64 / (2j+1))):
  - swap form: A in V (x) V (x) V <- V; A' = permute(A, (0, 2, 1), (3,))
    braids codomain legs 1 and 2, which recouples the inner line j01 (Multi
    groups with 6j coefficients); then C = contract(A', [2, 3], B, [2, 0])
    with B in V (x) V <- V (x) V.
  - cross form: A, B in V (x) V <- V (x) V, C = contract(A, [1, 3], B, [2, 0])
    -- both operand transforms cross the codomain/domain boundary, so inner
    lines recouple before the core GEMM.
  - the plain 2 <- 2 codomain swap permute(A, (1, 0), (2, 3)) is an R-move
    only (Single groups, coefficient +-1), timed on its own.

Entries are seeded on the device: normal(0,1) (Box-Muller on splitmix64) for
the environment, the reference's uniform [-1, 1) fill for the F-move case.
"""
import math

import numpy as np

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


def su2_bond(D=4096, twoj_max=20, sigma=3.0, count="dimension"):
    """Spins 0..twoj_max/2 (integer and half-integer); m_j Gaussian in j.

    count="dimension": D is the dimension of the leg, sum_j m_j (2j+1) = D;
    count="multiplets": D counts multiplets, sum_j m_j = D (the bond dimension
    SU(2) DMRG codes quote; the leg's dimension is then several times D)."""
    labels = list(range(0, twoj_max + 1))  # doubled spins
    if count == "multiplets":
        deg = gaussian_degeneracies(D, [l / 2.0 for l in labels], sigma)
    elif count == "dimension":
        deg = gaussian_degeneracies(D, [l / 2.0 for l in labels], sigma, weight=lambda j: int(2 * j + 1))
    else:
        raise ValueError(count)
    return {l: deg[l / 2.0] for l in labels}


def mps_env_spaces(rule_name="U1", D=4096, su2_count="dimension"):
    rule = Rule(rule_name)
    if rule_name == "U1":
        V = GradedSpace(rule, sorted(u1_bond(D).items()))
        P = GradedSpace(rule, [(0, 1), (1, 1)])
        M = GradedSpace(rule, [(0, 3), (1, 1), (-1, 1)])
    elif rule_name == "SU2":
        V = GradedSpace(rule, sorted(su2_bond(D, count=su2_count).items()))
        P = GradedSpace(rule, [(1, 1)])
        M = GradedSpace(rule, [(0, 2), (2, 1)])
    else:
        raise ValueError(rule_name)
    return rule, V, P, M


def heisenberg_mpo_payload(J=1.0):
    """Spin-1/2 Heisenberg MPO in U(1) form: (HomSpace, host payload) of
    W in M (x) P <- P (x) M with legs W[w_r, s_in; s_out, w_l] =
    <s_out| O(w_l, w_r) |s_in> and the lower-triangular operator matrix
        O(0,0) = I,  O(1,0) = S+,  O(2,0) = S-,  O(3,0) = Sz,
        O(4,1) = J/2 S-,  O(4,2) = J/2 S+,  O(4,3) = J Sz,  O(4,4) = I,
    so that left channel 4 and right channel 0 close the chain to
    sum over bonds of J S_i . S_(i+1).  Channels on M = {0: 3, -1: 1, +1: 1}:
    charge 0 holds channels 0, 3, 4 (in degeneracy order), charge -1 channel 1
    (it carries S+), charge +1 channel 2; P = {0: down, 1: up}."""
    rule, _, P, M = mps_env_spaces("U1", 64)
    sp, sm = np.zeros((2, 2)), np.zeros((2, 2))
    sp[1, 0] = 1.0  # <up| S+ |down>
    sm[0, 1] = 1.0
    sz, eye = np.diag([-0.5, 0.5]), np.eye(2)
    ops = {(0, 0): eye, (1, 0): sp, (2, 0): sm, (3, 0): sz, (4, 1): 0.5 * J * sm, (4, 2): 0.5 * J * sp,
           (4, 3): J * sz, (4, 4): eye}
    chan = {0: (0, 3, 4), -1: (1,), 1: (2,)}
    space = TensorMap._space([M, P], [P, M])
    data = np.zeros(space.len)
    for off, shape, strides, _, unc in space.blocks():
        qr, s_in, s_out, ql = (rule.label(x) for x in unc)
        for idx in np.ndindex(*shape):
            op = ops.get((chan[ql][idx[3]], chan[qr][idx[0]]))
            if op is not None:
                data[off + sum(i * st for i, st in zip(idx, strides))] = op[s_out, s_in]
    return space, data


def heisenberg_mpo(rt, J=1.0):
    """The U(1) spin-1/2 Heisenberg MPO tensor resident on the device."""
    space, data = heisenberg_mpo_payload(J)
    return TensorMap.from_numpy(rt, space, data)


class GraphStep:
    """Replays a workload step from a hipGraph: the step's launches are
    recorded once (every table and buffer already exists after one eager run)
    and each replay is one graph launch, without per-kernel launch gaps."""

    def capture(self):
        import ctypes as C
        from . import _native as N
        self.run()
        self.rt.synchronize()
        h = N.hip()
        self.rt._check(h.tnt_graph_begin(self.rt._ctx))
        try:
            self.run()
        finally:
            g = C.c_void_p()
            self.rt._check(h.tnt_graph_end(self.rt._ctx, C.byref(g)))
        self.graph = g
        return self

    def replay(self):
        from . import _native as N
        self.rt._check(N.hip().tnt_graph_launch(self.rt._ctx, self.graph))
        return self.out


class MPSEnvUpdate(GraphStep):
    """Left-environment update L' = A^dagger . ((L . A) . W) on one runtime."""

    def __init__(self, rt, rule_name="U1", D=4096, precision="f64", seed=3, su2_count="dimension"):
        self.rt = rt
        self.rule, V, P, M = mps_env_spaces(rule_name, D, su2_count)
        self.V, self.P, self.M = V, P, M
        self.precision = precision
        self.L = TensorMap.rand(rt, [V, M], [V], seed=seed, normal=True)
        self.A = TensorMap.rand(rt, [V], [P, V], seed=seed + 1, normal=True)
        self.W = TensorMap.rand(rt, [M, P], [P, M], seed=seed + 2, normal=True)
        self.out = None
        self._tmp = {}
        self.run()  # compile every structure (cold) and size the outputs
        c1 = rt.contraction(self.L.space, self.A.space, [2], [0], [0, 1, 2, 3], 2)
        c2 = rt.contraction(self._tmp["T1"].space, self.W.space, *self.T2_AXES)
        c3 = rt.contraction(self.A.space, self._tmp["T2"].space, *self.LP_AXES, True, False)
        self.contractions = (c1, c2, c3)
        self.flops = sum(c.flops for c in self.contractions)

    # T2 is produced as T2[a', s'; b, w]: the codomain is exactly the pair the
    # last contraction sums over, so that contraction reads T2 without a
    # recoupling pass, and the MPO step (narrow: k = n = dim(w) dim(s) <= 10)
    # runs as one fused pass from T1 to T2.  L' is the same tensor, in the same
    # layout, as with T2[a', w; s', b].
    T2_AXES = ([1, 2], [3, 1], [0, 3, 1, 2], 2)
    LP_AXES = ([2, 0], [0, 1], [0, 2, 1], 2)

    def run(self, L=None):
        L = L or self.L
        t1 = L.contract([2], self.A, [0], [0, 1, 2, 3], 2, out=self._tmp.get("T1"),
                        precision=self.precision)
        self._tmp["T1"] = t1
        t2 = t1.contract(self.T2_AXES[0], self.W, *self.T2_AXES[1:], out=self._tmp.get("T2"),
                         precision=self.precision)
        self._tmp["T2"] = t2
        self.out = self.A.adjoint().contract(self.LP_AXES[0], t2, *self.LP_AXES[1:], out=self.out,
                                             precision=self.precision)
        return self.out


def su2_fmove_space(mult=64, twoj_max=8):
    rule = Rule("SU2")
    V = GradedSpace(rule, [(j, max(1, mult // (j + 1))) for j in range(0, twoj_max + 1)])
    return rule, V


class SU2FMoveContract(GraphStep):
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


class SU2FMoveSwapContract(GraphStep):
    """A' = permute(A, (0, 2, 1), (3,)) then C = contract(A', [2, 3], B, [2, 0])."""

    SWAP = ((0, 2, 1), (3,))
    AXES = ([2, 3], [2, 0], [0, 1, 2, 3], 2)

    def __init__(self, rt, mult=64, twoj_max=8, precision="f64"):
        self.rt = rt
        self.rule, V = su2_fmove_space(mult, twoj_max)
        self.V = V
        self.precision = precision
        self.A = TensorMap.rand(rt, [V, V, V], [V], seed=1)
        self.B = TensorMap.rand(rt, [V, V], [V, V], seed=2)
        self.Ap = None
        self.out = None
        self.run()
        self.transform = rt.transform(self.A.space, *self.SWAP)
        self.contraction = rt.contraction(self.Ap.space, self.B.space, *self.AXES)
        self.flops = self.contraction.flops

    def run(self):
        self.Ap = self.A.permute(*self.SWAP, out=self.Ap)
        self.out = self.Ap.contract(self.AXES[0], self.B, *self.AXES[1:], out=self.out, precision=self.precision)
        return self.out


def u1z2_bond(rng, D, qmax=6):
    """U(1)xZ2 bond of dimension D: charges -qmax..qmax x parity, multinomial degeneracies."""
    labels = [(q, p) for q in range(-qmax, qmax + 1) for p in (0, 1)]
    w = np.exp(-np.array([q * q for q, _ in labels]) / (2.0 * (qmax / 2.0) ** 2))
    counts = rng.multinomial(max(0, D - len(labels)), w / w.sum()) + 1  # every sector present
    return [(l, int(c)) for l, c in zip(labels, counts)]


def batched_specs(count=256, dmin=64, dmax=1024, seed=4):
    """Bond legs (label, degeneracy) lists (V_l, V_m, V_r) of each merge, deterministic in seed."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        Dl, Dm, Dr = (int(x) for x in rng.integers(dmin, dmax + 1, size=3))
        out.append((u1z2_bond(rng, Dl), u1z2_bond(rng, Dm), u1z2_bond(rng, Dr)))
    return out


class BatchedMerge:
    """Many independent U(1)xZ2 two-site merges C_i = A_i . B_i in one grouped GEMM launch.

    A_i in V_l (x) P <- V_m, B_i in V_m <- P (x) V_r, P = {(0,0), (1,1)} (d = 2);
    bond dimensions uniform on [dmin, dmax] from seed 4.  All coupled-sector
    GEMMs of all contractions go to tnt_gemm_grouped as pointer jobs (one
    launch per tile shape), optionally captured in a hipGraph.
    """

    def __init__(self, rt, count=256, dmin=64, dmax=1024, seed=4, precision="f32"):
        import ctypes as C
        from . import _native as N
        self.rt = rt
        self.precision = precision
        rule = Rule("U1xZ2")
        P = GradedSpace(rule, [((0, 0), 1), ((1, 1), 1)])
        self.pairs = []
        jobs = []
        self.flops = 0.0
        h = N.hip()
        for i, (pl, pm, pr) in enumerate(batched_specs(count, dmin, dmax, seed)):
            Vl, Vm, Vr = GradedSpace(rule, pl), GradedSpace(rule, pm), GradedSpace(rule, pr)
            A = TensorMap.rand(rt, [Vl, P], [Vm], seed=1000 + 2 * i)
            B = TensorMap.rand(rt, [Vm], [P, Vr], seed=1001 + 2 * i)
            c = rt.contraction(A.space, B.space, [2], [0], [0, 1, 2, 3], 2)
            Cm = TensorMap.zeros(rt, c.dst_space)
            d = c.desc.contents
            ptr = {}
            for name, t in (("a", A), ("b", B), ("c", Cm)):
                p = C.c_void_p()
                h.tnt_buf_info(t._buf, C.byref(p), None, None, None)
                ptr[name] = p.value
            for k in range(d.n_jobs):
                j = d.jobs[k]
                jobs.append(N.GemmPtrJob(ptr["c"] + 8 * j.dst_offset, ptr["a"] + 8 * j.lhs_offset,
                                         ptr["b"] + 8 * j.rhs_offset, j.rows, j.contracted, j.cols,
                                         max(1, j.rows), max(1, j.rows), max(1, j.contracted), 0, 0))
            self.flops += c.flops
            self.pairs.append((A, B, Cm, c))
        self.jobs = (N.GemmPtrJob * len(jobs))(*jobs)
        self.n_jobs = len(jobs)
        # algorithmic bytes of one launch: every operand read once, every result written once
        self.bytes = float(sum(8 * (j.rows * j.contracted + j.contracted * j.cols + j.rows * j.cols) for j in jobs))
        self.graph = None

    def run(self, precision=None):
        from . import _native as N
        from .tensor import PRECISIONS
        prec = PRECISIONS[precision or self.precision]
        if self.graph is not None and precision is None:
            self.rt._check(N.hip().tnt_graph_launch(self.rt._ctx, self.graph))
            return
        self.rt._check(N.hip().tnt_gemm_grouped(self.rt._ctx, self.jobs, self.n_jobs, 1.0, 0.0, prec))

    def capture(self):
        """Record the grouped launch in a hipGraph (tables built by one eager run first)."""
        import ctypes as C
        from . import _native as N
        self.run()
        self.rt.synchronize()
        h = N.hip()
        self.rt._check(h.tnt_graph_begin(self.rt._ctx))
        self.run()
        g = C.c_void_p()
        self.rt._check(h.tnt_graph_end(self.rt._ctx, C.byref(g)))
        self.graph = g

    def results(self):
        return [Cm.to_numpy() for _, _, Cm, _ in self.pairs]
