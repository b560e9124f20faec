"""GPU parity of the two-site H_eff chain, the device Lanczos solver, the
stream-ordered dot and the continuous-batching server.

Bars: H_eff in FP64 within 1e-13 relative Frobenius error of the oracle chain
(split-BF16 within 1e-10); Lanczos lowest eigenvalue of the Gram operator
within 1e-9 (relative) of the eigenvalue of the dense matrix that the oracle
operator builds column by column; serving results bit-identical alone and in
the mix.
"""
import ctypes as C

import numpy as np
import pytest

from descs import ospace
from oracle.tensor import OTensor
from tenet_amd import _native as N
from tenet_amd.heff import TwoSiteHeff, lanczos
from tenet_amd.serving import HeffServer, poisson_trace
from tenet_amd.tensor import Runtime, TensorMap

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rt():
    r = Runtime(0)
    yield r
    r.close()


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb else 1.0)


def _o(t):
    return OTensor(ospace(t.space), t.to_numpy())


def _oracle_apply(H, theta):
    L, W1, W2, R = map(_o, (H.L, H.W1, H.W2, H.R))
    x1 = L.contract([2], theta, [0], [0, 1, 2, 3, 4], 2)
    x2 = x1.contract([1, 2], W1, [3, 1], [0, 3, 4, 1, 2], 2)
    x3 = x2.contract([1, 3], W2, [3, 1], [0, 3, 1, 4, 2], 2)
    return x3.contract([1, 4], R, [2, 0], [0, 1, 2, 3], 1)


@pytest.mark.parametrize("rule_name,D", [("U1", 40), ("SU2", 24)])
@pytest.mark.parametrize("prec", ["f64", "f32", "split_bf16", "split_i8"])
def test_heff_apply_matches_oracle(rt, rule_name, D, prec):
    H = TwoSiteHeff(rt, rule_name, D, seed=11, precision=prec)
    th = TensorMap.rand(rt, H.theta_space, seed=3, normal=True)
    y = H.apply(th)
    assert y.space == H.theta_space
    want = _oracle_apply(H, _o(th)).data
    tol = {"f64": 1e-13, "f32": 5e-6, "split_bf16": 1e-10, "split_i8": 1e-10}[prec]
    assert rel(y.to_numpy(), want) < tol, (rule_name, prec, rel(y.to_numpy(), want))
    # reuse of the scratch intermediates gives the same bits
    sc = {}
    a = H.apply(th, sc).to_numpy().copy()
    b = H.apply(th, sc).to_numpy()
    assert np.array_equal(a, b)


def _dense_gram(H):
    """The Gram operator L^T L of H's environment as a dense matrix, column by
    column from the oracle's contractions."""
    osp = ospace(H.theta_space)
    n = H.theta_space.len
    L = _o(H.L)
    Lad = L.adjoint()
    cols = []
    for i in range(n):
        e = np.zeros(n)
        e[i] = 1.0
        x = L.contract([2], OTensor(osp, e), [0], [0, 1, 2, 3, 4], 2)
        cols.append(Lad.contract([1, 2], x, [0, 1], [0, 1, 2, 3], 1).data)
    return np.array(cols).T


@pytest.mark.parametrize("rule_name,D", [("U1", 24), ("SU2", 24)])
def test_lanczos_gram_matches_dense_oracle(rt, rule_name, D):
    H = TwoSiteHeff(rt, rule_name, D, seed=5)
    n = H.theta_space.len
    G = _dense_gram(H)
    want = float(np.min(np.linalg.eigvals(G).real))
    v0 = TensorMap.rand(rt, H.theta_space, seed=9, normal=True)
    sc = {}
    ev, res, iters = lanczos(lambda v: H.gram(v, sc), v0, iters=n, tol=1e-12)
    assert abs(ev - want) <= 1e-9 * max(1.0, abs(want)), (ev, want, res, iters)


@pytest.mark.parametrize("rule_name,D", [("U1", 24), ("SU2", 24)])
def test_cg_damped_gram_solve_matches_dense_oracle(rt, rule_name, D):
    """The reference's CG (cg.rs:58-190) on device vectors: (L^T L + 0.5 I) x = b
    with the device Gram operator and weighted FP64 inner products, against a
    dense solve of the oracle's matrix; the residual is re-checked with a fresh
    device matvec."""
    from tenet_amd.krylov import CgOptions, DeviceVector, cg
    H = TwoSiteHeff(rt, rule_name, D, seed=5)
    b = TensorMap.rand(rt, H.theta_space, seed=13, normal=True)
    sc = {}
    res = cg(lambda v: DeviceVector(H.gram(v.t, sc)), DeviceVector(b), CgOptions(rtol=1e-12, damping=0.5))
    assert res.stats.converged and res.stats.breakdown is None, res.stats
    assert res.stats.matvecs == res.stats.iterations <= H.theta_space.len
    x = res.solution.t
    r = H.gram(x, {}).add(x, 1.0, 0.5).add(b, 1.0, -1.0)
    assert r.norm() <= 1e-10 * b.norm(), (r.norm(), b.norm())
    G = _dense_gram(H)
    want = np.linalg.solve(G + 0.5 * np.eye(len(G)), b.to_numpy())
    got = x.to_numpy()
    assert np.linalg.norm(got - want) <= 1e-9 * np.linalg.norm(want)


def _identity_env(rt, V, M, left):
    """Boundary environment delta(a', a) delta(w, boundary channel): left closes
    on MPO channel 4 (third slot of charge 0), right on channel 0 (first slot)."""
    from tenet_amd.tensor import TensorMap as T
    lab = V.rule.label
    if left:  # L[a', w; a]
        return T.from_block_fn(rt, [V, M], [V], lambda u, c, i: float(
            lab(u[1]) == 0 and i[1] == 2 and u[0] == u[2] and i[0] == i[2]))
    return T.from_block_fn(rt, [V], [V, M], lambda u, c, i: float(  # R[c; c', w]
        lab(u[2]) == 0 and i[2] == 0 and u[0] == u[1] and i[0] == i[1]))


def test_lanczos_heisenberg_bond_ground_energy(rt):
    """With identity boundary environments the two-site H_eff of the Heisenberg
    MPO is 1 (x) J S1.S2 (x) 1: symmetric, and its lowest eigenvalue is the
    singlet energy -3/4 J whenever a sector holds one up spin."""
    from tenet_amd.tensor import GradedSpace, Rule
    from tenet_amd.workloads import heisenberg_mpo, mps_env_spaces
    _, _, P, M = mps_env_spaces("U1", 64)
    Vl = GradedSpace(Rule("U1"), [(0, 2), (1, 3)])
    Vr = GradedSpace(Rule("U1"), [(-1, 2), (0, 1)])
    W = heisenberg_mpo(rt, J=1.0)
    H = TwoSiteHeff.from_tensors(rt, _identity_env(rt, Vl, M, True), W, W, _identity_env(rt, Vr, M, False))
    x = TensorMap.rand(rt, H.theta_space, seed=1, normal=True)
    y = TensorMap.rand(rt, H.theta_space, seed=2, normal=True)
    assert rel(H.apply(x).to_numpy(), _oracle_apply(H, _o(x)).data) < 1e-13
    hx, hy = H.apply(x).copy(), H.apply(y).copy()
    assert abs(x.inner(hy) - hx.inner(y)) < 1e-12 * abs(x.inner(hy))
    sc = {}
    ev, res, iters = lanczos(lambda v: H.apply(v, sc), x, iters=H.theta_space.len, tol=1e-12)
    assert abs(ev + 0.75) < 1e-10, (ev, res, iters)


def test_dot_async_matches_sync_and_events(rt):
    H = TwoSiteHeff(rt, "U1", 64, seed=2)
    a = TensorMap.rand(rt, H.theta_space, seed=1, normal=True)
    b = TensorMap.rand(rt, H.theta_space, seed=2, normal=True)
    want = a.inner(b)
    want_o = _o(a).inner(_o(b))
    assert abs(want - want_o) <= 1e-13 * abs(want_o)
    out = TensorMap.zeros(rt, HeffServer._vector_space(4))
    secs = a.space.sectors
    regs = (N.Region * len(secs))(*[N.Region(o, r * k) for _, o, r, k, _ in secs])
    ws = (C.c_double * len(secs))(*[d for *_, d in secs])
    h = N.hip()
    for slot in (0, 3, 1):
        rt._check(h.tnt_weighted_dot_async(rt._ctx, a._buf, b._buf, regs, ws, len(secs), out._buf, slot))
    ev = C.c_void_p()
    rt._check(h.tnt_event_create(rt._ctx, C.byref(ev)))
    rt._check(h.tnt_event_record(rt._ctx, ev))
    rt.synchronize()
    assert h.tnt_event_query(rt._ctx, ev) == 1
    rt._check(h.tnt_event_destroy(rt._ctx, ev))
    got = out.to_numpy()
    assert got[0] == want and got[3] == want and got[1] == want and got[2] == 0.0
    # out-of-range slot is rejected before anything is queued
    assert h.tnt_weighted_dot_async(rt._ctx, a._buf, b._buf, regs, ws, len(secs), out._buf, 4) == N.E_OOB


@pytest.mark.parametrize("rule,precision", [("U1", "split_bf16"), ("U1", "split_i8"), ("SU2", "split_i8"),
                                            ("SU2", "f64")])
def test_serving_same_alone_and_in_mix(rule, precision):
    """Requests arriving together are batched (stacked chains of up to 4 copies)
    and overlap on 3 streams; each result equals the request served alone
    (batch of one) bit for bit."""
    buckets = [32, 64, 96]
    srv = HeffServer(rule, buckets=buckets, n_streams=3, max_inflight=12, precision=precision, max_batch=4)
    try:
        srv.warm()
        trace = poisson_trace(18, rate=1e6, median_d=64, sigma=0.6, dmin=32, dmax=96, seed=5, buckets=buckets)
        stats = srv.serve(trace, keep_results=True)
        mix = dict(srv.results)
        assert stats["requests"] == 18 and stats["p99_ms"] >= stats["p50_ms"] > 0
        assert max(stats["batches"]) > 1, stats["batches"]  # batches formed
        for i, (_, D, seed) in enumerate(trace):
            srv.serve([(0.0, D, seed)], keep_results=True)
            alone = srv.results
            assert alone[0] == mix[i], (i, D)
            assert np.array_equal(alone[(0, "y")], mix[(i, "y")]), (i, D)
        # and the numerator matches an FP64 host evaluation of the same matvec
        _, D, seed = trace[0]
        v = srv._view(D, 0)
        th = TensorMap.rand(srv.rts[0], v.theta_space, seed=seed, normal=True)
        want = _o(th).inner(OTensor(ospace(v.theta_space), mix[(0, "y")]))
        assert abs(mix[0] - want) <= 1e-12 * max(1.0, abs(want))
        # the matvec itself against the oracle chain (split GEMMs within 1e-10)
        y_want = _oracle_apply(v, _o(th)).data
        assert rel(mix[(0, "y")], y_want) < (1e-13 if precision == "f64" else 1e-10)
    finally:
        srv.close()


def test_serving_hbm_gate_serialises_and_keeps_results():
    """With an HBM budget below two requests' buffers the server admits one
    request at a time (never more than the budget in flight, batching off by
    necessity); the results equal those of an unconstrained run."""
    buckets = [64]
    trace = poisson_trace(8, rate=1e6, median_d=64, sigma=0.1, dmin=64, dmax=64, seed=3, buckets=buckets)
    res = {}
    for gated in (False, True):
        srv = HeffServer("U1", buckets=buckets, n_streams=2, max_inflight=8, precision="split_i8", max_batch=4)
        try:
            srv.warm()
            if gated:
                srv.budget = int(1.5 * srv.fp[64])
            st = srv.serve(trace, keep_results=True)
            res[gated] = (dict(srv.results), st, srv.fp[64])
        finally:
            srv.close()
    (free, st_free, fp), (gate, st_gate, _) = res[False], res[True]
    assert st_gate["max_inflight"] == 1 and st_gate["batches"] == {1: 8}
    assert st_gate["max_footprint_bytes"] == fp  # one request's buffers at a time
    assert st_free["max_inflight"] > 1 and st_free["max_footprint_bytes"] > 1.5 * fp
    for i in range(len(trace)):
        assert gate[i] == free[i], i
        assert np.array_equal(gate[(i, "y")], free[(i, "y")]), i


def test_serving_graph_issue_matches_eager_issue():
    """Requests issued as hipGraph replays (four contractions captured per
    buffer set) give bit-for-bit the eager results; each graph is replayed
    several times with eager fills, dots and copies between the replays."""
    buckets = [32, 64]
    trace = poisson_trace(12, rate=1e6, median_d=48, sigma=0.5, dmin=32, dmax=64, seed=9, buckets=buckets)
    res = {}
    for graphs in (False, True):
        srv = HeffServer("U1", buckets=buckets, n_streams=2, max_inflight=4, precision="split_bf16",
                         graphs=graphs)
        try:
            srv.warm()
            if graphs:
                assert all(bs.graph is not None for sets in srv._free.values() for bs in sets)
            srv.serve(trace, keep_results=True)
            res[graphs] = dict(srv.results)
        finally:
            srv.close()
    for i in range(len(trace)):
        assert res[True][i] == res[False][i], i
        assert np.array_equal(res[True][(i, "y")], res[False][(i, "y")]), i


def _cg_cases(H, sc):
    from tenet_amd.krylov import CgOptions, DeviceVector
    pos = lambda v: DeviceVector(H.gram(v.t, sc))  # noqa: E731
    neg = lambda v: DeviceVector(H.gram(v.t, sc).copy().scale_(-1.0))  # noqa: E731
    return [("converge", pos, CgOptions(rtol=1e-12, damping=0.5)),
            ("max_iter", pos, CgOptions(rtol=1e-14, max_iter=3)),
            ("non_positive_curvature", neg, CgOptions()),
            ("unchecked_negative", neg, CgOptions(check_curvature=False, max_iter=4)),
            ("invalid_options", pos, CgOptions(rtol=-1.0))]


@pytest.mark.parametrize("rule_name,D", [("U1", 24), ("SU2", 24)])
def test_cg_device_loop_equals_host_synchronous_loop(rt, rule_name, D):
    """The stream-ordered CG (scalars on the device, one read per iteration)
    takes the same decisions on the same FP64 values as the loop with a host
    value per dot: identical statistics and a bit-identical solution, in
    convergence, iteration-limit, curvature-breakdown and invalid-option cases."""
    from tenet_amd.krylov import DeviceVector, cg_device, cg_generic
    H = TwoSiteHeff(rt, rule_name, D, seed=5)
    b = TensorMap.rand(rt, H.theta_space, seed=13, normal=True)
    for name, op, opt in _cg_cases(H, {}):
        dev = cg_device(op, DeviceVector(b), opt)
        host = cg_generic(op, DeviceVector(b), opt)
        for f in ("iterations", "matvecs", "converged", "initial_residual", "final_residual", "tolerance",
                  "breakdown"):
            assert getattr(dev.stats, f) == getattr(host.stats, f), (name, f, dev.stats, host.stats)
        assert np.array_equal(dev.solution.t.to_numpy(), host.solution.t.to_numpy()), name
        assert dev.stats.host_syncs == 1 + dev.stats.iterations + (
            1 if dev.stats.breakdown is not None and dev.stats.breakdown.value.endswith("Curvature") else 0), name
    assert cg_device(_cg_cases(H, {})[2][1], DeviceVector(b)).solution.t.norm() == 0.0  # rejected step


def test_lanczos_one_host_sync_per_iteration(rt):
    H = TwoSiteHeff(rt, "U1", 24, seed=5)
    v0 = TensorMap.rand(rt, H.theta_space, seed=9, normal=True)
    st = {}
    ev, res, iters = lanczos(lambda v: H.gram(v, {}), v0, iters=12, tol=0.0, stats=st)
    assert iters == 12 and st["host_syncs"] == 12
