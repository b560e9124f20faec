"""GPU parity of the benchmark workloads at small sizes against the oracle.

The environment update is checked against the oracle's chain written with the
textbook index order (T2[a', w; s', b]); the workload keeps T2 in the core
GEMM's output order instead, which must not change L'.
Bars: FP64 within 1e-13 relative Frobenius error; split-BF16 and split-INT8
within 1e-10.
"""
import numpy as np
import pytest

from descs import ospace
from oracle.tensor import OTensor
from tenet_amd.tensor import Runtime
from tenet_amd.workloads import MPSEnvUpdate, SU2FMoveContract, SU2FMoveSwapContract

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rt():
    r = Runtime(0)
    yield r
    r.close()


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _o(t):
    return OTensor(ospace(t.space), t.to_numpy())


@pytest.mark.parametrize("rule_name,D", [("U1", 128), ("SU2", 256)])
@pytest.mark.parametrize("prec", ["f64", "split_bf16", "split_i8"])
def test_env_update_matches_oracle(rt, rule_name, D, prec):
    wl = MPSEnvUpdate(rt, rule_name, D, prec)
    got = wl.run().to_numpy()
    assert wl.out.space == wl.L.space
    L, A, W = _o(wl.L), _o(wl.A), _o(wl.W)
    t1 = L.contract([2], A, [0], [0, 1, 2, 3], 2)
    t2 = t1.contract([1, 2], W, [3, 1], [0, 2, 3, 1], 2)
    want = A.adjoint().contract([2, 0], t2, [0, 2], [0, 1, 2], 2).data
    tol = {"f64": 1e-13, "split_bf16": 1e-10, "split_i8": 1e-10}[prec]
    assert rel(got, want) < tol, (rule_name, prec, rel(got, want))


def test_fmove_contract_matches_oracle(rt):
    wl = SU2FMoveContract(rt, mult=6, twoj_max=4, precision="f64")
    got = wl.run().to_numpy()
    want = _o(wl.A).contract([1, 3], _o(wl.B), [2, 0], [0, 1, 2, 3], 2).data
    assert rel(got, want) < 1e-13


@pytest.mark.parametrize("prec", ["f64", "f32", "split_bf16", "split_i8"])
def test_fmove_swap_contract_matches_oracle(rt, prec):
    """The 3 <- 1 codomain swap (inner line recoupled: Multi groups) then the
    contraction, against the oracle's permute and contract of the same
    payloads."""
    wl = SU2FMoveSwapContract(rt, mult=6, twoj_max=4, precision=prec)
    assert wl.transform.n_multi > 0
    got = wl.run().to_numpy()
    ap = _o(wl.A).permute([0, 2, 1], [3])
    assert np.abs(wl.Ap.to_numpy() - ap.data).max() <= 1e-14 * np.abs(ap.data).max()
    want = ap.contract([2, 3], _o(wl.B), [2, 0], [0, 1, 2, 3], 2).data
    tol = {"f64": 1e-13, "f32": 1e-5, "split_bf16": 1e-10, "split_i8": 1e-10}[prec]
    assert rel(got, want) < tol, (prec, rel(got, want))


def test_r_only_codomain_swap_matches_oracle(rt):
    """A 2 <- 2 codomain swap is an R-move only: Single groups whose
    coefficient is +-1.  Like the reference (fusion_tree.rs:3318-3372), the
    plan bends every leg into the codomain, braids, and bends back, so the
    coefficient is +-1 only to within rounding of the bend factors; the
    device result matches the oracle to one rounding per element."""
    from tenet_amd.tensor import TensorMap
    from tenet_amd.workloads import su2_fmove_space
    _, V = su2_fmove_space(6, 4)
    X = TensorMap.rand(rt, [V, V], [V, V], seed=7)
    tr = rt.transform(X.space, (1, 0), (2, 3))
    assert tr.n_multi == 0 and tr.n_single > 0
    d = tr.desc.contents
    co = np.array([d.coeffs[i] for i in range(d.n_coeffs)])
    assert np.abs(np.abs(co) - 1.0).max() <= 4 * np.finfo(float).eps
    got = X.permute((1, 0), (2, 3)).to_numpy()
    want = _o(X).permute([1, 0], [2, 3]).data
    assert np.all(np.abs(got - want) <= 8 * np.finfo(float).eps * np.abs(want))


@pytest.mark.parametrize("rule_name,D", [("U1", 128), ("SU2", 256)])
def test_snapshot_roundtrip_of_device_environment(rt, rule_name, D):
    """Checkpoint / resume: an environment written from the device and read
    back on another context is bit-identical, also as a lazy adjoint."""
    wl = MPSEnvUpdate(rt, rule_name, D, "f64")
    env = wl.run()
    data = env.to_bytes()
    rt2 = Runtime(0)
    back = adj = None
    try:
        back = type(env).from_bytes(rt2, data)
        assert back.space == env.space
        assert np.array_equal(back.to_numpy().view(np.uint64), env.to_numpy().view(np.uint64))
        assert back.to_bytes() == data
        adj = type(env).from_bytes(rt2, env.adjoint().to_bytes())
        assert adj._adjoint and np.array_equal(adj._base.to_numpy(), env.to_numpy())
    finally:
        del back, adj
        rt2.close()


def _env_network():
    from tenet_amd.network import Network
    # L[p, w; k] A[k; s, b] W[v, s; t, w] conj(A)[p; t, y]  ->  L'[y, v; b]
    return Network([["p", "w", "k"], ["k", "s", "b"], ["v", "s", "t", "w"], ["p", "t", "y"]],
                   [False, False, False, True], [2, 1, 2, 1], ["y", "v", "b"], 2)


@pytest.mark.parametrize("rule_name,D", [("U1", 128), ("SU2", 256)])
def test_network_env_update_matches_workload_and_oracle(rt, rule_name, D):
    from tenet_amd.network import LabelOrderOptimizer
    wl = MPSEnvUpdate(rt, rule_name, D, "f64")
    ref = wl.run().to_numpy()
    tensors = [wl.L, wl.A, wl.W, wl.A]
    net = _env_network()
    explicit = net.plan_with(tensors, [(0, 1), (4, 2), (3, 5)])
    got = explicit.execute(tensors)
    assert got.space == wl.out.space
    assert rel(got.to_numpy(), ref) < 1e-13
    L, A, W = _o(wl.L), _o(wl.A), _o(wl.W)
    t1 = L.contract([2], A, [0], [0, 1, 2, 3], 2)
    t2 = t1.contract([1, 2], W, [3, 1], [0, 2, 3, 1], 2)
    want = A.adjoint().contract([2, 0], t2, [0, 2], [0, 1, 2], 2).data
    for planned in (net.plan(tensors), net.plan(tensors, LabelOrderOptimizer(["w", "k", "s", "t", "p"]))):
        out = planned.execute(tensors)
        assert out.space == wl.out.space
        assert rel(out.to_numpy(), want) < 1e-13


@pytest.mark.parametrize("rule_name", ["U1", "SU2"])
def test_network_output_permutation_and_split(rt, rule_name):
    from tenet_amd.network import Network
    from descs import spaces
    leg = [(-1, 3), (0, 4), (1, 2)] if rule_name == "U1" else [(0, 3), (1, 2), (2, 2)]
    ps, _ = spaces(rule_name, [(leg, False), (leg, False)], [(leg, False), (leg, False)])
    from tenet_amd.tensor import TensorMap
    A = TensorMap.rand(rt, ps, seed=1)
    B = TensorMap.rand(rt, ps, seed=2)
    # C[e, b; a, f] = A[a, b; c, d] B[c, d; e, f]  (open labels reordered and re-split)
    net = Network([["a", "b", "c", "d"], ["c", "d", "e", "f"]], [False, False], [2, 2], ["e", "b", "a", "f"], 2)
    got = net.plan([A, B]).execute([A, B]).to_numpy()
    want = _o(A).contract([2, 3], _o(B), [0, 1], [2, 1, 0, 3], 2).data
    assert rel(got, want) < 1e-13


@pytest.mark.parametrize("prec", ["f64", "split_bf16"])
def test_fused_mpo_step_equals_general_route(rt, prec):
    """The MPO step of the U(1) environment update runs as one fused pass; the
    general route (TENET_AMD_FUSE_NARROW=0 at upload) gives the same bits, since
    U(1) recoupling coefficients are 1 and both sum the k products in order."""
    import os
    # both runtimes with the tiny route off: this compares the fused pass with
    # the general route (the tiny route has tests of its own)
    os.environ["TENET_AMD_TINY_FLOPS"] = "0"
    rt1 = Runtime(0)
    rt2 = None
    wl = wl2 = None
    try:
        wl = MPSEnvUpdate(rt1, "U1", 256, prec)
        assert wl.contractions[1].route == "fused_narrow"
        assert wl.contractions[0].route == wl.contractions[2].route == "transform_gemm"
        out = wl.run().to_numpy().copy()
        t2 = wl._tmp["T2"].to_numpy().copy()
        os.environ["TENET_AMD_FUSE_NARROW"] = "0"
        rt2 = Runtime(0)
        wl2 = MPSEnvUpdate(rt2, "U1", 256, prec)
        assert wl2.contractions[1].route == "transform_gemm"
        assert np.array_equal(wl2._tmp["T2"].to_numpy(), t2)
        assert np.array_equal(wl2.run().to_numpy(), out)
    finally:
        os.environ.pop("TENET_AMD_FUSE_NARROW", None)
        del os.environ["TENET_AMD_TINY_FLOPS"]
        del wl, wl2
        if rt2 is not None:
            rt2.close()
        rt1.close()


@pytest.mark.parametrize("count,D", [("dimension", 256), ("multiplets", 64)])
def test_su2_fused_mpo_step_matches_general_route_and_oracle(rt, count, D):
    """The SU(2) MPO step (lhs, rhs and output transforms with Multi blocks)
    runs as one fused pass of Multi groups whose coefficients are sums of
    gathered MPO terms: its T2 agrees with the general route
    (TENET_AMD_FUSE_NARROW=0 at upload) to rounding, and the whole update
    with the oracle."""
    import os
    from tenet_amd.tensor import Runtime
    os.environ["TENET_AMD_TINY_FLOPS"] = "0"
    rt2 = None
    wl = wl2 = None
    try:
        wl = MPSEnvUpdate(rt, "SU2", D, "f64", su2_count=count)
        assert wl.contractions[1].route == "fused_narrow"
        out = wl.run().to_numpy().copy()
        t2 = wl._tmp["T2"].to_numpy().copy()
        os.environ["TENET_AMD_FUSE_NARROW"] = "0"
        rt2 = Runtime(0)
        wl2 = MPSEnvUpdate(rt2, "SU2", D, "f64", su2_count=count)
        assert wl2.contractions[1].route == "transform_gemm"
        assert rel(wl2._tmp["T2"].to_numpy(), t2) < 1e-14
        assert rel(wl2.run().to_numpy(), out) < 1e-14
        L, A, W = _o(wl.L), _o(wl.A), _o(wl.W)
        t1 = L.contract([2], A, [0], [0, 1, 2, 3], 2)
        t2o = t1.contract(MPSEnvUpdate.T2_AXES[0], W, *MPSEnvUpdate.T2_AXES[1:])
        want = A.adjoint().contract(MPSEnvUpdate.LP_AXES[0], t2o, *MPSEnvUpdate.LP_AXES[1:])
        assert rel(out, want.data) < 1e-13
    finally:
        os.environ.pop("TENET_AMD_FUSE_NARROW", None)
        del os.environ["TENET_AMD_TINY_FLOPS"]
        del wl, wl2
        if rt2 is not None:
            rt2.close()


@pytest.mark.parametrize("rule_name,D,count", [("U1", 512, "dimension"), ("SU2", 512, "dimension"),
                                              ("SU2", 96, "multiplets")])
def test_output_transform_in_gemm_epilogue_equals_separate_pass(rt, rule_name, D, count):
    """A Single-only output transform that maps whole core columns (the last
    contraction of the environment update) is folded into the tile GEMM's
    epilogue for FP64 / FP32 calls: the same result as GEMM-then-transform
    (TENET_AMD_OUT_EPILOGUE=0 at upload) to rounding -- such an artifact also
    runs its few narrow-output jobs on workgroup tiles, which sum in another
    order than the narrow-output kernel -- also accumulating with alpha and
    beta, and the one-wave tiles (which keep the separate pass) agree."""
    import os
    from tenet_amd import _native as N
    os.environ["TENET_AMD_TINY_FLOPS"] = "0"
    rt2 = None
    wl = wl2 = None
    try:
        wl = MPSEnvUpdate(rt, rule_name, D, "f64", su2_count=count)
        os.environ["TENET_AMD_OUT_EPILOGUE"] = "0"
        rt2 = Runtime(0)
        wl2 = MPSEnvUpdate(rt2, rule_name, D, "f64", seed=3, su2_count=count)
        for prec in ("f64", "f32"):
            wl.precision = wl2.precision = prec
            a = wl.run().to_numpy().copy()
            b = wl2.run().to_numpy().copy()
            assert rel(a, b) < (1e-14 if prec == "f64" else 1e-6), prec
        from tenet_amd.tensor import TensorMap
        c3 = wl.contractions[2]
        wl.precision = wl2.precision = "f64"
        prod = wl.run().to_numpy().copy()  # T2 of this run is the FP64 one
        assert rel(wl2.run().to_numpy(), prod) < 1e-14
        X = TensorMap.rand(rt, wl.out.space, seed=77, normal=True)
        before = X.to_numpy().copy()
        wl.A.adjoint().contract(MPSEnvUpdate.LP_AXES[0], wl._tmp["T2"], *MPSEnvUpdate.LP_AXES[1:], out=X,
                                alpha=2.0, beta=-0.5)
        assert rel(X.to_numpy(), 2.0 * prod - 0.5 * before) < 1e-14
        # forced one-wave tiles keep the separate output pass: same result
        h = N.hip()
        rt._check(h.tnt_contract_set_tile_config(rt._ctx, c3.art, 6))
        wl.precision = "f64"
        small = wl.run().to_numpy()
        rt._check(h.tnt_contract_set_tile_config(rt._ctx, c3.art, -1))
        assert rel(small, prod) < 1e-14
    finally:
        os.environ.pop("TENET_AMD_OUT_EPILOGUE", None)
        del os.environ["TENET_AMD_TINY_FLOPS"]
        del wl, wl2
        if rt2 is not None:
            rt2.close()


def test_fused_route_accumulates_and_reads_adjoint_mpo(rt):
    """Fused narrow route on the device: out = alpha X.op(W) + beta out, with W
    as stored and as a lazy adjoint read with op = adjoint, against the oracle."""
    from tenet_amd.tensor import HomSpace, TensorMap
    from tenet_amd.workloads import mps_env_spaces
    r, V, P, M = mps_env_spaces("U1", 200)
    X = TensorMap.rand(rt, HomSpace(r, [V], [P, M]), seed=31, normal=True)
    W = TensorMap.rand(rt, HomSpace(r, [M, P], [P, M]), seed=32, normal=True)
    oX, oW = _o(X), _o(W)
    for adj, axes_b in ((False, [1, 0]), (True, [0, 1])):
        op = W.adjoint() if adj else W
        c = rt.contraction(X.space, W.space, [1, 2], axes_b, [0, 2, 1], 2, False, adj)
        assert c.route == "fused_narrow"
        out = X.contract([1, 2], op, axes_b, [0, 2, 1], 2)
        prod = oX.contract([1, 2], oW.adjoint() if adj else oW, axes_b, [0, 2, 1], 2).data
        assert rel(out.to_numpy(), prod) < 1e-14
        before = out.to_numpy()
        X.contract([1, 2], op, axes_b, [0, 2, 1], 2, alpha=2.0, beta=-0.5, out=out)
        assert rel(out.to_numpy(), 2.0 * prod - 0.5 * before) < 1e-14


@pytest.mark.parametrize("prec", ["f64", "split_bf16", "split_i8"])
def test_env_update_graph_replays_match_eager(rt, prec):
    """Every replay of the captured step equals the eager step bit for bit,
    with eager work and synchronisations between the replays."""
    wl = MPSEnvUpdate(rt, "U1", 256, prec)
    eager = wl.run().to_numpy().copy()
    wl.capture()
    for _ in range(3):
        got = wl.replay().to_numpy()
        assert np.array_equal(got, eager)
        wl.run()
        rt.synchronize()
        assert np.array_equal(wl.out.to_numpy(), eager)


def test_runtime_pool_leases_from_threads(rt):
    """Operations leased from one pool by four host threads run on separate
    contexts (streams) and give the same bits as one runtime; a failed lease
    quarantines its context."""
    from concurrent.futures import ThreadPoolExecutor

    from descs import spaces
    from tenet_amd.pool import RuntimePool
    from tenet_amd.tensor import TensorMap
    leg = [(-1, 12), (0, 20), (1, 10)]
    ps, _ = spaces("U1", [(leg, False), (leg, False)], [(leg, False), (leg, False)])

    def op(r, seed):
        A = TensorMap.rand(r, ps, seed=seed, normal=True)
        B = TensorMap.rand(r, ps, seed=seed + 1, normal=True)
        return A.contract([2, 3], B, [0, 1]).to_numpy()

    pool = RuntimePool(max_idle=4)
    try:
        def leased(seed):
            with pool.lease() as r:
                return id(r), op(r, seed)

        with ThreadPoolExecutor(4) as ex:
            got = list(ex.map(leased, range(0, 32, 2)))
        assert 1 < pool.minted <= 4 and pool.idle == pool.minted
        for seed, (_, out) in zip(range(0, 32, 2), got):
            assert np.array_equal(out, op(rt, seed))
        with pytest.raises(ValueError):
            with pool.lease() as r:
                bad = r
                TensorMap.from_numpy(r, ps, np.zeros(3))  # wrong payload length: rejected on the host
        assert pool.quarantined == 1 and bad._ctx is None
    finally:
        pool.close()


@pytest.mark.parametrize("prec", ["f64", "split_i8"])
def test_fmove_operand_transforms_on_two_streams_equal_one_stream(rt, monkeypatch, prec):
    """Operand transforms that each move >= 64 MiB run on two streams (the rhs
    one on the context's side stream, joined before the GEMM).  Same kernels,
    so the result equals the one-stream result bit for bit; two contractions
    issued back to back on the same artifact (the second reuses the transform
    scratch payloads the first one's GEMM reads) must each keep their result."""
    from tenet_amd.tensor import TensorMap
    wl = SU2FMoveContract(rt, mult=24, twoj_max=8, precision=prec)
    c = wl.contraction.desc.contents
    assert c.lhs and c.rhs
    A2 = TensorMap.rand(rt, wl.A.space, seed=11)
    B2 = TensorMap.rand(rt, wl.B.space, seed=12)

    def pair():
        o1 = wl.A.contract([1, 3], wl.B, [2, 0], [0, 1, 2, 3], 2, precision=prec)
        o2 = A2.contract([1, 3], B2, [2, 0], [0, 1, 2, 3], 2, precision=prec)
        return o1.to_numpy(), o2.to_numpy()

    forked = pair()
    monkeypatch.setenv("TENET_AMD_FORK_TRANSFORMS", "0")
    serial = pair()
    assert np.array_equal(forked[0], serial[0]) and np.array_equal(forked[1], serial[1])
    assert not np.array_equal(serial[0], serial[1])


def test_split_i8_single_transform_on_side_stream_equals_one_stream(rt, monkeypatch):
    """A contraction with one large operand transform (B's legs contracted in
    reversed order, A's domain as it is) in split-INT8: B's transform and its
    exponent and digit passes run on the side stream while A's passes run on
    the main stream; the result equals the one-stream result bit for bit and
    the FP64 one within 1e-10."""
    from tenet_amd.tensor import TensorMap
    from tenet_amd.workloads import su2_fmove_space
    _, V = su2_fmove_space(24, 8)
    A = TensorMap.rand(rt, [V, V], [V, V], seed=21)
    B = TensorMap.rand(rt, [V, V], [V, V], seed=22)
    args = ([2, 3], B, [1, 0], [0, 1, 2, 3], 2)
    c = rt.contraction(A.space, B.space, [2, 3], [1, 0], [0, 1, 2, 3], 2).desc.contents
    assert c.rhs and not c.lhs and not c.out
    assert c.rhs.contents.src_len * 16 >= 64 * 2 ** 20  # the transform moves >= 64 MiB
    forked = A.contract(*args, precision="split_i8").to_numpy()
    again = A.contract(*args, precision="split_i8").to_numpy()
    monkeypatch.setenv("TENET_AMD_FORK_TRANSFORMS", "0")
    serial = A.contract(*args, precision="split_i8").to_numpy()
    assert np.array_equal(forked, serial) and np.array_equal(again, serial)
    ref = A.contract(*args, precision="f64").to_numpy()
    assert rel(serial, ref) < 1e-10


@pytest.mark.parametrize("rule_name,D,count", [("U1", 256, "dimension"), ("SU2", 256, "dimension"),
                                               ("SU2", 64, "multiplets")])
def test_fused_mpo_step_gather_in_launch_equals_gather_launch(rt, monkeypatch, rule_name, D, count):
    """The fused MPO step gathers its coefficients inside the recoupling launch
    (per item, into LDS) when its plan is all few-source Multi items; the sums
    are formed as the separate gather launch forms them, so the result is the
    same bit for bit, with beta != 0 too."""
    from tenet_amd.tensor import TensorMap
    wl = MPSEnvUpdate(rt, rule_name, D, "f64", su2_count=count)
    assert wl.contractions[1].route == "fused_narrow"
    T1 = wl._tmp["T1"]
    X0 = TensorMap.rand(rt, wl._tmp["T2"].space, seed=5)

    def step():
        X = TensorMap.from_numpy(rt, X0.space, X0.to_numpy())
        T1.contract(MPSEnvUpdate.T2_AXES[0], wl.W, *MPSEnvUpdate.T2_AXES[1:], out=X, alpha=0.75, beta=-1.5)
        return X.to_numpy()

    folded = step()
    monkeypatch.setenv("TENET_AMD_GATHER_FOLD", "0")
    separate = step()
    assert np.array_equal(folded, separate)
