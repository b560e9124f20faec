"""GPU tests of the C-ABI boundary: the normal fill against the oracle, the
MatrixOp codes, FP32 / complex payloads, validation before any work is
queued, nothing-contracted jobs, the table cache under a tiny cap, memory kept
alive for graphs, and injected allocation / launch failures.
"""
import ctypes as C

import numpy as np
import pytest

from oracle.tensor import splitmix_normal
from tenet_amd import _native as N
from tenet_amd.pool import RuntimePool
from tenet_amd.tensor import GradedSpace, HomSpace, Rule, Runtime, TensorMap, TenetError

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rt():
    r = Runtime(0)
    yield r
    r.close()


def h():
    return N.hip()


def alloc(rt, n, dtype=N.F64):
    b = C.c_void_p()
    rt._check(h().tnt_buf_alloc(rt._ctx, max(1, n), dtype, C.byref(b)))
    return b


def put(rt, arr, dtype=N.F64):
    npt = np.float64 if dtype == N.F64 else np.float32
    a = np.ascontiguousarray(arr, dtype=npt)
    b = alloc(rt, len(a), dtype)
    if len(a):
        rt._check(h().tnt_buf_upload(rt._ctx, b, a.ctypes.data, len(a)))
    return b


def get(rt, b, n, dtype=N.F64):
    out = np.empty(n, dtype=np.float64 if dtype == N.F64 else np.float32)
    rt._check(h().tnt_buf_download(rt._ctx, b, out.ctypes.data, n))
    return out


def free(rt, *bs):
    for b in bs:
        h().tnt_buf_free(rt._ctx, b)


def jobs_of(shapes):
    js, lo, ro, do = [], 0, 0, 0
    for m, k, n in shapes:
        js.append(N.GemmJob(do, lo, ro, m, k, n))
        lo, ro, do = lo + m * k, ro + k * n, do + m * n
    return (N.GemmJob * len(js))(*js), lo, ro, do


def dense_ref(shapes, lhs, rhs, d0, alpha, beta, lt, rt_):
    """alpha op(A) op(B) + beta D per job; stored matrices column-major (ld = rows)."""
    out, lo, ro, do = [], 0, 0, 0
    for m, k, n in shapes:
        a = lhs[lo:lo + m * k].reshape(m, k) if lt else lhs[lo:lo + m * k].reshape(k, m).T
        b = rhs[ro:ro + k * n].reshape(k, n) if rt_ else rhs[ro:ro + k * n].reshape(n, k).T
        out.append((alpha * (a @ b) + beta * d0[do:do + m * n].reshape(n, m).T).T.reshape(-1))
        lo, ro, do = lo + m * k, ro + k * n, do + m * n
    return np.concatenate(out)


# ---------------------------------------------------------------------------
# normal fill
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("seed", [0, 3, 2 ** 63 + 5])
@pytest.mark.parametrize("n", [1, 2, 3, (1 << 20) + 3])
def test_normal_fill_bit_identical_to_oracle(rt, seed, n):
    b = alloc(rt, n)
    try:
        rt._check(h().tnt_buf_fill_random(rt._ctx, b, seed, 1))
        got = get(rt, b, n)
    finally:
        free(rt, b)
    want = splitmix_normal(seed, n)
    diff = np.flatnonzero(got.view(np.uint64) != want.view(np.uint64))
    assert diff.size == 0, f"{diff.size} of {n} differ, first at {diff[:4]}: {got[diff[:4]]} vs {want[diff[:4]]}"


# ---------------------------------------------------------------------------
# MatrixOp codes, payload types, validation before mutation
# ---------------------------------------------------------------------------
SHAPES = [(37, 19, 23), (64, 64, 64), (5, 130, 17), (129, 8, 3)]


@pytest.mark.parametrize("lop,rop", [(0, 0), (1, 2), (2, 1), (2, 2)])
def test_transpose_and_adjoint_codes_agree(rt, lop, rop):
    """MatrixOp::Transpose (1) and ::Adjoint (2) are the same operation on real
    payloads; both match numpy."""
    rng = np.random.default_rng(lop * 3 + rop)
    arr, nl, nr, nd = jobs_of(SHAPES)
    lhs, rhs, d0 = rng.normal(size=nl), rng.normal(size=nr), rng.normal(size=nd)
    bl, br = put(rt, lhs), put(rt, rhs)
    outs = []
    try:
        for ops in ((lop, rop), (min(lop, 1), min(rop, 1))):
            bd = put(rt, d0)
            rt._check(h().tnt_gemm_batch(rt._ctx, bd, bl, br, arr, len(SHAPES), None, 0, ops[0], ops[1], 0.5, -1.25, 0))
            outs.append(get(rt, bd, nd))
            free(rt, bd)
    finally:
        free(rt, bl, br)
    assert np.array_equal(outs[0], outs[1])
    want = dense_ref(SHAPES, lhs, rhs, d0, 0.5, -1.25, lop != 0, rop != 0)
    assert np.linalg.norm(outs[0] - want) / np.linalg.norm(want) < 1e-14


@pytest.mark.parametrize("prec,tol", [(N.PREC_F64, 2e-7), (N.PREC_F32, 2e-6)])
@pytest.mark.parametrize("lop,rop", [(0, 0), (2, 0), (0, 2)])
def test_f32_payload_gemm(rt, prec, tol, lop, rop):
    """DenseRead::F32 batches: FP32 in, FP32 out.  FP64 arithmetic rounds once
    (error ~2^-24); FP32 arithmetic accumulates in FP32."""
    rng = np.random.default_rng(7 + prec)
    arr, nl, nr, nd = jobs_of(SHAPES)
    lhs = rng.normal(size=nl).astype(np.float32)
    rhs = rng.normal(size=nr).astype(np.float32)
    d0 = rng.normal(size=nd).astype(np.float32)
    bl, br, bd = put(rt, lhs, N.F32), put(rt, rhs, N.F32), put(rt, d0, N.F32)
    try:
        rt._check(h().tnt_gemm_batch(rt._ctx, bd, bl, br, arr, len(SHAPES), None, 0, lop, rop, 1.5, 0.5, prec))
        got = get(rt, bd, nd, N.F32).astype(np.float64)
    finally:
        free(rt, bl, br, bd)
    want = dense_ref(SHAPES, lhs.astype(np.float64), rhs.astype(np.float64), d0.astype(np.float64), 1.5, 0.5,
                     lop != 0, rop != 0)
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < tol


def test_f32_zero_and_roundtrip(rt):
    x = np.arange(1001, dtype=np.float32) * 0.5
    b = put(rt, x, N.F32)
    try:
        assert np.array_equal(get(rt, b, 1001, N.F32), x)
        rt._check(h().tnt_buf_zero(rt._ctx, b))
        assert not get(rt, b, 1001, N.F32).any()
    finally:
        free(rt, b)


def test_payload_type_rejections(rt):
    b = C.c_void_p()
    assert h().tnt_buf_alloc(rt._ctx, 16, 9, C.byref(b)) == N.E_INVALID
    assert h().tnt_buf_wrap(rt._ctx, C.c_void_p(4096), 16, -1, C.byref(b)) == N.E_INVALID
    f32, f64 = put(rt, np.ones(64), N.F32), put(rt, np.ones(64))
    arr, *_ = jobs_of([(4, 4, 4)])
    try:
        # mixed payload types
        assert h().tnt_gemm_batch(rt._ctx, f64, f32, f64, arr, 1, None, 0, 0, 0, 1.0, 0.0, 0) == N.E_INVALID
        # split precisions need FP64 payloads
        for prec in (N.PREC_SPLIT_BF16, N.PREC_SPLIT_I8):
            assert h().tnt_gemm_batch(rt._ctx, f32, f32, f32, arr, 1, None, 0, 0, 0, 1.0, 0.0, prec) == N.E_UNSUPPORTED
        # FP32 payloads elsewhere
        out = C.c_double()
        reg = (N.Region * 1)(N.Region(0, 4))
        w = (C.c_double * 1)(1.0)
        assert h().tnt_weighted_dot(rt._ctx, f32, f32, reg, w, 1, C.byref(out)) == N.E_UNSUPPORTED
        assert h().tnt_buf_fill_random(rt._ctx, f32, 1, 0) == N.E_UNSUPPORTED
        assert h().tnt_axpby(rt._ctx, f32, f32, 1.0, 0.0) == N.E_UNSUPPORTED
    finally:
        free(rt, f32, f64)


def test_unknown_ops_rejected_everywhere_before_mutating(rt):
    d0 = np.arange(64.0)
    bd, bl, br = put(rt, d0), put(rt, np.ones(64)), put(rt, np.ones(64))
    arr, *_ = jobs_of([(4, 4, 4)])
    try:
        for lop, rop in ((3, 0), (0, -1), (7, 7)):
            assert h().tnt_gemm_batch(rt._ctx, bd, bl, br, arr, 1, None, 0, lop, rop, 1.0, 0.0, 0) == N.E_INVALID
        pd, pl = C.c_void_p(), C.c_void_p()
        rt._check(h().tnt_buf_info(bd, C.byref(pd), None, None, None))
        rt._check(h().tnt_buf_info(bl, C.byref(pl), None, None, None))
        gj = (N.GemmPtrJob * 1)(N.GemmPtrJob(pd.value, pl.value, pl.value, 4, 4, 4, 4, 4, 4, 3, 0))
        assert h().tnt_gemm_grouped(rt._ctx, gj, 1, 1.0, 0.0, 0) == N.E_INVALID
        # contraction artifact with an unknown op: rejected at upload
        desc = N.ContractDesc(None, None, None, 1, arr, 64, 64, 64, 0, None, 5, 0)
        art = C.c_void_p()
        assert h().tnt_contract_upload(rt._ctx, C.byref(desc), C.byref(art)) == N.E_INVALID
        assert get(rt, bd, 64).tolist() == d0.tolist()
    finally:
        free(rt, bd, bl, br)


def test_dimensions_beyond_int32_rejected(rt):
    bd = put(rt, np.zeros(16))
    big = (1 << 31)
    try:
        arr = (N.GemmJob * 1)(N.GemmJob(0, 0, 0, big, 1, 1))
        assert h().tnt_gemm_batch(rt._ctx, bd, bd, bd, arr, 1, None, 0, 0, 0, 1.0, 0.0, 0) == N.E_UNSUPPORTED
        desc = N.ContractDesc(None, None, None, 1, arr, 1 << 40, 1 << 40, 1 << 40, 0, None, 0, 0)
        art = C.c_void_p()
        assert h().tnt_contract_upload(rt._ctx, C.byref(desc), C.byref(art)) == N.E_UNSUPPORTED
        p = C.c_void_p()
        rt._check(h().tnt_buf_info(bd, C.byref(p), None, None, None))
        gj = (N.GemmPtrJob * 1)(N.GemmPtrJob(p.value, p.value, p.value, 1, 1, 1, big, 1, 1, 0, 0))
        assert h().tnt_gemm_grouped(rt._ctx, gj, 1, 1.0, 0.0, 0) == N.E_UNSUPPORTED
    finally:
        free(rt, bd)


# ---------------------------------------------------------------------------
# nothing contracted: D = beta D, every kernel family and precision
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("cfg", [-1, 1, 6])
@pytest.mark.parametrize("prec", [0, 1, 2, 3])
def test_zero_contracted_jobs_scale_destination(rt, cfg, prec):
    shapes = [(33, 0, 17), (5, 0, 40), (64, 3, 64)]
    rng = np.random.default_rng(cfg + 1 + 10 * prec)
    arr, nl, nr, nd = jobs_of(shapes)
    lhs, rhs, d0 = rng.normal(size=nl), rng.normal(size=nr), rng.normal(size=nd)
    bl, br, bd = put(rt, lhs), put(rt, rhs), put(rt, d0)
    try:
        rt._check(h().tnt_set_tile_config(rt._ctx, cfg))
        rt._check(h().tnt_gemm_batch(rt._ctx, bd, bl, br, arr, len(shapes), None, 0, 0, 0, 0.5, 2.0, prec))
        got = get(rt, bd, nd)
    finally:
        h().tnt_set_tile_config(rt._ctx, -1)
        free(rt, bl, br, bd)
    k0 = 33 * 17 + 5 * 40
    assert np.array_equal(got[:k0], 2.0 * d0[:k0])
    want = dense_ref(shapes, lhs, rhs, d0, 0.5, 2.0, False, False)
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < (1e-5 if prec == 1 else 1e-10)


# ---------------------------------------------------------------------------
# table cache: LRU under a tiny cap, entries of one call never evicted by it
# ---------------------------------------------------------------------------
def test_table_cache_tiny_cap_keeps_results_exact():
    r = Runtime(0)
    try:
        r._check(h().tnt_ctx_set_table_cache_cap(r._ctx, 1))
        rng = np.random.default_rng(5)
        x, y = rng.normal(size=5000), rng.normal(size=5000)
        bx, by = put(r, x), put(r, y)
        out = C.c_double()
        for trial in range(40):
            cuts = np.sort(rng.choice(np.arange(1, 5000), size=1 + trial % 7, replace=False))
            edges = np.concatenate([[0], cuts, [5000]])
            regs = (N.Region * (len(edges) - 1))(*[N.Region(int(a), int(b - a)) for a, b in zip(edges[:-1], edges[1:])])
            wts = rng.normal(size=len(edges) - 1)
            w = (C.c_double * len(wts))(*wts)
            r._check(h().tnt_weighted_dot(r._ctx, bx, by, regs, w, len(wts), C.byref(out)))
            want = sum(wt * float(x[a:b] @ y[a:b]) for wt, a, b in zip(wts, edges[:-1], edges[1:]))
            assert abs(out.value - want) <= 1e-12 * max(1.0, abs(want)), trial
            # a batched GEMM in between evicts the dot's tables (cap 1)
            arr, nl, nr, nd = jobs_of([(8 + trial % 5, 3, 4)])
            bd = put(r, np.zeros(nd))
            r._check(h().tnt_gemm_batch(r._ctx, bd, bx, by, arr, 1, None, 0, 0, 0, 1.0, 0.0, 0))
            m = 8 + trial % 5
            want_g = (x[:m * 3].reshape(3, m).T @ y[:12].reshape(4, 3).T).T.reshape(-1)
            assert np.allclose(get(r, bd, nd), want_g, rtol=1e-14, atol=1e-14)
            free(r, bd)
        free(r, bx, by)
    finally:
        r.close()


def test_graph_keeps_pinned_tables_and_retired_workspace():
    """A captured grouped GEMM replays correctly after the cache (cap 1) has
    turned over many times, and after the split workspace has grown (its old
    planes are retired, not freed, while the graph lives)."""
    r = Runtime(0)
    try:
        r._check(h().tnt_ctx_set_table_cache_cap(r._ctx, 1))
        rng = np.random.default_rng(9)
        m, k, n = 48, 40, 24
        a, b = rng.normal(size=m * k), rng.normal(size=k * n)
        ba, bb, bc = put(r, a), put(r, b), put(r, np.zeros(m * n))
        ptr = [C.c_void_p() for _ in range(3)]
        for p, buf in zip(ptr, (bc, ba, bb)):
            r._check(h().tnt_buf_info(buf, C.byref(p), None, None, None))
        gj = (N.GemmPtrJob * 1)(N.GemmPtrJob(ptr[0].value, ptr[1].value, ptr[2].value, m, k, n, m, m, k, 0, 0))
        want = (a.reshape(k, m).T @ b.reshape(n, k).T).T.reshape(-1)
        for prec, tol in ((0, 1e-14), (3, 1e-10)):
            r._check(h().tnt_gemm_grouped(r._ctx, gj, 1, 1.0, 0.0, prec))  # tables exist before capture
            r._check(h().tnt_graph_begin(r._ctx))
            r._check(h().tnt_gemm_grouped(r._ctx, gj, 1, 1.0, 0.0, prec))
            g = C.c_void_p()
            r._check(h().tnt_graph_end(r._ctx, C.byref(g)))
            try:
                for rep in range(3):
                    # eager work that turns the cache over and grows the split workspace
                    big = 200 + 100 * rep
                    x = put(r, rng.normal(size=big * big))
                    y = put(r, np.zeros(big * big))
                    arr = (N.GemmJob * 1)(N.GemmJob(0, 0, 0, big, big, big))
                    r._check(h().tnt_gemm_batch(r._ctx, y, x, x, arr, 1, None, 0, 0, 0, 1.0, 0.0, 3))
                    r.synchronize()
                    free(r, x, y)
                    r._check(h().tnt_buf_zero(r._ctx, bc))
                    r._check(h().tnt_graph_launch(r._ctx, g))
                    r.synchronize()
                    got = get(r, bc, m * n)
                    assert np.linalg.norm(got - want) / np.linalg.norm(want) < tol, (prec, rep)
            finally:
                r._check(h().tnt_graph_destroy(r._ctx, g))
        free(r, ba, bb, bc)
    finally:
        r.close()


# ---------------------------------------------------------------------------
# injected failures: validate before mutate, clean unwinding, quarantine
# ---------------------------------------------------------------------------
def test_injected_allocation_failure_unwinds_upload():
    r = Runtime(0)
    try:
        V = GradedSpace(Rule("U1"), [(-1, 5), (0, 7), (1, 4)])
        A = TensorMap.rand(r, [V, V], [V, V], seed=1)
        B = TensorMap.rand(r, [V, V], [V, V], seed=2)
        want = A.contract([1, 3], B, [2, 0], [0, 2, 1, 3], 2).to_numpy()
        for after in range(0, 6):
            r.clear_caches()  # every plan and artifact is uploaded again
            r._check(h().tnt_ctx_inject_fault(r._ctx, N.FAULT_ALLOC, after))
            try:
                A.permute([0, 1], [2, 3]).contract([1, 3], B.permute([1, 0], [2, 3]), [2, 0], [0, 2, 1, 3], 2)
            except TenetError as e:
                assert e.code == N.E_OOM
            finally:
                r._check(h().tnt_ctx_inject_fault(r._ctx, N.FAULT_NONE, -1))
        # the context still works and gives the same result
        got = A.contract([1, 3], B, [2, 0], [0, 2, 1, 3], 2).to_numpy()
        assert np.array_equal(got, want)
    finally:
        r.close()


def test_injected_failure_before_gemm_leaves_destination_untouched(rt):
    """An allocation failure inside tnt_gemm_batch (its job-table upload) comes
    before any kernel: the destination keeps its values."""
    rng = np.random.default_rng(3)
    shapes = [(11, 13, 17), (3, 5, 7)]
    arr, nl, nr, nd = jobs_of(shapes)
    d0 = rng.normal(size=nd)
    bl, br, bd = put(rt, rng.normal(size=nl)), put(rt, rng.normal(size=nr)), put(rt, d0)
    try:
        rt._check(h().tnt_ctx_inject_fault(rt._ctx, N.FAULT_ALLOC, 0))
        rc = h().tnt_gemm_batch(rt._ctx, bd, bl, br, arr, 2, None, 0, 0, 0, 1.0, 0.0, 0)
        rt._check(h().tnt_ctx_inject_fault(rt._ctx, N.FAULT_NONE, -1))
        assert rc == N.E_OOM
        assert np.array_equal(get(rt, bd, nd), d0)
    finally:
        h().tnt_ctx_inject_fault(rt._ctx, N.FAULT_NONE, -1)
        free(rt, bl, br, bd)


def test_injected_launch_failure_quarantines_leased_runtime():
    pool = RuntimePool(device=0, max_idle=2)
    V = GradedSpace(Rule("U1"), [(-1, 3), (0, 4), (1, 2)])
    with pytest.raises(TenetError) as ei:
        with pool.lease() as bad:
            A = TensorMap.rand(bad, [V, V], [V, V], seed=1)
            bad._check(h().tnt_ctx_inject_fault(bad._ctx, N.FAULT_LAUNCH, 0))
            A.contract([1, 3], A, [2, 0], [0, 2, 1, 3], 2)
    assert ei.value.code == N.E_HIP
    assert pool.quarantined == 1 and pool.idle == 0
    with pool.lease() as good:
        assert good is not bad
        A = TensorMap.rand(good, [V, V], [V, V], seed=1)
        assert np.isfinite(A.contract([1, 3], A, [2, 0], [0, 2, 1, 3], 2).to_numpy()).all()
    assert pool.idle == 1
    pool.close()


@pytest.mark.parametrize("cfg", [0, 3, 5])
@pytest.mark.parametrize("beta", [0.0, 0.75])
def test_workgroup_tiles_long_k_and_ragged_edges(rt, cfg, beta):
    """Long K (many slabs, odd tail), several tiles per job and ragged edges:
    the workgroup tiles against numpy."""
    shapes = [(300, 1001, 260), (129, 16, 1), (1, 17, 129), (256, 257, 128)]
    rng = np.random.default_rng(cfg * 7 + int(4 * beta))
    arr, nl, nr, nd = jobs_of(shapes)
    lhs, rhs, d0 = rng.normal(size=nl), rng.normal(size=nr), rng.normal(size=nd)
    bl, br, bd = put(rt, lhs), put(rt, rhs), put(rt, d0)
    try:
        rt._check(h().tnt_set_tile_config(rt._ctx, cfg))
        rt._check(h().tnt_gemm_batch(rt._ctx, bd, bl, br, arr, len(shapes), None, 0, 0, 0, 1.25, beta, 0))
        got = get(rt, bd, nd)
    finally:
        h().tnt_set_tile_config(rt._ctx, -1)
        free(rt, bl, br, bd)
    want = dense_ref(shapes, lhs, rhs, d0, 1.25, beta, False, False)
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-14
