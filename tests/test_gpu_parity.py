"""GPU parity of the gfx950 backend against the CPU oracle (through the C ABI).

Bars: integer/index/layout work and U(1) permutations bit-exact; FP64 GEMM
results within 1e-13 relative Frobenius error of the FP64 oracle (accumulation
order differs from numpy's); FP32 GEMM within 1e-5.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

from descs import spaces
from oracle.fusion import HomSpace as OHomSpace, Leg as OLeg
from oracle.symmetry import make_rule
from oracle.tensor import OTensor, splitmix_unit
from tenet_amd import _native as N
from tenet_amd.tensor import GradedSpace, HomSpace, Rule, Runtime, TensorMap

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(__file__)
KAT = json.load(open(os.path.join(HERE, "golden", "kat.json")))
GOLD = json.load(open(os.path.join(HERE, "golden", "tensorkit_oracle.json")))


@pytest.fixture(scope="module")
def rt():
    r = Runtime(0)
    yield r
    r.close()


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb else 1.0)


def _buf(rt, arr):
    b = C.c_void_p()
    rt._check(N.hip().tnt_buf_alloc(rt._ctx, max(1, len(arr)), 0, C.byref(b)))
    if len(arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        rt._check(N.hip().tnt_buf_upload(rt._ctx, b, a.ctypes.data, len(a)))
    return b


def _get(rt, b, n):
    out = np.empty(n)
    rt._check(N.hip().tnt_buf_download(rt._ctx, b, out.ctypes.data, n))
    return out


def _free(rt, *bs):
    for b in bs:
        N.hip().tnt_buf_free(rt._ctx, b)


# ---------------------------------------------------------------------------
# GEMM seam (DenseExecutor::matmul_batch_axpby_with_ops_into)
# ---------------------------------------------------------------------------
GEMM_SHAPES = [(16, 4, 16), (1, 1, 1), (17, 5, 33), (64, 64, 64), (130, 70, 129), (300, 17, 5),
               (5, 300, 7), (256, 256, 256), (129, 1000, 65)]


@pytest.mark.parametrize("lop,rop", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("prec", [0, 1, 2, 3])
def test_gemm_batch_matches_numpy(rt, lop, rop, prec):
    rng = np.random.default_rng(7 + 2 * lop + rop)
    jobs, lhs, rhs, dst_parts, want = [], [], [], [], []
    lo = ro = do = 0
    for m, k, n in GEMM_SHAPES:
        A = rng.standard_normal((m, k))
        B = rng.standard_normal((k, n))
        D0 = rng.standard_normal((m, n))
        sa = A.T if lop else A  # stored matrix
        sb = B.T if rop else B
        lhs.append(sa.reshape(-1, order="F"))
        rhs.append(sb.reshape(-1, order="F"))
        dst_parts.append(D0.reshape(-1, order="F"))
        want.append((0.5 * A @ B + 2.0 * D0).reshape(-1, order="F"))
        jobs.append(N.GemmJob(do, lo, ro, m, k, n))
        lo += m * k
        ro += k * n
        do += m * n
    L, R, D = (np.concatenate(x) for x in (lhs, rhs, dst_parts))
    bl, br, bd = _buf(rt, L), _buf(rt, R), _buf(rt, D)
    arr = (N.GemmJob * len(jobs))(*jobs)
    rt._check(N.hip().tnt_gemm_batch(rt._ctx, bd, bl, br, arr, len(jobs), None, 0, lop, rop, 0.5, 2.0, prec))
    got = _get(rt, bd, len(D))
    _free(rt, bl, br, bd)
    w = np.concatenate(want)
    tol = {0: 1e-13, 1: 1e-5, 2: 1e-10, 3: 1e-10}[prec]
    o = 0
    for (m, k, n), part in zip(GEMM_SHAPES, want):
        assert rel(got[o:o + m * n], part) < tol, ((m, k, n), lop, rop, prec)
        o += m * n
    assert rel(got, w) < tol


@pytest.mark.parametrize("cfg", [0, 1, 2, 3, 5, 6])
@pytest.mark.parametrize("lop,rop", [(0, 0), (1, 1)])
@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (0.5, -1.25)])
def test_gemm_tile_configs_match_numpy(rt, cfg, lop, rop, alpha, beta):
    """Every tile configuration (forced) on ragged FP64 jobs, both operand
    ops, overwriting (beta = 0) and accumulating into a nonzero D."""
    rng = np.random.default_rng(cfg)
    shapes = [(130, 70, 129), (300, 17, 65), (64, 64, 64), (257, 33, 200)]
    jobs, lhs, rhs, want = [], [], [], []
    lo = ro = do = 0
    for m, k, n in shapes:
        a, b = rng.normal(size=(m, k)), rng.normal(size=(k, n))
        lhs.append((a.T if lop else a).reshape(-1, order="F"))
        rhs.append((b.T if rop else b).reshape(-1, order="F"))
        jobs.append(N.GemmJob(do, lo, ro, m, k, n))
        want.append((a @ b).reshape(-1, order="F"))
        lo, ro, do = lo + m * k, ro + k * n, do + m * n
    d0 = rng.normal(size=do)
    L, R, D = _buf(rt, np.concatenate(lhs)), _buf(rt, np.concatenate(rhs)), _buf(rt, d0)
    h = N.hip()
    try:
        h.tnt_set_tile_config(rt._ctx, cfg)
        arr = (N.GemmJob * len(jobs))(*jobs)
        rt._check(h.tnt_gemm_batch(rt._ctx, D, L, R, arr, len(jobs), None, 0, lop, rop, alpha, beta, 0))
        got = _get(rt, D, do)
    finally:
        h.tnt_set_tile_config(rt._ctx, -1)
        _free(rt, L, R, D)
    w = alpha * np.concatenate(want) + beta * d0
    assert rel(got, w) < 1e-14


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("lop,rop", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_gemm_small_tiles_bit_identical_to_workgroup_tiles(rt, prec, lop, rop):
    """The one-wave 32 x 32 kernel (registers only) and the 64 x 64 LDS kernel
    issue the same MFMA sequence over the same k groups of four, so their
    results agree bit for bit (ragged jobs, k tails, beta != 0)."""
    rng = np.random.default_rng(100 + 4 * lop + 2 * rop + prec)
    shapes = [(30, 16, 30), (1, 7, 33), (33, 1, 17), (64, 64, 64), (65, 9, 31), (100, 37, 90)]
    jobs, lhs, rhs, dst = [], [], [], []
    lo = ro = do = 0
    for m, k, n in shapes:
        a, b = rng.normal(size=(m, k)), rng.normal(size=(k, n))
        lhs.append((a.T if lop else a).reshape(-1, order="F"))
        rhs.append((b.T if rop else b).reshape(-1, order="F"))
        dst.append(rng.normal(size=m * n))
        jobs.append(N.GemmJob(do, lo, ro, m, k, n))
        lo, ro, do = lo + m * k, ro + k * n, do + m * n
    L, R = _buf(rt, np.concatenate(lhs)), _buf(rt, np.concatenate(rhs))
    D0 = np.concatenate(dst)
    h = N.hip()
    out = {}
    try:
        for cfg in (1, 6):
            D = _buf(rt, D0)
            h.tnt_set_tile_config(rt._ctx, cfg)
            arr = (N.GemmJob * len(jobs))(*jobs)
            rt._check(h.tnt_gemm_batch(rt._ctx, D, L, R, arr, len(jobs), None, 0, lop, rop, 0.75, -1.5, prec))
            out[cfg] = _get(rt, D, do)
            _free(rt, D)
    finally:
        h.tnt_set_tile_config(rt._ctx, -1)
        _free(rt, L, R)
    assert np.array_equal(out[1], out[6])


def test_gemm_batch_rejects_out_of_bounds_before_mutating(rt):
    D = np.arange(16.0)
    bd, bl, br = _buf(rt, D), _buf(rt, np.ones(16)), _buf(rt, np.ones(16))
    jobs = (N.GemmJob * 2)(N.GemmJob(0, 0, 0, 2, 2, 2), N.GemmJob(10, 0, 0, 4, 2, 2))
    rc = N.hip().tnt_gemm_batch(rt._ctx, bd, bl, br, jobs, 2, None, 0, 0, 0, 1.0, 0.0, 0)
    assert rc == -4
    assert np.array_equal(_get(rt, bd, 16), D)
    _free(rt, bd, bl, br)


def test_gemm_grouped_pointer_jobs(rt):
    rng = np.random.default_rng(3)
    tensors = []
    jobs = []
    wants = []
    for m, k, n in [(70, 33, 90), (128, 64, 128), (5, 200, 9), (257, 129, 66)]:
        A, B = rng.standard_normal((m, k)), rng.standard_normal((k, n))
        ba, bb, bd = _buf(rt, A.reshape(-1, order="F")), _buf(rt, B.reshape(-1, order="F")), _buf(rt, np.zeros(m * n))
        pa, pb, pd = C.c_void_p(), C.c_void_p(), C.c_void_p()
        for b, p in ((ba, pa), (bb, pb), (bd, pd)):
            N.hip().tnt_buf_info(b, C.byref(p), None, None, None)
        jobs.append(N.GemmPtrJob(pd.value, pa.value, pb.value, m, k, n, m, m, k, 0, 0))
        tensors.append((ba, bb, bd, m, n))
        wants.append(A @ B)
    arr = (N.GemmPtrJob * len(jobs))(*jobs)
    rt._check(N.hip().tnt_gemm_grouped(rt._ctx, arr, len(jobs), 1.0, 0.0, 0))
    for (ba, bb, bd, m, n), w in zip(tensors, wants):
        got = _get(rt, bd, m * n).reshape(n, m).T
        assert rel(got, w) < 1e-13
        _free(rt, ba, bb, bd)


# ---------------------------------------------------------------------------
# Replay known-answer tests through tnt_plan_upload / tnt_tree_transform
# ---------------------------------------------------------------------------
def _desc_from(dst_blocks, src_blocks, multis, dst_len, src_len):
    from oracle.replay import packed_column_major
    dl = packed_column_major(dst_blocks)
    sl = packed_column_major(src_blocks)
    layouts, blocks, coeffs = [], [], []

    def lay(off, shape, strides):
        L = N.Layout()
        L.offset, L.rank = off, len(shape)
        for i, (s, t) in enumerate(zip(shape, strides)):
            L.shape[i], L.strides[i] = s, t
        return L
    for dsts, srcs, U in multis:
        b = N.TransformBlock(1, len(layouts), len(dsts), 0, len(srcs), len(coeffs))
        for d in dsts:
            layouts.append(lay(*dl[d]))
        b.src_layout = len(layouts)
        for s in srcs:
            # source layouts are read in destination axis order: reshape to the
            # destination block's shape (packed column = column-major index)
            layouts.append(lay(*sl[s]))
        coeffs.extend(U)
        blocks.append(b)
    keep = (layouts, blocks, coeffs)
    desc = N.TransformDesc()
    desc.dst_len, desc.src_len = dst_len, src_len
    desc.n_layouts = len(layouts)
    desc.layouts = (N.Layout * len(layouts))(*layouts)
    desc.n_blocks = len(blocks)
    desc.blocks = (N.TransformBlock * len(blocks))(*blocks)
    desc.n_coeffs = len(coeffs)
    desc.coeffs = (C.c_double * len(coeffs))(*coeffs)
    desc.n_inactive = 0
    return desc, keep


@pytest.mark.parametrize("name", ["multi_2x2", "multi_orientation", "multi_sorted_offsets"])
def test_replay_kats_exact(rt, name):
    k = KAT[name]
    if name == "multi_sorted_offsets":
        dstb = srcb = k["blocks"]
        multis = [(m["dst_idx"], m["src_idx"], m["U_dst_src"]) for m in k["multis"]]
        n = len(k["src"])
        dlen = slen = n
    else:
        dstb, srcb = k["dst_blocks"], k["src_blocks"]
        multis = [(k["dst_idx"], k["src_idx"], k["U_dst_src"])]
        dlen = sum(int(np.prod(s)) for s in dstb)
        slen = len(k["src"])
    desc, keep = _desc_from(dstb, srcb, multis, dlen, slen)
    plan = C.c_void_p()
    rt._check(N.hip().tnt_plan_upload(rt._ctx, C.byref(desc), C.byref(plan)))
    bs = _buf(rt, np.array(k["src"], dtype=float))
    bd = _buf(rt, np.full(dlen, k["fill"], dtype=float))
    ow = 1 if k["beta"] == 0.0 else 0
    rt._check(N.hip().tnt_tree_transform(rt._ctx, plan, bd, bs, k["alpha"], k["beta"], ow))
    got = _get(rt, bd, dlen)
    N.hip().tnt_plan_destroy(rt._ctx, plan)
    _free(rt, bs, bd)
    assert list(got) == k["expected"]


# ---------------------------------------------------------------------------
# Layout-level parity: seeded fill, permute, contraction
# ---------------------------------------------------------------------------
U1V = [(-2, 3), (-1, 5), (0, 7), (1, 5), (2, 3)]
SU2V = [(0, 3), (1, 2), (2, 2), (3, 1)]


def test_seeded_fill_bitwise(rt):
    V = GradedSpace(Rule("U1"), U1V)
    t = TensorMap.rand(rt, [V, V], [V, V], seed=123456789)
    assert np.array_equal(t.to_numpy(), splitmix_unit(123456789, t.space.len))


@pytest.mark.parametrize("rule_name,pairs", [("U1", U1V), ("SU2", SU2V),
                                             ("U1xZ2", [((0, 0), 3), ((1, 1), 2), ((-1, 1), 2), ((2, 0), 1)])])
@pytest.mark.parametrize("p1,p2", [([1, 0], [2, 3]), ([0, 2], [1, 3]), ([2, 0, 1], [3]), ([3], [0, 1, 2]),
                                   ([1, 3], [2, 0]), ([3, 2, 1, 0], [])])
def test_permute_matches_oracle(rt, rule_name, pairs, p1, p2):
    ps, os_ = spaces(rule_name, [(pairs, False), (pairs, True)], [(pairs, False), (pairs, False)])
    t = TensorMap.rand(rt, ps, seed=11)
    src = t.to_numpy()
    got = t.permute(p1, p2).to_numpy()
    want = OTensor(os_, src).permute(p1, p2).data
    if rule_name == "SU2":
        assert np.abs(got - want).max() <= 1e-14 * max(1.0, np.abs(want).max())
    else:
        assert np.array_equal(got, want), "abelian permutation must be bit-exact"


BIGV = [(-1, 70), (0, 130), (1, 40)]  # ragged blocks of 16..130 rows: full and partial 64 x 64 tiles


@pytest.mark.parametrize("cod,dom,p1,p2", [(1, 1, [1], [0]), (2, 1, [2, 0], [1]), (2, 1, [1], [2, 0]),
                                           (1, 2, [2, 1], [0])])
def test_transposed_blocks_through_lds_tiles(rt, cod, dom, p1, p2):
    """Blocks whose source is contiguous along another axis than the
    destination go through 64 x 64 LDS tiles: bit-exact, also accumulating."""
    ps, os_ = spaces("U1", [(BIGV, False)] * cod, [(BIGV, False)] * dom)
    t = TensorMap.rand(rt, ps, seed=21)
    src = t.to_numpy()
    out = t.permute(p1, p2)
    want = OTensor(os_, src).permute(p1, p2).data
    assert np.array_equal(out.to_numpy(), want), "abelian permutation must be bit-exact"
    before = out.to_numpy()
    t.permute(p1, p2, out=out, alpha=2.0, beta=-0.5)
    assert np.array_equal(out.to_numpy(), 2.0 * want + -0.5 * before)


SMALLV = [(-1, 3), (0, 5), (1, 2)]


@pytest.mark.parametrize("legs,cod,p1,p2", [(BIGV, 2, [0, 2], [1]), (BIGV, 1, [0, 2], [1]),
                                            (SMALLV, 2, [0, 3], [2, 1]), (SMALLV, 2, [0, 2, 1], [3])])
def test_single_class_flattened_and_table_items_accumulate(rt, legs, cod, p1, p2):
    """Single-only U(1) permutations (one launch of the Single class: items of
    one outer index as flattened records, items of several outer indices
    through the block table), overwriting and accumulating: bit-exact."""
    rank = 3 if legs is BIGV else 4
    ps, os_ = spaces("U1", [(legs, False)] * cod, [(legs, False)] * (rank - cod))
    t = TensorMap.rand(rt, ps, seed=31)
    src = t.to_numpy()
    want = OTensor(os_, src).permute(p1, p2).data
    out = t.permute(p1, p2)
    assert np.array_equal(out.to_numpy(), want)
    before = out.to_numpy()
    t.permute(p1, p2, out=out, alpha=2.0, beta=-0.5)
    assert np.array_equal(out.to_numpy(), 2.0 * want + -0.5 * before)


@pytest.mark.parametrize("small_plan_bytes", ["0", "1e12"])
def test_multi_class_plans_per_class_and_in_one_launch(small_plan_bytes):
    """SU(2) permutations mixing Single, Multi and Scale items: replayed per
    item class and as one launch of every class, same results."""
    import os
    os.environ["TENET_AMD_SMALL_PLAN_BYTES"] = small_plan_bytes
    r = Runtime(0)
    try:
        ps, os_ = spaces("SU2", [(SU2V, False), (SU2V, True)], [(SU2V, False), (SU2V, False)])
        t = TensorMap.rand(r, ps, seed=17)
        src = t.to_numpy()
        for p1, p2 in (([0, 2], [1, 3]), ([3, 0], [2, 1]), ([2], [0, 1, 3])):
            got = t.permute(p1, p2).to_numpy()
            want = OTensor(os_, src).permute(p1, p2).data
            assert np.abs(got - want).max() <= 1e-14 * max(1.0, np.abs(want).max())
        del t
    finally:
        del os.environ["TENET_AMD_SMALL_PLAN_BYTES"]
        r.close()


def test_permute_axpby_into_existing(rt):
    ps, os_ = spaces("SU2", [(SU2V, False), (SU2V, False)], [(SU2V, False), (SU2V, False)])
    t = TensorMap.rand(rt, ps, seed=5)
    out = t.permute([0, 2], [1, 3])
    before = out.to_numpy()
    t.permute([0, 2], [1, 3], out=out, alpha=2.0, beta=-0.5)
    want = 2.0 * OTensor(os_, t.to_numpy()).permute([0, 2], [1, 3]).data - 0.5 * before
    assert np.abs(out.to_numpy() - want).max() < 1e-13


CONTRACTS = [([2, 3], [0, 1], [0, 1, 2, 3], 2), ([3, 2], [0, 1], [0, 1, 2, 3], 2),
             ([3, 2], [0, 1], [1, 0, 2, 3], 2), ([1, 3], [2, 0], [0, 2, 1, 3], 2),
             ([0], [2], [0, 1, 2, 3, 4, 5], 3)]


@pytest.mark.parametrize("rule_name,pairs", [("U1", U1V), ("SU2", SU2V)])
@pytest.mark.parametrize("ci", range(len(CONTRACTS)))
@pytest.mark.parametrize("prec", ["f64", "f32", "split_bf16", "split_i8"])
def test_contract_matches_oracle(rt, rule_name, pairs, ci, prec):
    axa, axb, perm, nout = CONTRACTS[ci]
    ps, os_ = spaces(rule_name, [(pairs, False), (pairs, False)], [(pairs, False), (pairs, False)])
    A = TensorMap.rand(rt, ps, seed=1)
    B = TensorMap.rand(rt, ps, seed=2)
    got = A.contract(axa, B, axb, perm, nout, precision=prec).to_numpy()
    want = OTensor(os_, A.to_numpy()).contract(axa, OTensor(os_, B.to_numpy()), axb, perm, nout).data
    tol = {"f64": 1e-13, "f32": 2e-6, "split_bf16": 1e-10, "split_i8": 1e-10}[prec]
    assert rel(got, want) < tol, (rule_name, ci, prec, rel(got, want))


@pytest.mark.parametrize("axb", [[0, 1], [1, 0]])
def test_quickstart_scale_compose_and_swap(rt, axb):
    """U(1) legs -2..2 x 8 (dim 40), uniform fills with seeds 1 and 2:
    A[a b; c d] . B[c d; g h] (no recoupling) and B[d c; g h] (rhs recoupled)."""
    V = [(q, 8) for q in range(-2, 3)]
    ps, os_ = spaces("U1", [(V, False), (V, False)], [(V, False), (V, False)])
    A, B = TensorMap.rand(rt, ps, seed=1), TensorMap.rand(rt, ps, seed=2)
    got = A.contract([2, 3], B, axb).to_numpy()
    want = OTensor(os_, A.to_numpy()).contract([2, 3], OTensor(os_, B.to_numpy()), axb, [0, 1, 2, 3], 2).data
    assert rel(got, want) < 1e-13


def test_contract_rejects_short_operands_before_mutating(rt):
    """Validate before mutate: an operand shorter than the artifact expects is
    an error and the destination keeps its contents."""
    V = [(q, 4) for q in range(-1, 2)]
    ps, _ = spaces("U1", [(V, False), (V, False)], [(V, False), (V, False)])
    A, B = TensorMap.rand(rt, ps, seed=1), TensorMap.rand(rt, ps, seed=2)
    c = rt.contraction(A.space, B.space, [2, 3], [1, 0], [0, 1, 2, 3], 2)
    D = TensorMap.rand(rt, c.dst_space, seed=3)
    before = D.to_numpy()
    short = _buf(rt, np.zeros(5))
    try:
        rc = N.hip().tnt_contract(rt._ctx, c.art, D._buf, A._buf, short, 1.0, 0.0, 0)
        assert rc == N.E_SHAPE
        rc = N.hip().tnt_contract(rt._ctx, c.art, short, A._buf, B._buf, 1.0, 0.0, 0)
        assert rc == N.E_SHAPE
        rc = N.hip().tnt_contract(rt._ctx, c.art, D._buf, A._buf, B._buf, 1.0, 0.0, 7)
        assert rc == N.E_INVALID
    finally:
        _free(rt, short)
    assert np.array_equal(D.to_numpy(), before)


def _oracle_fill_fn(rule, c0):
    def fn(unc, coupled, idx):
        lab = [rule.label(s) for s in unc] + [rule.label(coupled)]
        l1, l2, m1, m2, lc = lab
        w = (c0 + 7 * l1 + 11 * l2 + 13 * m1 + 17 * m2 + 19 * lc + 23 * (idx[0] + 1)
             + 29 * (idx[1] + 1) + 31 * (idx[2] + 1) + 37 * (idx[3] + 1))
        return float(w % 41 - 20)
    return fn


@pytest.mark.parametrize("name,pairs", [("U1", [(-1, 2), (0, 3), (1, 2)]), ("SU2", [(0, 2), (1, 2), (2, 1)])])
def test_tensorkit_invariant_stream_on_gpu(rt, name, pairs):
    """benchmarks/tensorkit_semantic_oracle.out section 3 (semantic_suite.rs:839-1030)."""
    rule = Rule(name)
    V = GradedSpace(rule, pairs)
    A = TensorMap.from_block_fn(rt, [V, V], [V, V], _oracle_fill_fn(rule, 3))
    B = TensorMap.from_block_fn(rt, [V, V], [V, V], _oracle_fill_fn(rule, 5))
    Cm = A.compose(B)
    D = A.permute([1, 0], [3, 2])
    E = D.compose(Cm)
    G = A.adjoint().compose(A)
    H = E.add(A, 1.0, 0.5)
    got = {n: (t.norm(), t.tr()) for n, t in [("s1a", A), ("s1b", B), ("s2", Cm), ("s3", D), ("s4", E),
                                               ("s5", G), ("s7", H)]}
    s8 = H.compose(H).tr()
    want = GOLD["stream"][name]
    for step, (nv, tv) in got.items():
        assert abs(nv - want[step][0]) <= 1e-12 * abs(want[step][0]), step
        assert abs(tv - want[step][1]) <= 1e-12 * max(1.0, abs(want[step][1])), step
    assert abs(s8 - want["s8"]) <= 1e-12 * abs(want["s8"])


@pytest.mark.parametrize("d", [4, 16])
def test_su2_microbench_checksums_on_gpu(rt, d):
    """docs/su2_authority.md:81-83 on the workloads of microbench_fusion.rs:24-50."""
    rule = Rule("SU2")
    V = GradedSpace(rule, [(0, d), (1, d), (2, d)])
    sp = HomSpace(rule, [V, V], [V, V])
    n = sp.len
    A = TensorMap.from_numpy(rt, sp, (np.arange(n) % 17) * 0.25 - 2.0)
    B = TensorMap.from_numpy(rt, sp, (np.arange(n) % 13) * 0.5 - 3.0)
    want = KAT["microbench_checksums"]["SU2"][str(d)]
    for (la, lb, perm), w in zip([([2, 3], [0, 1], [0, 1, 2, 3]), ([3, 2], [0, 1], [0, 1, 2, 3]),
                                  ([3, 2], [0, 1], [1, 0, 2, 3])], want):
        out = A.contract(la, B, lb, perm, 2).to_numpy()
        assert abs(out.sum() - w) <= 1e-13 * np.abs(out).sum() + 1e-12


def test_quickstart_inner_50(rt):
    rule = Rule("U1")
    V = GradedSpace(rule, [(-1, 1), (0, 2), (1, 1)])
    a = TensorMap.from_block_fn(rt, [V], [V], lambda u, c, i: float(2 + rule.label(c)) if i[0] == i[1] else 0.0)
    b = TensorMap.from_block_fn(rt, [V], [V], lambda u, c, i: float(2 - rule.label(c)) if i[0] == i[1] else 0.0)
    c = a.contract([1], b, [0])
    assert c.inner(c) == 50.0


def test_weighted_inner_matches_oracle(rt):
    ps, os_ = spaces("SU2", [(SU2V, False), (SU2V, False)], [(SU2V, False), (SU2V, False)])
    A, B = TensorMap.rand(rt, ps, seed=3), TensorMap.rand(rt, ps, seed=4)
    want = OTensor(os_, A.to_numpy()).inner(OTensor(os_, B.to_numpy()))
    got = A.inner(B)
    assert abs(got - want) <= 1e-12 * abs(want)
    assert A.inner(B) == got  # deterministic reduction order


def test_graph_capture_replays_match_eager(rt):
    ps, os_ = spaces("SU2", [(SU2V, False), (SU2V, False)], [(SU2V, False), (SU2V, False)])
    A, B = TensorMap.rand(rt, ps, seed=8), TensorMap.rand(rt, ps, seed=9)
    eager = A.contract([1, 3], B, [2, 0], [0, 2, 1, 3], 2)
    ref = eager.to_numpy()
    out = TensorMap.zeros(rt, eager.space)
    h = N.hip()
    rt._check(h.tnt_graph_begin(rt._ctx))
    A.contract([1, 3], B, [2, 0], [0, 2, 1, 3], 2, out=out)
    g = C.c_void_p()
    rt._check(h.tnt_graph_end(rt._ctx, C.byref(g)))
    for rep in range(4):
        out.scale_(0.0)
        rt.synchronize()
        rt._check(h.tnt_graph_launch(rt._ctx, g))
        rt.synchronize()
        assert np.array_equal(out.to_numpy(), ref), rep
    h.tnt_graph_destroy(rt._ctx, g)


@pytest.mark.parametrize("prec", [2, 3])
@pytest.mark.parametrize("slices,tol", [(4, 1e-7), (5, 1e-10), (6, 1e-12)])
def test_split_error_vs_slices(rt, prec, slices, tol):
    """Relative Frobenius error of the split GEMMs (BF16 and INT8 digit planes)
    against FP64, per slice count."""
    rng = np.random.default_rng(slices)
    m, k, n = 200, 700, 150
    A = rng.standard_normal((m, k)) * np.exp(rng.uniform(-3, 3, (m, 1)))  # uneven row scales
    B = rng.standard_normal((k, n))
    bl, br = _buf(rt, A.reshape(-1, order="F")), _buf(rt, B.reshape(-1, order="F"))
    bd = _buf(rt, np.zeros(m * n))
    jobs = (N.GemmJob * 1)(N.GemmJob(0, 0, 0, m, k, n))
    rt._check(N.hip().tnt_set_split_slices(rt._ctx, slices))
    try:
        rt._check(N.hip().tnt_gemm_batch(rt._ctx, bd, bl, br, jobs, 1, None, 0, 0, 0, 1.0, 0.0, prec))
        got = _get(rt, bd, m * n).reshape(n, m).T
    finally:
        N.hip().tnt_set_split_slices(rt._ctx, 5)
        _free(rt, bl, br, bd)
    err = rel(got, A @ B)
    print(f"split prec={prec} slices={slices}: rel err {err:.3e}")
    assert err < tol


def _carry_digit_matrix(rng, rows, cols, digits=5):
    """Rows x cols with entries 2^-7 (d_0 + sum_{p=1..digits-1} d_p 2^-8p),
    |d_0| <= 120 and the lower digits drawn from {-128, -127, 0, 1, 127, 128},
    column 0 = 0.75 (every row's maximum is in [0.5, 126.5 / 128), exponent 0).
    `digits` INT8 digit planes hold such entries exactly; with five, their
    round-to-nearest digits often reach +-128, the case that needs a carry
    into the digit above.  With more than one digit, every fourth row instead
    has maximum 1 - 2^-20 (in the top of its binade: the line exponent goes up
    by one, the first digit shrinks to [-64, 64]) and its last digit even, so
    it stays exact."""
    val = rng.integers(-120, 121, (rows, cols)).astype(np.float64) * 2.0 ** -7
    for p in range(1, digits):
        d = rng.choice([-128, -127, 0, 1, 127, 128], size=(rows, cols))
        if p == digits - 1:
            d[::4] &= ~1
        val += d * 2.0 ** (-7 - 8 * p)
    val[:, 0] = 0.75
    if digits > 1:
        val[::4, 0] = 1.0 - 2.0 ** -20
    return val


@pytest.mark.parametrize("carry_side", ["lhs", "rhs"])
@pytest.mark.parametrize("lop,rop", [(0, 0), (1, 1)])
def test_split_i8_digits_on_the_carry_boundary_are_exact(rt, carry_side, lop, rop):
    """One operand exact in five INT8 planes with digits on the carry
    boundary, the other exact in its first plane (multiples of 2^-7): every
    nonzero digit product then has level p + q <= 4 and is kept, so the
    split-INT8 GEMM has no truncation error and matches the FP64 product to
    rounding.  A digit of 128 stored without its carry would be off by ~2^-7."""
    rng = np.random.default_rng(31 + lop + (2 if carry_side == "rhs" else 0))
    m, k, n = 70, 300, 50
    A = _carry_digit_matrix(rng, m, k, 5 if carry_side == "lhs" else 1)
    B = _carry_digit_matrix(rng, n, k, 5 if carry_side == "rhs" else 1).T  # column maxima as above
    sa, sb = (A.T if lop else A), (B.T if rop else B)
    bl, br = _buf(rt, sa.reshape(-1, order="F")), _buf(rt, sb.reshape(-1, order="F"))
    bd = _buf(rt, np.zeros(m * n))
    jobs = (N.GemmJob * 1)(N.GemmJob(0, 0, 0, m, k, n))
    try:
        rt._check(N.hip().tnt_gemm_batch(rt._ctx, bd, bl, br, jobs, 1, None, 0, lop, rop, 1.0, 0.0, 3))
        got = _get(rt, bd, m * n).reshape(n, m).T
    finally:
        _free(rt, bl, br, bd)
    assert rel(got, A @ B) < 1e-14, rel(got, A @ B)


@pytest.mark.parametrize("prec", [2, 3])
def test_split_line_exponents_absorb_extreme_scales(rt, prec):
    """Rows of A and columns of B scaled by 10^u, u uniform in [-100, 100],
    one all-zero row and column: the line exponents absorb the scales, so
    every C_ij is as accurate relative to ||a_i|| ||b_j|| as without them."""
    rng = np.random.default_rng(8 + prec)
    m, k, n = 90, 700, 70
    ra, cb = 10.0 ** rng.uniform(-100, 100, m), 10.0 ** rng.uniform(-100, 100, n)
    A = rng.standard_normal((m, k)) * ra[:, None]
    B = rng.standard_normal((k, n)) * cb[None, :]
    A[7] = 0.0
    B[:, 5] = 0.0
    bl, br = _buf(rt, A.reshape(-1, order="F")), _buf(rt, B.reshape(-1, order="F"))
    bd = _buf(rt, np.zeros(m * n))
    jobs = (N.GemmJob * 1)(N.GemmJob(0, 0, 0, m, k, n))
    try:
        rt._check(N.hip().tnt_gemm_batch(rt._ctx, bd, bl, br, jobs, 1, None, 0, 0, 0, 1.0, 0.0, prec))
        got = _get(rt, bd, m * n).reshape(n, m).T
    finally:
        _free(rt, bl, br, bd)
    assert np.all(np.isfinite(got))
    assert not got[7].any() and not got[:, 5].any()
    scale = np.outer(np.linalg.norm(A, axis=1), np.linalg.norm(B, axis=0))
    scale[scale == 0] = 1.0
    assert np.max(np.abs(got - A @ B) / scale) < 1e-11


def test_split_i8_long_k_spans_two_exact_windows(rt):
    """K = 16684 > 16384: the INT32 level sums are flushed after the first
    window and the second window adds into D (alpha, beta != 0)."""
    rng = np.random.default_rng(5)
    m, k, n = 40, 16684, 24
    A, B, D0 = rng.standard_normal((m, k)), rng.standard_normal((k, n)), rng.standard_normal((m, n))
    bl, br = _buf(rt, A.reshape(-1, order="F")), _buf(rt, B.reshape(-1, order="F"))
    bd = _buf(rt, D0.reshape(-1, order="F"))
    jobs = (N.GemmJob * 1)(N.GemmJob(0, 0, 0, m, k, n))
    try:
        rt._check(N.hip().tnt_gemm_batch(rt._ctx, bd, bl, br, jobs, 1, None, 0, 0, 0, 0.75, -0.5, 3))
        got = _get(rt, bd, m * n).reshape(n, m).T
    finally:
        _free(rt, bl, br, bd)
    assert rel(got, 0.75 * A @ B - 0.5 * D0) < 1e-10


@pytest.mark.parametrize("prec", [2, 3])
def test_split_tile_shapes_match_numpy(rt, prec):
    """Jobs sized so that the split GEMMs use each of their four tile shapes
    (128x128 on 8 waves, 128x64, 64x128, 64x64: K >= 1024 and enough tiles),
    accumulating into a nonzero D."""
    rng = np.random.default_rng(40 + prec)
    shapes = [(1100, 1100, 1100), (1050, 1030, 1000), (600, 1030, 200), (200, 1030, 600), (100, 1030, 90)]
    jobs, lhs, rhs, want = [], [], [], []
    lo = ro = do = 0
    d_parts = []
    for m, k, n in shapes:
        a, b, d0 = rng.standard_normal((m, k)), rng.standard_normal((k, n)), rng.standard_normal((m, n))
        lhs.append(a.reshape(-1, order="F"))
        rhs.append(b.reshape(-1, order="F"))
        d_parts.append(d0.reshape(-1, order="F"))
        want.append((1.5 * a @ b + 0.25 * d0).reshape(-1, order="F"))
        jobs.append(N.GemmJob(do, lo, ro, m, k, n))
        lo, ro, do = lo + m * k, ro + k * n, do + m * n
    L, R, D = _buf(rt, np.concatenate(lhs)), _buf(rt, np.concatenate(rhs)), _buf(rt, np.concatenate(d_parts))
    try:
        arr = (N.GemmJob * len(jobs))(*jobs)
        rt._check(N.hip().tnt_gemm_batch(rt._ctx, D, L, R, arr, len(jobs), None, 0, 0, 0, 1.5, 0.25, prec))
        got = _get(rt, D, do)
    finally:
        _free(rt, L, R, D)
    o = 0
    for (m, k, n), part in zip(shapes, want):
        assert rel(got[o:o + m * n], part) < 1e-10, ((m, k, n), prec)
        o += m * n


@pytest.mark.parametrize("prec,tol", [("f64", 1e-13), ("f32", 1e-5), ("split_bf16", 1e-10), ("split_i8", 1e-10)])
def test_batched_grouped_launch_matches_oracle(rt, prec, tol):
    """Batched U(1)xZ2 merges at small sizes: every merge of one grouped launch vs the FP64 oracle."""
    from oracle.fusion import HomSpace as OH, Leg as OL
    from tenet_amd.workloads import BatchedMerge, batched_specs
    count, dmin, dmax = 6, 32, 64
    wl = BatchedMerge(rt, count, dmin, dmax, seed=11, precision=prec)
    wl.run()
    rt.synchronize()
    got = wl.results()
    orule = make_rule("U1xZ2")
    oP = OL([(orule.encode((0, 0)), 1), (orule.encode((1, 1)), 1)])
    for i, ((pl, pm, pr), g) in enumerate(zip(batched_specs(count, dmin, dmax, 11), got)):
        Vl, Vm, Vr = (OL([(orule.encode(l), d) for l, d in x]) for x in (pl, pm, pr))
        a = OTensor(OH(orule, [Vl, oP], [Vm]), wl.pairs[i][0].to_numpy())
        b = OTensor(OH(orule, [Vm], [oP, Vr]), wl.pairs[i][1].to_numpy())
        assert rel(g, a.compose(b).data) < tol, (i, prec)
    # hipGraph replay of the same grouped launch reproduces the eager result
    wl.capture()
    for rep in range(3):
        for _, _, Cm, _ in wl.pairs:
            Cm.scale_(0.0)
        rt.synchronize()
        wl.run()
        rt.synchronize()
        assert all(np.array_equal(x, y) for x, y in zip(wl.results(), got)), rep
