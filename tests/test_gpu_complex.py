"""GPU tests of complex payloads (DenseRead / DenseWrite::{C32, C64},
tenet-dense/src/executor.rs:228-285) at the batched GEMM seam, against numpy's
complex matmul: every MatrixOp pair (Adjoint = conjugate transpose), complex
alpha and beta, ragged and empty jobs, nothing-contracted jobs, every
precision, beta = 0 not reading the destination, graph replay, and the
rejections of complex payloads elsewhere.
"""
import ctypes as C

import numpy as np
import pytest

from tenet_amd import _native as N
from tenet_amd.tensor import Runtime

pytestmark = pytest.mark.gpu

SHAPES = [(37, 19, 23), (64, 64, 64), (5, 130, 17), (129, 8, 3), (0, 4, 5), (6, 0, 7), (9, 3, 0), (1, 1, 1)]
NP = {N.C64: np.complex128, N.C32: np.complex64}


@pytest.fixture(scope="module")
def rt():
    r = Runtime(0)
    yield r
    r.close()


def h():
    return N.hip()


def put(rt, arr, dtype):
    a = np.ascontiguousarray(arr, dtype=NP[dtype])
    b = C.c_void_p()
    rt._check(h().tnt_buf_alloc(rt._ctx, max(1, len(a)), dtype, C.byref(b)))
    if len(a):
        rt._check(h().tnt_buf_upload(rt._ctx, b, a.ctypes.data, len(a)))
    return b


def get(rt, b, n, dtype):
    out = np.empty(n, dtype=NP[dtype])
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


def op_of(stored, rows, cols, op):
    """op(X) of a column-major stored matrix; stored is rows x cols for the
    identity and cols x rows otherwise."""
    if op == N.OP_IDENTITY:
        return stored.reshape(cols, rows).T
    x = stored.reshape(rows, cols)  # row-major view of the column-major cols x rows matrix = its transpose
    return np.conj(x) if op == N.OP_ADJOINT else x


def complex_ref(shapes, lhs, rhs, d0, alpha, beta, lop, rop):
    out, lo, ro, do = [], 0, 0, 0
    for m, k, n in shapes:
        a = op_of(lhs[lo:lo + m * k].astype(np.complex128), m, k, lop)
        b = op_of(rhs[ro:ro + k * n].astype(np.complex128), k, n, rop)
        d = d0[do:do + m * n].astype(np.complex128).reshape(n, m).T
        r = alpha * (a @ b) + (beta * d if beta != 0 else 0)
        out.append(np.asarray(r).T.reshape(-1))
        lo, ro, do = lo + m * k, ro + k * n, do + m * n
    return np.concatenate(out)


def crandn(rng, n):
    return rng.normal(size=n) + 1j * rng.normal(size=n)


def run(rt, shapes, lhs, rhs, d0, dtype, lop, rop, alpha, beta, prec):
    arr, nl, nr, nd = jobs_of(shapes)
    bl, br, bd = put(rt, lhs, dtype), put(rt, rhs, dtype), put(rt, d0, dtype)
    try:
        rt._check(h().tnt_gemm_batch_complex(rt._ctx, bd, bl, br, arr, len(shapes), None, 0, lop, rop,
                                             alpha.real, alpha.imag, beta.real, beta.imag, prec))
        return get(rt, bd, nd, dtype)
    finally:
        free(rt, bl, br, bd)


def rel(got, want):
    return np.linalg.norm(got - want) / np.linalg.norm(want)


OPS = [N.OP_IDENTITY, N.OP_TRANSPOSE, N.OP_ADJOINT]


@pytest.mark.parametrize("lop", OPS)
@pytest.mark.parametrize("rop", OPS)
def test_c64_every_op_pair_matches_numpy(rt, lop, rop):
    rng = np.random.default_rng(10 * lop + rop)
    _, nl, nr, nd = jobs_of(SHAPES)
    lhs, rhs, d0 = crandn(rng, nl), crandn(rng, nr), crandn(rng, nd)
    alpha, beta = 0.75 - 1.5j, -0.25 + 2.0j
    got = run(rt, SHAPES, lhs, rhs, d0, N.C64, lop, rop, alpha, beta, N.PREC_F64)
    want = complex_ref(SHAPES, lhs, rhs, d0, alpha, beta, lop, rop)
    # FP64 products and sums: a few ulps of the row norms
    assert rel(got, want) < 1e-14


@pytest.mark.parametrize("prec,tol", [(N.PREC_F64, 1e-14), (N.PREC_F32, 2e-6), (N.PREC_SPLIT_I8, 1e-10)])
@pytest.mark.parametrize("lop", OPS)
@pytest.mark.parametrize("beta", [-0.5, 0.0])
def test_c64_left_expanded_form(rt, prec, tol, lop, beta):
    """op(B) = identity with a real beta: only op(A) is expanded (alpha folded
    in) and the real GEMM reads B and writes D in their real views."""
    rng = np.random.default_rng(300 + 7 * prec + lop)
    shapes = SHAPES + [(260, 300, 170)]
    _, nl, nr, nd = jobs_of(shapes)
    lhs, rhs, d0 = crandn(rng, nl), crandn(rng, nr), crandn(rng, nd)
    alpha = 0.5 - 1.25j
    got = run(rt, shapes, lhs, rhs, d0, N.C64, lop, N.OP_IDENTITY, alpha, complex(beta), prec)
    want = complex_ref(shapes, lhs, rhs, d0, alpha, beta, lop, N.OP_IDENTITY)
    assert rel(got, want) < tol


@pytest.mark.parametrize("prec,tol", [(N.PREC_F32, 2e-6), (N.PREC_SPLIT_BF16, 1e-10), (N.PREC_SPLIT_I8, 1e-10)])
@pytest.mark.parametrize("lop,rop", [(N.OP_IDENTITY, N.OP_IDENTITY), (N.OP_ADJOINT, N.OP_TRANSPOSE)])
def test_c64_other_precisions(rt, prec, tol, lop, rop):
    """FP32 MFMA (operands rounded to FP32) and the split schemes (bar 1e-10
    relative Frobenius, the same as for FP64 payloads) on the real planes."""
    rng = np.random.default_rng(100 + prec)
    shapes = SHAPES + [(300, 700, 260)]
    _, nl, nr, nd = jobs_of(shapes)
    lhs, rhs, d0 = crandn(rng, nl), crandn(rng, nr), crandn(rng, nd)
    alpha, beta = 1.0 + 0.5j, 0.5 - 0.5j
    got = run(rt, shapes, lhs, rhs, d0, N.C64, lop, rop, alpha, beta, prec)
    want = complex_ref(shapes, lhs, rhs, d0, alpha, beta, lop, rop)
    assert rel(got, want) < tol


@pytest.mark.parametrize("prec,tol", [(N.PREC_F64, 2e-7), (N.PREC_F32, 4e-6)])
@pytest.mark.parametrize("lop,rop", [(N.OP_IDENTITY, N.OP_IDENTITY), (N.OP_TRANSPOSE, N.OP_ADJOINT)])
def test_c32_payloads(rt, prec, tol, lop, rop):
    """C32 in, C32 out: FP64 arithmetic rounds once (~2^-24); FP32 arithmetic
    accumulates in FP32."""
    rng = np.random.default_rng(200 + prec)
    _, nl, nr, nd = jobs_of(SHAPES)
    lhs, rhs, d0 = (crandn(rng, x).astype(np.complex64) for x in (nl, nr, nd))
    alpha, beta = -1.25 + 0.5j, 0.5 + 0.25j
    got = run(rt, SHAPES, lhs, rhs, d0, N.C32, lop, rop, alpha, beta, prec)
    want = complex_ref(SHAPES, lhs, rhs, d0, alpha, beta, lop, rop)
    assert got.dtype == np.complex64
    assert rel(got.astype(np.complex128), want) < tol


def test_real_scalars_through_tnt_gemm_batch(rt):
    """tnt_gemm_batch on complex payloads = tnt_gemm_batch_complex with zero
    imaginary parts, bit for bit."""
    rng = np.random.default_rng(5)
    arr, nl, nr, nd = jobs_of(SHAPES)
    lhs, rhs, d0 = crandn(rng, nl), crandn(rng, nr), crandn(rng, nd)
    bl, br = put(rt, lhs, N.C64), put(rt, rhs, N.C64)
    outs = []
    try:
        for form in (0, 1):
            bd = put(rt, d0, N.C64)
            if form == 0:
                rt._check(h().tnt_gemm_batch(rt._ctx, bd, bl, br, arr, len(SHAPES), None, 0, 2, 0, 1.5, -0.5, 0))
            else:
                rt._check(h().tnt_gemm_batch_complex(rt._ctx, bd, bl, br, arr, len(SHAPES), None, 0, 2, 0,
                                                     1.5, 0.0, -0.5, 0.0, 0))
            outs.append(get(rt, bd, nd, N.C64))
            free(rt, bd)
    finally:
        free(rt, bl, br)
    assert np.array_equal(outs[0], outs[1])
    want = complex_ref(SHAPES, lhs, rhs, d0, 1.5, -0.5, N.OP_ADJOINT, N.OP_IDENTITY)
    assert rel(outs[0], want) < 1e-14


def test_beta_zero_does_not_read_destination(rt):
    """beta = 0: D is written, never read (NaN in D does not propagate), for
    complex and for real alpha."""
    rng = np.random.default_rng(6)
    shapes = [(33, 17, 29), (4, 0, 3)]
    _, nl, nr, nd = jobs_of(shapes)
    lhs, rhs = crandn(rng, nl), crandn(rng, nr)
    d0 = np.full(nd, np.nan + 1j * np.nan)
    for alpha in (2.0 - 1.0j, 1.0 + 0.0j):
        got = run(rt, shapes, lhs, rhs, d0, N.C64, 0, 0, alpha, 0j, N.PREC_F64)
        want = complex_ref(shapes, lhs, rhs, np.zeros(nd), alpha, 0, 0, 0)
        assert np.isfinite(got).all()
        assert rel(got[:33 * 29], want[:33 * 29]) < 1e-14
        assert not got[33 * 29:].any()  # nothing contracted, beta = 0: zeros


def test_nothing_contracted_scales_destination(rt):
    rng = np.random.default_rng(7)
    shapes = [(6, 0, 7), (3, 0, 2)]
    _, nl, nr, nd = jobs_of(shapes)
    d0 = crandn(rng, nd)
    beta = 0.5 - 2.0j
    got = run(rt, shapes, np.zeros(0), np.zeros(0), d0, N.C64, 0, 0, 1.0 + 1.0j, beta, N.PREC_F64)
    assert np.allclose(got, beta * d0, rtol=1e-15, atol=0)


def test_workspace_growth_and_shrink_keep_results(rt):
    """A bigger batch after a small one grows the workspace; a smaller one
    after that reuses it."""
    rng = np.random.default_rng(8)
    for shapes in ([(8, 8, 8)], [(250, 300, 270), (70, 3, 90)], [(12, 5, 7)]):
        _, nl, nr, nd = jobs_of(shapes)
        lhs, rhs, d0 = crandn(rng, nl), crandn(rng, nr), crandn(rng, nd)
        got = run(rt, shapes, lhs, rhs, d0, N.C64, 1, 2, 1.0 - 1.0j, 0.25j, N.PREC_F64)
        want = complex_ref(shapes, lhs, rhs, d0, 1.0 - 1.0j, 0.25j, 1, 2)
        assert rel(got, want) < 1e-14


def test_graph_replay_matches_eager(rt):
    """The complex batch captured in a hipGraph (workspace sized by an eager
    call first) replays to the eager result on every replay."""
    rng = np.random.default_rng(9)
    shapes = [(40, 24, 56), (17, 33, 9)]
    arr, nl, nr, nd = jobs_of(shapes)
    lhs, rhs, d0 = crandn(rng, nl), crandn(rng, nr), crandn(rng, nd)
    bl, br = put(rt, lhs, N.C64), put(rt, rhs, N.C64)
    bd = put(rt, d0, N.C64)
    g = C.c_void_p()
    try:
        def call():  # adjoint x transpose, alpha = 0.5 + 0.5i, beta = 0
            return h().tnt_gemm_batch_complex(rt._ctx, bd, bl, br, arr, len(shapes), None, 0, 2, 1,
                                              0.5, 0.5, 0.0, 0.0, 0)

        rt._check(call())
        eager = get(rt, bd, nd, N.C64)
        rt._check(h().tnt_graph_begin(rt._ctx))
        rc = call()
        rt._check(h().tnt_graph_end(rt._ctx, C.byref(g)))
        rt._check(rc)
        for _ in range(3):
            rt._check(h().tnt_buf_zero(rt._ctx, bd))
            rt._check(h().tnt_graph_launch(rt._ctx, g))
            rt._check(h().tnt_ctx_synchronize(rt._ctx))
            assert np.array_equal(get(rt, bd, nd, N.C64), eager)
    finally:
        if g.value:
            h().tnt_graph_destroy(rt._ctx, g)
        free(rt, bl, br, bd)
    want = complex_ref(shapes, lhs, rhs, d0, 0.5 + 0.5j, 0, 2, 1)
    assert rel(eager, want) < 1e-14


def test_roundtrip_zero_and_rejections(rt):
    for dt in (N.C32, N.C64):
        x = (np.arange(513) * 0.5 - 1j * np.arange(513)).astype(NP[dt])
        b = put(rt, x, dt)
        try:
            assert np.array_equal(get(rt, b, 513, dt), x)
            n, dtype = C.c_size_t(), C.c_int()
            rt._check(h().tnt_buf_info(b, None, C.byref(n), C.byref(dtype), None))
            assert (n.value, dtype.value) == (513, dt)
            rt._check(h().tnt_buf_zero(rt._ctx, b))
            assert not get(rt, b, 513, dt).any()
            # complex payloads are for the GEMM seam only
            out = C.c_double()
            reg = (N.Region * 1)(N.Region(0, 4))
            w = (C.c_double * 1)(1.0)
            assert h().tnt_weighted_dot(rt._ctx, b, b, reg, w, 1, C.byref(out)) == N.E_UNSUPPORTED
            assert h().tnt_buf_fill_random(rt._ctx, b, 1, 0) == N.E_UNSUPPORTED
            assert h().tnt_axpby(rt._ctx, b, b, 1.0, 0.0) == N.E_UNSUPPORTED
        finally:
            free(rt, b)
    f64, c64, c32 = C.c_void_p(), put(rt, np.ones(64), N.C64), put(rt, np.ones(64), N.C32)
    rt._check(h().tnt_buf_alloc(rt._ctx, 64, N.F64, C.byref(f64)))
    arr, *_ = jobs_of([(4, 4, 4)])
    try:
        # mixed payload types, complex scalars on real payloads, split on C32
        assert h().tnt_gemm_batch(rt._ctx, c64, c32, c64, arr, 1, None, 0, 0, 0, 1.0, 0.0, 0) == N.E_INVALID
        assert h().tnt_gemm_batch(rt._ctx, c64, f64, c64, arr, 1, None, 0, 0, 0, 1.0, 0.0, 0) == N.E_INVALID
        assert h().tnt_gemm_batch_complex(rt._ctx, f64, f64, f64, arr, 1, None, 0, 0, 0,
                                          1.0, 1.0, 0.0, 0.0, 0) == N.E_INVALID
        for prec in (N.PREC_SPLIT_BF16, N.PREC_SPLIT_I8):
            assert h().tnt_gemm_batch(rt._ctx, c32, c32, c32, arr, 1, None, 0, 0, 0, 1.0, 0.0, prec) == N.E_UNSUPPORTED
        # out of bounds in complex elements (64 complex = 128 doubles; a 9x9 job needs 81)
        big = (N.GemmJob * 1)(N.GemmJob(0, 0, 0, 9, 9, 9))
        assert h().tnt_gemm_batch(rt._ctx, c64, c64, c64, big, 1, None, 0, 0, 0, 1.0, 0.0, 0) == N.E_OOB
        assert np.array_equal(get(rt, c64, 64, N.C64), np.ones(64))
    finally:
        free(rt, f64, c64, c32)
