"""Uncovered core regions on the general route (transforms, grouped GEMM).

Coupled sectors of the core result that no GEMM job writes get dst = beta dst
(zero for beta = 0).  Without an output transform they run as GEMM jobs with
nothing contracted, in the launch of the contraction's own tiles; with one
they are regions of the core scratch, zeroed before the GEMM.  Both against
the oracle for beta = 0 and beta != 0 in FP64, FP32 and the split precisions,
on every forced tile shape, the region elements exactly beta * dst, and under
the shadow write count (every core element written once).
"""
import ctypes as C
import os

import numpy as np
import pytest

from descs import ospace
from oracle.tensor import OTensor
from tenet_amd import _native as N
from tenet_amd.tensor import GradedSpace, Rule, Runtime, TensorMap

pytestmark = pytest.mark.gpu

TOL = {"f64": 1e-13, "f32": 1e-5, "split_bf16": 1e-10, "split_i8": 1e-10}


@pytest.fixture(scope="module")
def rt():
    r = Runtime(0)
    yield r
    r.close()


def _operands(rt):
    r = Rule("U1")
    Va = GradedSpace(r, [(-1, 40), (0, 23)])
    Vd = GradedSpace(r, [(0, 37), (1, 19)])
    A = TensorMap.rand(rt, [Va, Va], [Va, Vd], seed=5, normal=True)
    B = TensorMap.rand(rt, [Va, Vd], [Va, Va], seed=6, normal=True)
    return A, B


def _general(rt, A, ax, B, bx, perm):
    os.environ["TENET_AMD_TINY_FLOPS"] = "0"
    try:
        return rt.contraction(A.space, B.space, ax, bx, perm, 2)
    finally:
        del os.environ["TENET_AMD_TINY_FLOPS"]


def _regions(c):
    d = c.desc.contents
    return [(d.uncovered[i].offset, d.uncovered[i].len) for i in range(d.n_uncovered)]


@pytest.mark.parametrize("ax,bx,perm,has_out", [([2, 3], [0, 1], [0, 1, 2, 3], False),
                                                ([3, 2], [1, 0], [1, 0, 2, 3], True)])
def test_uncovered_regions_on_the_general_route(rt, ax, bx, perm, has_out):
    A, B = _operands(rt)
    c = _general(rt, A, ax, B, bx, perm)
    assert c.route == "transform_gemm"
    assert bool(c.desc.contents.out) == has_out
    regions = _regions(c)
    assert regions and sum(n for _, n in regions) > 0
    want = OTensor(ospace(A.space), A.to_numpy()).contract(ax, OTensor(ospace(B.space), B.to_numpy()),
                                                           bx, perm, 2).data
    d0 = np.random.default_rng(7).normal(size=want.size)
    h = N.hip()
    try:
        for cfg in (-1, 0, 1, 2, 3, 4, 5, 6):
            rt._check(h.tnt_contract_set_tile_config(rt._ctx, c.art, cfg))
            for prec in ("f64", "f32", "split_i8", "split_bf16"):
                if cfg >= 0 and prec != "f64":
                    continue  # forced shapes: the FP64 table
                for beta in (0.0, 2.0):
                    D = TensorMap.from_numpy(rt, c.dst_space, d0)
                    A.contract(ax, B, bx, perm, 2, alpha=1.25, beta=beta, out=D, precision=prec)
                    got = D.to_numpy()
                    ref = 1.25 * want + beta * d0
                    err = np.abs(got - ref).max() / np.abs(ref).max()
                    assert err <= TOL[prec], (cfg, prec, beta, err)
                    if not has_out:  # the core is the destination: regions exactly beta * dst
                        for off, n in regions:
                            assert np.array_equal(got[off:off + n], beta * d0[off:off + n]), (cfg, prec, beta)
    finally:
        rt._check(h.tnt_contract_set_tile_config(rt._ctx, c.art, -1))


def test_uncovered_region_jobs_write_each_core_element_once():
    r = Runtime(0)
    try:
        r._check(N.hip().tnt_ctx_set_shadow(r._ctx, 1))
        A, B = _operands(r)
        c = _general(r, A, [2, 3], B, [0, 1], [0, 1, 2, 3])
        assert c.route == "transform_gemm" and _regions(c)
        for prec in ("f64", "f32"):
            A.contract([2, 3], B, [0, 1], [0, 1, 2, 3], 2, precision=prec)
        w, e, d, p = C.c_uint64(), C.c_uint64(), C.c_uint64(), C.c_uint64()
        r._check(N.hip().tnt_ctx_shadow_stats(r._ctx, C.byref(w), C.byref(e), C.byref(d), C.byref(p)))
        assert p.value >= 2 and w.value > 0 and d.value == 0 and w.value == e.value, (w, e, d, p)
        r._check(N.hip().tnt_ctx_set_shadow(r._ctx, 0))
    finally:
        r.close()
