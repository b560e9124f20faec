"""The gathered small-GEMM route of tiny contractions against the general route.

A tiny contraction (Single-only transforms, at most 16 Mi flops and 2^20
elements per payload) runs its FP64 calls as one launch: operand transforms
folded into the GEMM's operand loads, the output transform into its stores,
uncovered core regions and their destinations handled as fix-ups
(tnt_contract_upload's TNT_ROUTE_TINY_GATHERED).  The same descriptors
uploaded with TENET_AMD_TINY_FLOPS=0 take the general route (transform, GEMM,
transform).  Both must agree -- the route forms every product and sum in the
same order as the general route's 32 x 32 tiles, so the bar is one rounding
per element -- and both must match the oracle; alpha / beta, adjoint operands,
uncovered regions and the split precisions (which keep the general route) are
covered.
"""
import os

import numpy as np
import pytest

from descs import ospace
from oracle.tensor import OTensor
from tenet_amd.tensor import GradedSpace, HomSpace, Rule, Runtime, TensorMap

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rts():
    tiny = Runtime(0)
    gen = Runtime(0)
    yield tiny, gen
    tiny.close()
    gen.close()


def _general(rt, A, ax, B, bx, perm, nout):
    """Upload the contraction on `rt` with the tiny route off."""
    os.environ["TENET_AMD_TINY_FLOPS"] = "0"
    try:
        rt.contraction(A.space, B.space, ax, bx, perm, nout)
    finally:
        del os.environ["TENET_AMD_TINY_FLOPS"]


def _copy(rt, t):
    return TensorMap.from_numpy(rt, t.space, t.to_numpy())


CELLS = [("compose", [2, 3], [0, 1], [0, 1, 2, 3]), ("swap", [3, 2], [0, 1], [0, 1, 2, 3]),
         ("swap+out", [3, 2], [0, 1], [1, 0, 2, 3]), ("cross", [1, 3], [2, 0], [0, 2, 1, 3])]


@pytest.mark.parametrize("sym,pairs", [("U1", [(-1, 4), (0, 4), (1, 4)]), ("SU2", [(0, 4), (1, 4), (2, 4)]),
                                       ("U1", [(-2, 1), (0, 3), (1, 2)])])
@pytest.mark.parametrize("cell", CELLS, ids=[c[0] for c in CELLS])
def test_tiny_route_matches_general_route_and_oracle(rts, sym, pairs, cell):
    rt, rg = rts
    name, ax, bx, perm = cell
    rule = Rule(sym)
    V = GradedSpace(rule, pairs)
    sp = HomSpace(rule, [V, V], [V, V])
    A = TensorMap.rand(rt, [V, V], [V, V], seed=11)
    B = TensorMap.rand(rt, [V, V], [V, V], seed=12)
    c = rt.contraction(A.space, B.space, ax, bx, perm, 2)
    Ag, Bg = _copy(rg, A), _copy(rg, B)
    _general(rg, Ag, ax, Bg, bx, perm, 2)
    cg = rg.contraction(Ag.space, Bg.space, ax, bx, perm, 2)
    assert cg.route in ("transform_gemm", "fused_narrow")
    if cg.route == "fused_narrow":
        assert c.route == "fused_narrow"  # the narrow pass takes precedence
    elif rule.name == "SU2" and name == "cross":
        assert c.route == "transform_gemm"  # Multi groups: not a tiny route
    else:
        assert c.route == "tiny_gathered", name
    want = OTensor(ospace(sp), A.to_numpy()).contract(ax, OTensor(ospace(sp), B.to_numpy()), bx, perm, 2).data
    got = A.contract(ax, B, bx, perm, 2).to_numpy()
    gen = Ag.contract(ax, Bg, bx, perm, 2).to_numpy()
    scale = np.abs(want).max()
    assert np.abs(got - gen).max() <= 1e-14 * scale
    assert np.abs(got - want).max() <= 1e-13 * scale
    # alpha / beta on a destination with data, both routes from the same start
    d0 = np.random.default_rng(3).normal(size=got.size)
    D = TensorMap.from_numpy(rt, c.dst_space, d0)
    Dg = TensorMap.from_numpy(rg, cg.dst_space, d0)
    A.contract(ax, B, bx, perm, 2, alpha=0.75, beta=-1.5, out=D)
    Ag.contract(ax, Bg, bx, perm, 2, alpha=0.75, beta=-1.5, out=Dg)
    ref = 0.75 * want - 1.5 * d0
    assert np.abs(D.to_numpy() - ref).max() <= 1e-13 * np.abs(ref).max()
    assert np.abs(D.to_numpy() - Dg.to_numpy()).max() <= 1e-14 * np.abs(ref).max()
    # the split precisions keep the general route
    s8 = A.contract(ax, B, bx, perm, 2, precision="split_i8").to_numpy()
    assert np.linalg.norm(s8 - want) <= 1e-10 * np.linalg.norm(want)


def test_tiny_route_with_uncovered_regions(rts):
    """Coupled sectors of the output that no job writes: zero for beta = 0,
    beta * dst otherwise, with and without an output transform."""
    rt, rg = rts
    r = Rule("U1")
    Va = GradedSpace(r, [(-1, 2), (0, 1)])
    Vd = GradedSpace(r, [(0, 2), (1, 1)])
    A = TensorMap.rand(rt, [Va, Va], [Va, Vd], seed=5)
    B = TensorMap.rand(rt, [Va, Vd], [Va, Va], seed=6)
    for ax, bx, perm in (([2, 3], [0, 1], [0, 1, 2, 3]), ([3, 2], [1, 0], [1, 0, 2, 3])):
        c = rt.contraction(A.space, B.space, ax, bx, perm, 2)
        assert c.route == "tiny_gathered"
        assert c.desc.contents.n_uncovered > 0
        Ag, Bg = _copy(rg, A), _copy(rg, B)
        _general(rg, Ag, ax, Bg, bx, perm, 2)
        want = OTensor(ospace(A.space), A.to_numpy()).contract(ax, OTensor(ospace(B.space), B.to_numpy()),
                                                               bx, perm, 2).data
        for beta in (0.0, 2.0):
            d0 = np.random.default_rng(7).normal(size=want.size)
            D = TensorMap.from_numpy(rt, c.dst_space, d0)
            Dg = TensorMap.from_numpy(rg, c.dst_space, d0)
            A.contract(ax, B, bx, perm, 2, alpha=1.25, beta=beta, out=D)
            Ag.contract(ax, Bg, bx, perm, 2, alpha=1.25, beta=beta, out=Dg)
            ref = 1.25 * want + beta * d0
            assert np.abs(D.to_numpy() - ref).max() <= 1e-13 * np.abs(ref).max()
            assert np.abs(D.to_numpy() - Dg.to_numpy()).max() <= 1e-14 * np.abs(ref).max()


def test_tiny_route_adjoint_operand(rts):
    rt, rg = rts
    r = Rule("SU2")
    V = GradedSpace(r, [(0, 3), (1, 2), (2, 2)])
    A = TensorMap.rand(rt, [V, V], [V, V], seed=21)
    B = TensorMap.rand(rt, [V, V], [V, V], seed=22)
    got = A.adjoint().contract([2, 3], B, [0, 1], [0, 1, 2, 3], 2)
    c = rt.contraction(A.space, B.space, [2, 3], [0, 1], [0, 1, 2, 3], 2, True, False)
    assert c.route == "tiny_gathered"
    want = OTensor(ospace(A.space), A.to_numpy()).adjoint().contract(
        [2, 3], OTensor(ospace(B.space), B.to_numpy()), [0, 1], [0, 1, 2, 3], 2).data
    assert np.abs(got.to_numpy() - want).max() <= 1e-13 * np.abs(want).max()


def test_tiny_route_replays_in_a_graph(rts):
    """The one-launch call captured in a hipGraph: every replay equals eager."""
    import ctypes as C

    from tenet_amd import _native as N
    rt, _ = rts
    r = Rule("U1")
    V = GradedSpace(r, [(-1, 4), (0, 4), (1, 4)])
    A = TensorMap.rand(rt, [V, V], [V, V], seed=31)
    B = TensorMap.rand(rt, [V, V], [V, V], seed=32)
    out = A.contract([3, 2], B, [0, 1], [1, 0, 2, 3], 2)
    eager = out.to_numpy()
    h = N.hip()
    rt._check(h.tnt_graph_begin(rt._ctx))
    try:
        A.contract([3, 2], B, [0, 1], [1, 0, 2, 3], 2, out=out)
    finally:
        g = C.c_void_p()
        rt._check(h.tnt_graph_end(rt._ctx, C.byref(g)))
    try:
        for _ in range(3):
            out.scale_(0.0)
            rt.synchronize()
            rt._check(h.tnt_graph_launch(rt._ctx, g))
            rt.synchronize()
            assert np.array_equal(out.to_numpy(), eager)
    finally:
        rt._check(h.tnt_graph_destroy(rt._ctx, g))
