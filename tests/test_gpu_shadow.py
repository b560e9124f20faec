"""Shadow write counts (tnt_ctx_set_shadow) over the work decompositions.

With the debug option on, every recoupling replay, FP64 GEMM table run (with
the uncovered core regions) and tiny-route call is followed by a pass that
counts the writes its decomposition makes to each destination element (plan
items, GEMM tiles, regions, output-map stores).  Every element must be
written exactly once: the counted writes equal the writes the compiled plans
prescribe and no element is counted twice.  Covered: U(1) and SU(2)
permutations (Single, Multi, transposed and flattened items, several item
sizes), contractions on the general route with and without an output
transform, the fused narrow route, the tiny route, lazy adjoints, uncovered
regions, and the benchmark workloads at reduced size.  The results with the
option on equal those without it bit for bit.
"""
import ctypes as C

import numpy as np
import pytest

from tenet_amd import _native as N
from tenet_amd.tensor import GradedSpace, HomSpace, Rule, Runtime, TensorMap
from tenet_amd.workloads import MPSEnvUpdate, SU2FMoveContract, SU2FMoveSwapContract

pytestmark = pytest.mark.gpu


def _stats(rt):
    w, e, d, p = C.c_uint64(), C.c_uint64(), C.c_uint64(), C.c_uint64()
    rt._check(N.hip().tnt_ctx_shadow_stats(rt._ctx, C.byref(w), C.byref(e), C.byref(d), C.byref(p)))
    return w.value, e.value, d.value, p.value


@pytest.fixture()
def rt():
    r = Runtime(0)
    r._check(N.hip().tnt_ctx_set_shadow(r._ctx, 1))
    yield r
    r._check(N.hip().tnt_ctx_set_shadow(r._ctx, 0))
    r.close()


def _check(rt, min_passes):
    w, e, d, p = _stats(rt)
    assert p >= min_passes, (w, e, d, p)
    assert w > 0 and d == 0 and w == e, (w, e, d, p)
    return p


@pytest.mark.parametrize("rule,pairs", [("U1", [(-2, 3), (-1, 5), (0, 6), (1, 4), (2, 2)]),
                                        ("SU2", [(0, 5), (1, 4), (2, 3), (3, 2)])])
def test_permutations_write_each_element_once(rt, rule, pairs):
    V = GradedSpace(Rule(rule), pairs)
    A = TensorMap.rand(rt, [V, V, V], [V], seed=3)
    X = TensorMap.rand(rt, [V, V], [V, V], seed=4)
    n = 0
    for T, p1, p2 in ((A, [0, 2, 1], [3]), (A, [3, 0], [2, 1]), (A, [], [3, 2, 1, 0]),
                      (X, [1, 0], [2, 3]), (X, [2, 0], [1, 3]), (X, [3], [0, 1, 2])):
        T.permute(p1, p2)
        n += 1
    _check(rt, n)


def test_contraction_routes_write_each_element_once(rt):
    r = Rule("U1")
    V = GradedSpace(r, [(-1, 4), (0, 4), (1, 4)])
    Va, Vd = GradedSpace(r, [(-1, 2), (0, 1)]), GradedSpace(r, [(0, 2), (1, 1)])
    A = TensorMap.rand(rt, [V, V], [V, V], seed=1)
    B = TensorMap.rand(rt, [V, V], [V, V], seed=2)
    routes = set()
    for ax, bx, perm in (([2, 3], [0, 1], [0, 1, 2, 3]), ([3, 2], [0, 1], [1, 0, 2, 3]), ([1, 3], [2, 0], [0, 2, 1, 3])):
        A.contract(ax, B, bx, perm, 2)
        routes.add(rt.contraction(A.space, B.space, ax, bx, perm, 2).route)
    P = TensorMap.rand(rt, [Va, Va], [Va, Vd], seed=5)
    Q = TensorMap.rand(rt, [Va, Vd], [Va, Va], seed=6)
    P.contract([3, 2], Q, [1, 0], [1, 0, 2, 3], 2, alpha=0.5, beta=0.0)  # uncovered core regions
    A.adjoint().contract([2, 3], B, [0, 1], [0, 1, 2, 3], 2)
    assert "tiny_gathered" in routes
    _check(rt, 5)


@pytest.mark.parametrize("rule_name,D", [("U1", 256), ("SU2", 256)])
def test_environment_update_writes_each_element_once(rt, rule_name, D):
    wl = MPSEnvUpdate(rt, rule_name, D, "f64")
    routes = {c.route for c in wl.contractions}
    out = wl.run().to_numpy()
    _check(rt, 4)
    # the same result with the option off
    rt2 = Runtime(0)
    try:
        wl2 = MPSEnvUpdate(rt2, rule_name, D, "f64")
        assert np.array_equal(wl2.run().to_numpy(), out)
        del wl2
    finally:
        rt2.close()
    assert routes


def test_fmove_workloads_write_each_element_once(rt):
    for wl in (SU2FMoveContract(rt, mult=12, twoj_max=6), SU2FMoveSwapContract(rt, mult=12, twoj_max=6)):
        wl.run()
    _check(rt, 6)


def test_option_off_counts_nothing():
    rt = Runtime(0)
    try:
        V = GradedSpace(Rule("U1"), [(0, 3), (1, 2)])
        X = TensorMap.rand(rt, [V, V], [V, V], seed=1)
        X.permute([1, 0], [2, 3])
        assert _stats(rt) == (0, 0, 0, 0)
    finally:
        rt.close()
