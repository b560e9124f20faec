"""Host planner of the fused narrow-contraction route.

A narrow contraction (every core job k <= 32, n <= 16) runs on the device as
one recoupling pass lhs -> dst whose Multi coefficients are gathered from the
rhs payload.  Checked on the CPU: the planned transform, replayed by the
oracle's replay with U[t] = sum of scale[p] * rhs[idx[p]] over the gather
terms of coefficient t, equals the general route
(lhs transform, per-job GEMM, output transform, replayed the same way) and the
oracle contraction; ineligible contractions are refused with the failed
condition named.
"""
import ctypes as C

import numpy as np
import pytest

from descs import contract_desc_replay, ospace, replay_desc
from oracle.tensor import OTensor
from tenet_amd import _native as N
from tenet_amd.tensor import HomSpace, _HostContraction
from tenet_amd.workloads import MPSEnvUpdate, mps_env_spaces


def fused(cdesc):
    h = N.host()
    f = C.c_void_p()
    rc = h.tnth_narrow_fusion(cdesc, C.byref(f))
    if rc != N.OK:
        return None, h.tnth_last_error().decode()
    return f, None


def gather_terms(f):
    """(ptr, idx, scale) of the plan's coefficient gather: coefficient t is
    the sum over ptr[t] <= p < ptr[t+1] of scale[p] * rhs[idx[p]]."""
    h = N.host()
    ptr, idx, scale = N.I64P(), N.I64P(), C.POINTER(C.c_double)()
    n = h.tnth_fusion_gather(f, C.byref(ptr), C.byref(idx), C.byref(scale))
    pt = np.ctypeslib.as_array(ptr, (n + 1,)).copy() if n else np.zeros(1, np.int64)
    nt = int(pt[-1])
    ix = np.ctypeslib.as_array(idx, (nt,)).copy() if nt else np.zeros(0, np.int64)
    sc = np.ctypeslib.as_array(scale, (nt,)).copy() if nt else np.zeros(0)
    return pt, ix, sc


def fused_replay(f, lhs, rhs):
    pt, ix, sc = gather_terms(f)
    U = [sum(sc[p] * rhs[ix[p]] for p in range(pt[t], pt[t + 1])) for t in range(len(pt) - 1)]
    return replay_desc(N.host().tnth_fusion_desc(f), lhs, coeffs=U), ix, sc


def env_spaces(rule, D):
    r, V, P, M = mps_env_spaces(rule, D)
    return r, HomSpace(r, [V, M], [V]), HomSpace(r, [V], [P, V]), HomSpace(r, [M, P], [P, M])


@pytest.mark.parametrize("out_perm", [[0, 1, 2, 3], [0, 3, 1, 2], [3, 0, 2, 1]])
@pytest.mark.parametrize("nout", [2, 1])
def test_env_mpo_step_fuses_and_matches(out_perm, nout):
    r, L, A, W = env_spaces("U1", 48)
    T1 = _HostContraction(L, A, [2], [0], [0, 1, 2, 3], 2, False, False).dst_space
    c = _HostContraction(T1, W, [1, 2], [3, 1], out_perm, nout, False, False)
    f, why = fused(c.desc)
    assert f is not None, why
    try:
        rng = np.random.default_rng(1)
        lhs = rng.normal(size=T1.len)
        rhs = rng.normal(size=W.len)
        got, ix, sc = fused_replay(f, lhs, rhs)
        want = contract_desc_replay(c.desc, lhs, rhs, c.desc.contents.core_len)
        assert got.shape == want.shape == (c.dst_space.len,)
        assert np.max(np.abs(got - want)) <= 1e-13 * np.max(np.abs(want))
        # U(1) recoupling coefficients are 1: every gathered scale is 1 or a structural 0
        assert set(np.unique(sc)) <= {0.0, 1.0}
        assert np.all((ix >= 0) == (sc != 0))
        ref = OTensor(ospace(T1), lhs).contract([1, 2], OTensor(ospace(W), rhs), [3, 1], out_perm, nout)
        assert np.max(np.abs(got - ref.data)) <= 1e-13 * np.max(np.abs(ref.data))
    finally:
        N.host().tnth_fusion_destroy(f)


def test_heff_mpo_steps_fuse():
    r, V, P, M = mps_env_spaces("U1", 40)
    L = HomSpace(r, [V, M], [V])
    th = HomSpace(r, [V], [P, P, V])
    W = HomSpace(r, [M, P], [P, M])
    x1 = _HostContraction(L, th, [2], [0], [0, 1, 2, 3, 4], 2, False, False)
    x2 = _HostContraction(x1.dst_space, W, [1, 2], [3, 1], [0, 3, 4, 1, 2], 2, False, False)
    x3 = _HostContraction(x2.dst_space, W, [1, 3], [3, 1], [0, 3, 1, 4, 2], 2, False, False)
    rng = np.random.default_rng(2)
    for c, lhs_space in ((x2, x1.dst_space), (x3, x2.dst_space)):
        f, why = fused(c.desc)
        assert f is not None, why
        try:
            lhs, rhs = rng.normal(size=lhs_space.len), rng.normal(size=W.len)
            got, _, _ = fused_replay(f, lhs, rhs)
            want = contract_desc_replay(c.desc, lhs, rhs, c.desc.contents.core_len)
            assert np.max(np.abs(got - want)) <= 1e-13 * np.max(np.abs(want))
        finally:
            N.host().tnth_fusion_destroy(f)


def test_wide_contractions_keep_the_general_route():
    r, L, A, W = env_spaces("U1", 48)
    c1 = _HostContraction(L, A, [2], [0], [0, 1, 2, 3], 2, False, False)
    f, why = fused(c1.desc)
    assert f is None and "plain GEMM" in why
    # A^dagger . T2 contracts k up to hundreds: general route
    T1 = c1.dst_space
    T2 = _HostContraction(T1, W, *MPSEnvUpdate.T2_AXES, False, False).dst_space
    c3 = _HostContraction(A, T2, *MPSEnvUpdate.LP_AXES, True, False)
    f, why = fused(c3.desc)
    assert f is None and "not narrow" in why


@pytest.mark.parametrize("count,D", [("dimension", 32), ("dimension", 64), ("multiplets", 12)])
def test_su2_mpo_step_fuses_and_matches(count, D):
    """The SU(2) MPO step bends legs across the codomain/domain boundary, so
    its lhs, rhs and output transforms have Multi blocks (2 x 2, mixing row
    trees of both core jobs).  It fuses into groups that are connected
    components of row trees under those blocks, with U entries that are sums
    of rhs terms; the pass equals the general route and the oracle."""
    r, V, P, M = mps_env_spaces("SU2", D, count)
    L, A, W = HomSpace(r, [V, M], [V]), HomSpace(r, [V], [P, V]), HomSpace(r, [M, P], [P, M])
    T1 = _HostContraction(L, A, [2], [0], [0, 1, 2, 3], 2, False, False).dst_space
    c = _HostContraction(T1, W, [1, 2], [3, 1], [0, 3, 1, 2], 2, False, False)
    d = c.desc.contents
    assert any(d.lhs.contents.blocks[i].kind == 1 for i in range(d.lhs.contents.n_blocks))
    assert any(d.out.contents.blocks[i].kind == 1 for i in range(d.out.contents.n_blocks))
    f, why = fused(c.desc)
    assert f is not None, why
    try:
        fd = N.host().tnth_fusion_desc(f).contents
        assert max(fd.blocks[i].src_count for i in range(fd.n_blocks)) > 1 < max(
            fd.blocks[i].dst_count for i in range(fd.n_blocks))
        pt, _, _ = gather_terms(f)
        assert np.any(np.diff(pt) > 1)  # some coefficients are sums of rhs terms
        rng = np.random.default_rng(3)
        lhs, rhs = rng.normal(size=T1.len), rng.normal(size=W.len)
        got, _, _ = fused_replay(f, lhs, rhs)
        want = contract_desc_replay(c.desc, lhs, rhs, c.desc.contents.core_len)
        assert np.max(np.abs(got - want)) <= 1e-13 * np.max(np.abs(want))
        ref = OTensor(ospace(T1), lhs).contract([1, 2], OTensor(ospace(W), rhs), [3, 1], [0, 3, 1, 2], 2)
        assert np.max(np.abs(got - ref.data)) <= 1e-13 * np.max(np.abs(ref.data))
    finally:
        N.host().tnth_fusion_destroy(f)


@pytest.mark.parametrize("out_perm,nout", [([0, 2, 1], 2), ([2, 0, 1], 1)])
def test_adjoint_mpo_operand_fuses(out_perm, nout):
    """X[a; s, w] . W^dagger contracted in W^dagger's own leg order: the core
    reads the stored W with op = adjoint, and the gather indexes it transposed."""
    r, V, P, M = mps_env_spaces("U1", 48)
    X, W = HomSpace(r, [V], [P, M]), HomSpace(r, [M, P], [P, M])
    c = _HostContraction(X, W, [1, 2], [0, 1], out_perm, nout, False, True)
    assert c.desc.contents.rhs_op == N.OP_ADJOINT
    f, why = fused(c.desc)
    assert f is not None, why
    try:
        rng = np.random.default_rng(4)
        lhs, rhs = rng.normal(size=X.len), rng.normal(size=W.len)
        got, _, _ = fused_replay(f, lhs, rhs)
        want = contract_desc_replay(c.desc, lhs, rhs, c.desc.contents.core_len)
        ref = OTensor(ospace(X), lhs).contract([1, 2], OTensor(ospace(W), rhs).adjoint(), [0, 1], out_perm, nout)
        assert np.max(np.abs(got - want)) <= 1e-14 * np.max(np.abs(want))
        assert np.max(np.abs(got - ref.data)) <= 1e-14 * np.max(np.abs(ref.data))
    finally:
        N.host().tnth_fusion_destroy(f)
