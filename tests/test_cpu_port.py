"""The C++/OpenMP CPU port of the host route (oracle/cpu_port.cpp), CPU only.

The port is the CPU baseline and a second checker, so it is pinned here:
* its cache-blocked DGEMM against numpy matmul for every operand op, ragged
  sizes around the 8 x 6 micro-tile and the 192 / 256 panel edges, k = 0,
  and alpha / beta;
* its descriptor replay (transforms, jobs, uncovered regions, output
  transform) against the numpy replay of the same descriptors (tests/descs.py)
  and against the oracle's own contraction, for U(1) and SU(2) plans with
  Single and Multi groups, an adjoint operand and output permutations.
"""
import numpy as np
import pytest

from descs import contract_desc_replay, ospace
from oracle import cpu_port as P
from oracle.tensor import OTensor
from tenet_amd.tensor import GradedSpace, HomSpace, Rule, _HostContraction
from tenet_amd.workloads import mps_env_spaces, su2_fmove_space


@pytest.mark.parametrize("m,k,n", [(1, 1, 1), (7, 5, 3), (8, 6, 6), (9, 257, 7), (193, 300, 191),
                                   (200, 513, 13), (64, 0, 5), (0, 4, 4)])
@pytest.mark.parametrize("lo,ro", [(0, 0), (1, 0), (0, 1), (2, 1)])
def test_gemm_matches_numpy(m, k, n, lo, ro):
    rng = np.random.default_rng(m * 1000 + k * 10 + n)
    A, B, D0 = rng.normal(size=m * k), rng.normal(size=k * n), rng.normal(size=m * n)
    Am = A.reshape(m, k, order="F") if lo == 0 else A.reshape(k, m, order="F").T
    Bm = B.reshape(k, n, order="F") if ro == 0 else B.reshape(n, k, order="F").T
    D = P.gemm_jobs([(0, 0, 0, m, k, n)], A, B, D0.copy(), lo, ro, 0.7, -0.3)
    want = (0.7 * (Am @ Bm) - 0.3 * D0.reshape(m, n, order="F")).reshape(-1, order="F")
    assert np.abs(D - want).max(initial=0.0) <= 1e-13 * max(1, k)
    D = P.gemm_jobs([(0, 0, 0, m, k, n)], A, B, np.full(m * n, np.nan), lo, ro, 1.0, 0.0)
    assert np.all(np.isfinite(D))  # beta = 0 never reads the destination


def _cases():
    out = []
    for rule in ("U1", "SU2"):
        _, V, Pp, M = mps_env_spaces(rule, 64)
        out.append((f"{rule} L.A", HomSpace(V.rule, [V, M], [V]), HomSpace(V.rule, [V], [Pp, V]),
                    [2], [0], [0, 1, 2, 3], 2, False))
        if rule == "SU2":
            _, V = su2_fmove_space(6, 4)
        else:
            V = GradedSpace(Rule("U1"), [(-1, 3), (0, 4), (1, 2)])
        X, X3 = HomSpace(V.rule, [V, V], [V, V]), HomSpace(V.rule, [V, V, V], [V])
        out += [(f"{rule} cross", X, X, [1, 3], [2, 0], [0, 1, 2, 3], 2, False),
                (f"{rule} swap+out", X, X, [3, 2], [0, 1], [1, 0, 2, 3], 2, False),
                (f"{rule} 3<-1", X3, X, [2, 3], [2, 0], [0, 1, 2, 3], 2, False),
                (f"{rule} adjoint", X, X, [2, 3], [0, 1], [0, 1, 2, 3], 2, True)]
    return out


@pytest.mark.parametrize("case", _cases(), ids=lambda c: c[0])
def test_contract_matches_descriptor_replay_and_oracle(case):
    _, A, B, ax, bx, perm, nout, adj = case
    c = _HostContraction(A, B, ax, bx, perm, nout, adj, False)
    d = c.desc.contents
    rng = np.random.default_rng(5)
    a, b = rng.normal(size=A.len), rng.normal(size=B.len)
    want = contract_desc_replay(c.desc, a, b, d.core_len)
    f = P.Contract(c.desc, c.dst_space.len)
    got = f(a, b)
    assert np.linalg.norm(got - want) <= 1e-14 * np.linalg.norm(want)
    d0 = rng.normal(size=c.dst_space.len)
    got2 = f(a, b, dst=d0.copy(), alpha=0.5, beta=-2.0)
    assert np.linalg.norm(got2 - (0.5 * want - 2.0 * d0)) <= 1e-14 * np.linalg.norm(want)
    # the oracle's own structure and replay, from the same payloads
    oa = OTensor(ospace(A), a)
    oa = oa.adjoint() if adj else oa
    ref = oa.contract(ax, OTensor(ospace(B), b), bx, perm, nout).data
    assert np.linalg.norm(got - ref) <= 1e-13 * np.linalg.norm(ref)


def test_transform_overwrite_and_beta():
    from tenet_amd import _native as N
    import ctypes as C
    _, V = su2_fmove_space(6, 4)
    sp = HomSpace(V.rule, [V, V, V], [V])
    h = N.host()
    t = C.c_void_p()
    N.check_host(h.tnth_transform_compile(sp._h, N.sz_array((0, 2, 1)), 3, N.sz_array((3,)), 1, C.byref(t)))
    try:
        desc = h.tnth_transform_desc(t)
        rng = np.random.default_rng(2)
        x = rng.normal(size=sp.len)
        y = P.transform(desc, x)
        ref = OTensor(ospace(sp), x).permute([0, 2, 1], [3]).data
        assert np.abs(y - ref).max() <= 1e-14 * np.abs(ref).max()
        d0 = rng.normal(size=len(y))
        z = P.transform(desc, x, dst=d0.copy(), alpha=2.0, beta=0.5, overwrite=False)
        assert np.abs(z - (2.0 * ref + 0.5 * d0)).max() <= 1e-13 * np.abs(ref).max()
    finally:
        h.tnth_transform_destroy(t)
