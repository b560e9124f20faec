"""CPU parity of the C++ host structure compiler against the oracle.

* layouts: every block's offset/shape/strides/uncoupled sectors, in order;
* recoupling: the compiled descriptor, replayed on the CPU, equals the
  oracle's permute -- bitwise for U(1) and U(1)xZ2 (pure copies), 1e-14 for SU(2);
* contraction: the compiled artifact (operand transforms, core GEMM jobs,
  output transform), replayed on the CPU, equals the oracle's contraction;
* symbols: SU(2) F/R from the C++ provider equal the oracle's (which is pinned
  to the TensorKit table) over all labels 2j <= 8.
"""
import itertools

import numpy as np
import pytest

from descs import contract_desc_replay, replay_desc, spaces
from oracle.symmetry import SU2 as OSU2
from oracle.tensor import OTensor
from tenet_amd import _native as N
from tenet_amd.tensor import GradedSpace, HomSpace, Rule, Runtime, _HostContraction  # noqa: F401

U1V = [(-1, 2), (0, 3), (1, 2)]
U1W = [(0, 1), (1, 2), (2, 2), (-2, 1)]
SU2V = [(0, 2), (1, 2), (2, 1)]
SU2W = [(0, 1), (1, 2), (3, 1)]
PZV = [((0, 0), 2), ((1, 1), 1), ((-1, 1), 2), ((1, 0), 1)]

CASES = [
    ("U1", [(U1V, False), (U1W, False)], [(U1W, False), (U1V, False)]),
    ("U1", [(U1V, False), (U1V, True), (U1W, False)], [(U1V, False)]),
    ("SU2", [(SU2V, False), (SU2V, False)], [(SU2V, False), (SU2W, False)]),
    ("SU2", [(SU2V, False), (SU2W, False), (SU2V, False)], [(SU2V, False)]),
    ("U1xZ2", [(PZV, False), (PZV, False)], [(PZV, False), (PZV, True)]),
]

PERMS4 = [([1, 0], [2, 3]), ([0, 1], [3, 2]), ([0, 2], [1, 3]), ([2, 0, 1], [3]), ([3], [0, 1, 2]),
          ([1, 3], [0, 2]), ([], [0, 1, 2, 3]), ([3, 2, 1, 0], [])]


def test_abi_symbols_exported():
    """Every function declared in include/*.h is exported by the built libraries."""
    import os
    import re
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    for hdr, lib in [("tenet_hip.h", "libtenet_hip.so"), ("tenet_host.h", "libtenet_host.so")]:
        text = open(os.path.join(root, "include", hdr)).read()
        names = set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\**\s*(tnth?_[a-z_0-9]+)\(", text, re.M))
        assert names, hdr
        import ctypes
        so = ctypes.CDLL(os.path.join(root, "tenet_amd", "lib", lib))
        missing = [n for n in names if not hasattr(so, n)]
        assert not missing, (hdr, missing)


@pytest.mark.parametrize("case", range(len(CASES)))
def test_layout_matches_oracle(case):
    rule, cod, dom = CASES[case]
    ps, os_ = spaces(rule, cod, dom)
    assert ps.len == os_.required_len
    pb = ps.blocks()
    assert len(pb) == len(os_.blocks)
    for (off, shape, strides, coupled, unc), ob in zip(pb, os_.blocks):
        assert off == ob.offset and shape == ob.shape and strides == ob.strides
        assert coupled == ob.key[0].coupled
        assert unc == ob.key[0].unc + ob.key[1].unc
    assert [(c, o, r, k) for c, o, r, k, _ in ps.sectors] == [tuple(s) for s in os_.sectors]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_recoupling_descriptor_matches_oracle(case):
    from tenet_amd.tensor import _Transform  # noqa: F401
    rule, cod, dom = CASES[case]
    ps, os_ = spaces(rule, cod, dom)
    n = ps.rank
    rng = np.random.default_rng(case)
    src = rng.standard_normal(ps.len)
    perms = PERMS4 if n == 4 else [([1, 0, 2], [3]), ([0, 2], [1, 3]), ([2], [0, 1, 3]), ([3, 1], [2, 0])]
    h = N.host()
    import ctypes as C
    for p1, p2 in perms:
        out = C.c_void_p()
        N.check_host(h.tnth_transform_compile(ps._h, N.sz_array(p1), len(p1), N.sz_array(p2), len(p2),
                                              C.byref(out)))
        try:
            got = replay_desc(h.tnth_transform_desc(out), src)
            want = OTensor(os_, src).permute(p1, p2).data
            if rule == "SU2":
                np.testing.assert_allclose(got, want, rtol=0, atol=1e-13 * np.abs(src).max())
            else:
                assert np.array_equal(got, want), (p1, p2)
        finally:
            h.tnth_transform_destroy(out)


CONTRACTS = [
    # rule, A legs, B legs, axes_a, axes_b, out_perm, out_nout, adj_a, adj_b
    ("U1", 2, [2, 3], [0, 1], [0, 1, 2, 3], 2, False, False),
    ("U1", 2, [3, 2], [0, 1], [1, 0, 2, 3], 2, False, False),
    ("U1", 2, [1, 3], [2, 0], [0, 2, 1, 3], 2, False, False),
    ("SU2", 2, [2, 3], [0, 1], [0, 1, 2, 3], 2, False, False),
    ("SU2", 2, [3, 2], [0, 1], [1, 0, 2, 3], 2, False, False),
    ("SU2", 2, [1, 3], [2, 0], [0, 2, 1, 3], 2, False, False),
    ("SU2", 2, [2, 3], [0, 1], [0, 1, 2, 3], 2, True, False),
    ("U1", 2, [2, 3], [0, 1], [0, 1, 2, 3], 2, False, True),
    ("U1", 2, [3, 2], [1, 0], [0, 2, 1, 3], 2, True, True),
    ("SU2", 2, [3, 2], [0, 1], [0, 1, 2, 3], 2, True, False),
    ("U1xZ2", 2, [1, 2], [3, 0], [2, 0, 1, 3], 1, False, False),
]


@pytest.mark.parametrize("case", range(len(CONTRACTS)))
def test_contraction_descriptor_matches_oracle(case):
    import ctypes as C
    rule, _, axa, axb, perm, nout, adja, adjb = CONTRACTS[case]
    V = {"U1": U1V, "SU2": SU2V, "U1xZ2": PZV}[rule]
    pa, oa = spaces(rule, [(V, False), (V, False)], [(V, False), (V, False)])
    pb, ob = pa, oa
    rng = np.random.default_rng(100 + case)
    A = rng.standard_normal(pa.len)
    B = rng.standard_normal(pb.len)
    h = N.host()
    out = C.c_void_p()
    N.check_host(h.tnth_contract_compile(pa._h, pb._h, N.sz_array(axa), N.sz_array(axb), len(axa),
                                         N.sz_array(perm), nout, int(adja), int(adjb), C.byref(out)))
    try:
        d = h.tnth_contract_desc(out)
        got = contract_desc_replay(d, A, B, d.contents.core_len)
        OA = OTensor(oa, A)
        OB = OTensor(ob, B)
        if adja:
            OA = OA.adjoint()
        if adjb:
            OB = OB.adjoint()
        want = OA.contract(axa, OB, axb, perm, nout).data
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * max(1.0, np.abs(want).max()))
    finally:
        h.tnth_contract_destroy(out)


def test_su2_symbols_match_oracle():
    h = N.host()
    o = OSU2()
    rng = range(0, 9)
    for a, b, c in itertools.product(rng, rng, rng):
        for d in range(0, min(a + b + c, 16) + 1):
            for e in o.channels(a, b):
                for f in o.channels(b, c):
                    got = h.tnth_symbol_f(1, a, b, c, d, e, f)
                    want = o.F(a, b, c, d, e, f)
                    assert abs(got - want) <= 2e-15, (a, b, c, d, e, f, got, want)
        for cc in o.channels(a, b):
            assert h.tnth_symbol_r(1, a, b, cc) == o.R(a, b, cc)


def test_host_compiler_rejects_bad_input():
    import ctypes as C
    ps, _ = spaces("U1", [(U1V, False), (U1V, False)], [(U1V, False)])
    h = N.host()
    out = C.c_void_p()
    rc = h.tnth_transform_compile(ps._h, N.sz_array([0, 0]), 2, N.sz_array([1]), 1, C.byref(out))
    assert rc == -1 and b"permutation" in h.tnth_last_error()
    # contracting a codomain leg with a codomain leg of the same orientation is a space mismatch
    rc = h.tnth_contract_compile(ps._h, ps._h, N.sz_array([0]), N.sz_array([0]), 1,
                                 N.sz_array([0, 1, 2, 3]), 2, 0, 0, C.byref(out))
    assert rc == -1 and b"match" in h.tnth_last_error()


def test_quickstart_scale_compose_closed_form():
    """Two U(1) rank-4 maps, charges -2..2 x degeneracy 8 per leg: coupled
    charge c has n_c = 64 p(c) rows (p(c) = pairs q1 + q2 = c: 1,2,3,4,5,4,3,2,1),
    so the payload is sum n_c^2 = 4096 * 85 elements and the compose
    A[a b; c d] . B[c d; g h] costs sum 2 n_c^3 = 2 * 64^3 * 325 flops
    (170.4 MFLOP); the swapped order B[d c; g h] adds an rhs recoupling and
    the same GEMM work."""
    r = Rule("U1")
    V = GradedSpace(r, [(q, 8) for q in range(-2, 3)])
    A = HomSpace(r, [V, V], [V, V])
    assert A.len == 4096 * 85
    c = _HostContraction(A, A, [2, 3], [0, 1], [0, 1, 2, 3], 2, False, False)
    assert c.flops == 2 * 64 ** 3 * 325
    d = c.desc.contents
    assert not d.lhs and not d.rhs and not d.out  # Core route: no recoupling
    sw = _HostContraction(A, A, [2, 3], [1, 0], [0, 1, 2, 3], 2, False, False)
    assert sw.flops == c.flops and sw.desc.contents.rhs
