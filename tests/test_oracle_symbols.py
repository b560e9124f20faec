"""Pin the oracle's symmetry providers (CPU only).

Golden data: TensorKit symbol tables in benchmarks/tensorkit_semantic_oracle.out
(181 SU(2) F values with 2j<=3, R with 2j<=4, dim/FS with 2j<=6) and the frozen
values of tenet-sectors/src/su2.rs:377-430.  What the tables leave open
(2j up to 8 for the SU(2) workload) is pinned by the pentagon and hexagon
identities and by orthogonality of the F-move matrices.
"""
import itertools
import json
import math
import os

import pytest

from oracle.symmetry import SU2, U1, Product, Z2, cantor_decode, cantor_encode, zigzag, unzigzag

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "tensorkit_oracle.json")))
KAT = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "kat.json")))


def test_fixture_hash_matches_reference_manifest():
    # benchmarks/tensorkit_semantic_oracle.sha256:2
    assert GOLD["sha256"].startswith("82f8e506") and GOLD["sha256"].endswith("e411")


def test_su2_f_table_181_values():
    su2 = SU2()
    assert len(GOLD["F_SU2"]) == 181
    for a, b, c, d, e, f, val in GOLD["F_SU2"]:
        got = su2.F(a, b, c, d, e, f)
        assert abs(got - val) <= 4e-16 * max(1.0, abs(val)) + 1e-16, (a, b, c, d, e, f, got, val)


def test_su2_r_dim_fs_tables():
    su2 = SU2()
    for a, b, c, val in GOLD["R_SU2"]:
        assert su2.R(a, b, c) == val
    for twoj, dim, twist, fs in GOLD["dim_SU2"]:
        assert su2.dim(twoj) == dim and su2.fs(twoj) == fs and twist == 1.0


def test_su2_frozen_symbols():
    su2 = SU2()
    for a, b, c, d, e, f, val in KAT["su2_frozen_symbols"]["F"]:
        assert su2.F(a, b, c, d, e, f) == val
    for a, b, c, val in KAT["su2_frozen_symbols"]["R"]:
        assert su2.R(a, b, c) == val
    assert su2.fs(253) == -1.0
    assert su2.channels(1, 1) == [0, 2]


def _adm(rule, a, b, c):
    return rule.nsym(a, b, c) == 1


def test_su2_pentagon_up_to_2j_4():
    """Pentagon identity, multiplicity-free form, over 2j <= 4 external labels."""
    su2 = SU2()
    labels = range(0, 5)
    F = su2.F
    count = 0
    for a, b, c, d in itertools.product(labels, repeat=4):
        for e in range(0, a + b + c + d + 1):
            for f in su2.channels(a, b):
                for g in su2.channels(f, c):
                    if not _adm(su2, g, d, e):
                        continue
                    for h in su2.channels(c, d):
                        for k in su2.channels(b, h):
                            if not _adm(su2, a, k, e):
                                continue
                            lhs = F(f, c, d, e, g, h) * F(a, b, h, e, f, k)
                            rhs = 0.0
                            for l in su2.channels(b, c):
                                rhs += F(a, b, c, g, f, l) * F(a, l, d, e, g, k) * F(b, c, d, k, l, h)
                            assert abs(lhs - rhs) < 1e-12
                            count += 1
    assert count > 1000


def test_su2_hexagon_up_to_2j_4():
    su2 = SU2()
    F, R = su2.F, su2.R
    for a, b, c in itertools.product(range(5), repeat=3):
        for d in range(0, a + b + c + 1):
            for e in su2.channels(c, a):
                for g in su2.channels(c, b):
                    if not (_adm(su2, e, b, d) and _adm(su2, a, g, d)):
                        continue
                    lhs = R(c, a, e) * F(a, c, b, d, e, g) * R(c, b, g)
                    rhs = 0.0
                    for f in su2.channels(a, b):
                        if _adm(su2, c, f, d):
                            rhs += F(c, a, b, d, e, f) * R(c, f, d) * F(a, b, c, d, f, g)
                    assert abs(lhs - rhs) < 1e-12


def test_su2_f_matrix_orthogonal_up_to_2j_8():
    """Each F^{abc}_d is a real orthogonal matrix; covers the 2j<=8 workload labels."""
    su2 = SU2()
    for a, b, c in itertools.product(range(0, 9, 2), range(0, 9, 3), range(0, 9)):
        for d in range(0, a + b + c + 1):
            es = [e for e in su2.channels(a, b) if _adm(su2, e, c, d)]
            fs = [f for f in su2.channels(b, c) if _adm(su2, a, f, d)]
            if not es:
                continue
            assert len(es) == len(fs)
            for e1 in es:
                for e2 in es:
                    dot = sum(su2.F(a, b, c, d, e1, f) * su2.F(a, b, c, d, e2, f) for f in fs)
                    assert abs(dot - (1.0 if e1 == e2 else 0.0)) < 1e-12


def test_sixj_special_case_closed_form():
    """{a b 0; c d f} reduces to (-1)^(a+d+f)/sqrt((2a+1)(2d+1)) when a=b, c=d."""
    from oracle.symmetry import sixj_doubled
    for a in range(0, 9):
        for c in range(0, 9):
            for f in range(abs(a - c), a + c + 1, 2):
                want = (-1) ** ((a + c + f) // 2) / math.sqrt((a + 1) * (c + 1))
                assert abs(sixj_doubled(a, a, 0, c, c, f) - want) < 1e-15


def test_u1_zigzag_and_fusion():
    u1 = U1()
    assert [zigzag(q) for q in (0, -1, 1, -2, 2)] == [0, 1, 2, 3, 4]
    assert all(unzigzag(zigzag(q)) == q for q in range(-20, 21))
    assert u1.channels(u1.encode(-2), u1.encode(3)) == [u1.encode(1)]
    assert u1.dual(u1.encode(5)) == u1.encode(-5)


def test_product_codec_roundtrip():
    p = Product(U1(), Z2())
    for l, r in itertools.product(range(30), range(2)):
        assert cantor_decode(cantor_encode(l, r)) == (l, r)
    a = p.encode((3, 1))
    b = p.encode((-1, 1))
    assert p.channels(a, b) == [p.encode((2, 0))]
    assert p.label(p.dual(a)) == (-3, 1)
