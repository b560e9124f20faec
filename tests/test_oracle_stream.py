"""Pin the oracle's layout, recoupling and contraction against golden vectors (CPU only).

* TensorKit invariant stream (norm/tr after compose, permute, adjoint-compose,
  add), benchmarks/tensorkit_semantic_oracle.out section 3, mirrored by
  tenet/tests/semantic_suite.rs:839-1030 (tolerance 1e-9 there; we hold 1e-12).
* TensorKit planar repartition data (bends), section 4.
* SU(2) output recoupling KAT, tenet-tensors/src/tests/contract_fusion.rs:4693-4776.
* Replay KATs, tenet-tensors/src/tests/mod.rs:1099-1320, tree_transform_exec.rs:184-600.
* Microbench checksums, docs/su2_authority.md:81-83.
* Quickstart inner product 50, tenet-network/examples/quickstart.rs:35-39.
* U(1): every permute and contraction equals the dense-tensor computation.
"""
import json
import os

import numpy as np
import pytest

from oracle.fusion import HomSpace, Tree, permute_pair
from oracle.replay import packed_column_major, replay
from oracle.tensor import OTensor, leg, make_rule, splitmix_unit

HERE = os.path.dirname(__file__)
GOLD = json.load(open(os.path.join(HERE, "golden", "tensorkit_oracle.json")))
KAT = json.load(open(os.path.join(HERE, "golden", "kat.json")))


def oracle_fill(c0, labels, idx):
    """tenet/tests/semantic_suite.rs:742-756 (= fill_value of the TensorKit script)."""
    l1, l2, m1, m2, lc = labels
    w = (c0 + 7 * l1 + 11 * l2 + 13 * m1 + 17 * m2 + 19 * lc + 23 * (idx[0] + 1)
         + 29 * (idx[1] + 1) + 31 * (idx[2] + 1) + 37 * (idx[3] + 1))
    return float(w % 41 - 20)


def _stream(rule, V):
    sp = HomSpace(rule, [V, V], [V, V])
    lab = rule.label

    def fn(c0):
        return lambda key, idx: oracle_fill(
            c0, [lab(key[0].unc[0]), lab(key[0].unc[1]), lab(key[1].unc[0]), lab(key[1].unc[1]),
                 lab(key[0].coupled)], idx)

    A = OTensor.from_block_fn(sp, fn(3))
    B = OTensor.from_block_fn(sp, fn(5))
    C = A.compose(B)
    D = A.permute([1, 0], [3, 2])
    E = D.compose(C)
    G = A.adjoint().compose(A)
    H = E.add(A, 1.0, 0.5)
    got = {n: [t.norm(), t.tr()] for n, t in
           [("s1a", A), ("s1b", B), ("s2", C), ("s3", D), ("s4", E), ("s5", G), ("s7", H)]}
    got["s8"] = H.compose(H).tr()
    return got


@pytest.mark.parametrize("name,pairs", [
    ("U1", [(-1, 2), (0, 3), (1, 2)]),
    ("SU2", [(0, 2), (1, 2), (2, 1)]),
])
def test_tensorkit_invariant_stream(name, pairs):
    rule = make_rule(name)
    got = _stream(rule, leg(rule, pairs))
    want = GOLD["stream"][name]
    for step, val in want.items():
        if step == "s8":
            assert abs(got[step] - val) <= 1e-12 * abs(val)
        elif step in got:
            for g, w in zip(got[step], val):
                assert abs(g - w) <= 1e-12 * max(1.0, abs(w)), (step, g, w)


def _sequential(rule, legs):
    sp = HomSpace(rule, legs[:2], legs[2:])
    return OTensor(sp, np.arange(1, sp.required_len + 1, dtype=float))


def test_tensorkit_repartition_su2_and_u1():
    su2, u1 = make_rule("SU2"), make_rule("U1")
    S = [leg(su2, [(0, 1), (1, 1)])] * 3 + [leg(su2, [(0, 1), (1, 2)])]
    # TensorKit repartition(t, 3): codomain (1,2,4), domain (3)  (1-based)
    t = _sequential(su2, S).permute([0, 1, 3], [2])
    np.testing.assert_allclose(t.data, GOLD["repartition"]["SU2 2|2 -> 3|1"], rtol=0, atol=1e-13)
    U = [leg(u1, [(0, k)]) for k in (1, 2, 3, 4)]
    t = _sequential(u1, U).permute([0, 1, 3], [2])
    assert list(t.data) == GOLD["repartition"]["U1 2|2 -> 3|1"]


def test_su2_output_recoupling_kat():
    k = KAT["su2_output_recoupling"]
    su2 = make_rule("SU2")
    src = Tree(tuple(k["src_tree"]["unc"]), (False,) * 4, 0, tuple(k["src_tree"]["inner"]))
    empty = Tree((), (), 0, ())
    rows = permute_pair(su2, (src, empty), k["perm"], [])
    inner_to_coeff = {pair[0].inner: c for pair, c in rows}
    coeffs = [inner_to_coeff[tuple(d["inner"])] for d in k["dst_trees"]]
    np.testing.assert_allclose(coeffs, k["coefficients"], atol=1e-12)
    # dst = alpha * coeff * (lhs * rhs) + beta * dst
    out = [k["alpha"] * c * k["lhs"] * k["rhs"] + k["beta"] * f for c, f in zip(coeffs, k["dst_fill"])]
    np.testing.assert_allclose(out, k["expected"], atol=k["tol"])


@pytest.mark.parametrize("name", ["multi_2x2", "multi_orientation"])
def test_replay_multi_kats(name):
    k = KAT[name]
    srcl = packed_column_major(k["src_blocks"])
    dstl = packed_column_major(k["dst_blocks"])
    n = sum(int(np.prod(s)) for s in k["dst_blocks"])
    out = replay([k["fill"]] * n, k["src"], dstl, srcl,
                 [("multi", k["dst_idx"], k["src_idx"], k["U_dst_src"])], k["alpha"], k["beta"])
    assert list(out) == k["expected"]


def test_replay_sorted_offsets_kat():
    k = KAT["multi_sorted_offsets"]
    lay = packed_column_major(k["blocks"])
    specs = [("multi", m["dst_idx"], m["src_idx"], m["U_dst_src"]) for m in k["multis"]]
    out = replay([k["fill"]] * 8, k["src"], lay, lay, specs, k["alpha"], k["beta"])
    assert list(out) == k["expected"]


def _microbench(rule, labels, d):
    V = leg(rule, [(l, d) for l in labels])
    sp = HomSpace(rule, [V, V], [V, V])
    n = sp.required_len
    A = OTensor(sp, (np.arange(n) % 17) * 0.25 - 2.0)
    B = OTensor(sp, (np.arange(n) % 13) * 0.5 - 3.0)
    out = []
    # microbench_fusion.rs:24-50: compose, swap, swap+out
    for la, lb, perm in [([2, 3], [0, 1], [0, 1, 2, 3]), ([3, 2], [0, 1], [0, 1, 2, 3]),
                         ([3, 2], [0, 1], [1, 0, 2, 3])]:
        C = A.contract(la, B, lb, perm, 2)
        out.append((C.data.sum(), np.abs(C.data).sum()))
    return out


@pytest.mark.parametrize("d", [4, 16])
def test_su2_microbench_checksums(d):
    got = _microbench(make_rule("SU2"), [0, 1, 2], d)
    want = KAT["microbench_checksums"]["SU2"][str(d)]
    for (s, l1), w in zip(got, want):
        assert abs(s - w) <= 1e-13 * l1 + 1e-12, (s, w)


def test_quickstart_inner_is_50():
    u1 = make_rule("U1")
    V = leg(u1, [(-1, 1), (0, 2), (1, 1)])
    sp = HomSpace(u1, [V], [V])
    a = OTensor.from_block_fn(sp, lambda key, idx: float(2 + u1.label(key[0].coupled)) if idx[0] == idx[1] else 0.0)
    b = OTensor.from_block_fn(sp, lambda key, idx: float(2 - u1.label(key[0].coupled)) if idx[0] == idx[1] else 0.0)
    c = a.contract([1], b, [0])
    assert c.inner(c) == KAT["quickstart"]["inner"]


def test_splitmix_fill_closed_form():
    """tenet/src/typed.rs:2086-2092, scalar restatement vs the vectorised one."""
    def scalar(seed, n):
        st, out = seed, []
        for _ in range(n):
            st = (st + 0x9E3779B97F4A7C15) & ((1 << 64) - 1)
            z = st
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & ((1 << 64) - 1)
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & ((1 << 64) - 1)
            z ^= z >> 31
            out.append((z >> 11) / float(1 << 52) - 1.0)
        return out
    for seed in (0, 1, 2, 12345, (1 << 64) - 1):
        assert list(splitmix_unit(seed, 50)) == scalar(seed, 50)
    v = splitmix_unit(7, 100000)
    assert v.min() >= -1.0 and v.max() < 1.0


@pytest.mark.parametrize("p1,p2", [([1, 0], [2, 3]), ([0, 3], [2, 1]), ([2, 0, 1], [3]),
                                   ([3], [0, 1, 2]), ([1, 2], [0, 3]), ([], [3, 2, 1, 0])])
def test_u1_permute_equals_dense_transpose(p1, p2):
    u1 = make_rule("U1")
    V = leg(u1, [(-1, 2), (0, 3), (1, 1)])
    W = leg(u1, [(0, 1), (1, 2), (2, 2)])
    sp = HomSpace(u1, [V, W], [W, V])
    t = OTensor.rand(sp, 11)
    out = t.permute(p1, p2)
    np.testing.assert_array_equal(out.to_dense(), np.transpose(t.to_dense(), p1 + p2))


def test_u1_contract_equals_dense_einsum():
    u1 = make_rule("U1")
    V = leg(u1, [(-1, 2), (0, 3), (1, 2)])
    sp = HomSpace(u1, [V, V], [V, V])
    A = OTensor.rand(sp, 1)
    B = OTensor.rand(sp, 2)
    C = A.contract([3, 0], B, [0, 2], [2, 0, 1, 3], 2)
    Ad, Bd = A.to_dense(), B.to_dense()
    # A axes (a0,a1,a2,a3): contract a3 with b0 and a0 with b2 -> open (a1,a2,b1,b3)
    ref = np.einsum("xabq,qcxd->abcd", Ad, Bd)
    np.testing.assert_allclose(C.to_dense(), np.transpose(ref, [2, 0, 1, 3]), atol=1e-12)


def test_su2_inner_line_swap_invariants():
    """The 3 <- 1 codomain swap (0, 2, 1) recouples the inner line (Multi
    groups): it is an involution, keeps the dimension-weighted norm (the F/R
    moves are unitary), and contracting the swapped tensor equals contracting
    the original with the swapped axis -- the fused operand transform and the
    explicit F-move take different group plans to the same result."""
    su2 = make_rule("SU2")
    V = leg(su2, [(0, 3), (1, 2), (2, 2), (3, 1)])
    A = OTensor.rand(HomSpace(su2, [V, V, V], [V]), 1)
    B = OTensor.rand(HomSpace(su2, [V, V], [V, V]), 2)
    Ap = A.permute([0, 2, 1], [3])
    assert np.abs(Ap.permute([0, 2, 1], [3]).data - A.data).max() < 1e-14
    assert abs(Ap.norm() - A.norm()) < 1e-12 * A.norm()
    C1 = Ap.contract([2, 3], B, [2, 0], [0, 1, 2, 3], 2)
    C2 = A.contract([1, 3], B, [2, 0], [0, 1, 2, 3], 2)
    assert np.linalg.norm(C1.data - C2.data) < 1e-13 * np.linalg.norm(C2.data)


def test_layout_closed_form_z2():
    """tenet-core/src/tests.rs:189-260: strides [1,2,12,24], exact coverage."""
    z2 = make_rule("Z2")
    L = lambda d: leg(z2, [(0, d), (1, d)])
    sp = HomSpace(z2, [L(2), L(3)], [L(2), L(3)])
    covered = np.zeros(sp.required_len, dtype=int)
    for b in sp.blocks:
        assert b.strides == (1, 2, 12, 24)
        for idx in np.ndindex(*b.shape):
            covered[b.offset + sum(i * s for i, s in zip(idx, b.strides))] += 1
    assert (covered == 1).all()
