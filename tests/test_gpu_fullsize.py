"""GPU properties of the benchmark workloads at their full sizes.

The oracle is too slow at these sizes, so these tests check properties that
hold at any size (the small-size parity against the oracle lives in
test_gpu_workloads.py):

* U(1) environment update, D = 4096: linear in the environment; split-BF16
  and split-INT8 within 1e-10 of FP64; an abelian permutation followed by its inverse gives
  back the payload bit for bit.
* SU(2) F-move contraction, multiplicity floor(64/(2j+1)) (2.74 GB per
  operand): the contraction with both operands recoupled inside it equals
  explicit F-move permutations followed by a canonical compose (the
  reference's route invariance, contract_fusion.rs:532-700, at 1e-10 there);
  a permutation followed by its inverse returns the payload within 1e-14;
  split-BF16 and split-INT8 within 1e-10 of FP64.

* The batched workload at its full size (256 U(1)xZ2 merges, bond dimensions
  64..1024, 6656 sector GEMMs in one grouped launch): every merge against the
  C++/OpenMP CPU port replaying the same descriptors on the downloaded
  payloads (the port is pinned against the oracle in tests/test_cpu_port.py),
  in FP64, FP32 and both split precisions, eager and as a hipGraph replay.

Norms are taken on the device (FP64 weighted inner products), so nothing of
gigabyte size crosses PCIe.
"""
import math

import numpy as np
import pytest

from tenet_amd.tensor import Runtime, TensorMap
from tenet_amd.workloads import MPSEnvUpdate, SU2FMoveContract, SU2FMoveSwapContract

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rt():
    r = Runtime(0)
    yield r
    r.close()


def dev_rel(a, b):
    """||a - b|| / ||b|| on the device."""
    return a.add(b, 1.0, -1.0).norm() / b.norm()


@pytest.fixture(scope="module")
def u1_env(rt):
    return MPSEnvUpdate(rt, "U1", 4096, "f64")


def test_u1_env_update_full_size_is_linear_in_environment(rt, u1_env):
    wl = u1_env
    L1 = wl.L
    L2 = TensorMap.rand(rt, [wl.V, wl.M], [wl.V], seed=99, normal=True)
    f1 = wl.run(L1).copy()
    f2 = wl.run(L2).copy()
    f12 = wl.run(L1.add(L2, 0.75, -1.25)).copy()
    want = f1.add(f2, 0.75, -1.25)
    err = dev_rel(f12, want)
    assert err < 1e-13, err
    assert math.isfinite(f1.norm()) and f1.norm() > 0


@pytest.mark.parametrize("prec", ["split_bf16", "split_i8"])
def test_u1_env_update_full_size_split_within_1e10(rt, u1_env, prec):
    ref = u1_env.run().copy()
    split = MPSEnvUpdate(rt, "U1", 4096, prec)
    got = split.run()
    assert np.array_equal(split.L.to_numpy(), u1_env.L.to_numpy())  # same seeded inputs
    err = dev_rel(got, ref)
    assert err < 1e-10, err


def test_u1_permutation_roundtrip_full_size_bit_exact(rt, u1_env):
    T = u1_env.L  # L[a, w; b], 5.97 M elements
    P = T.permute([2, 0], [1])  # codomain (b, a), domain (w): crosses codomain / domain
    Q = P.permute([1, 2], [0])
    assert Q.space == T.space
    assert np.array_equal(Q.to_numpy().view(np.uint64), T.to_numpy().view(np.uint64))


@pytest.fixture(scope="module")
def su2_fmove(rt):
    return SU2FMoveContract(rt, mult=64, twoj_max=8, precision="f64")


def test_su2_fmove_full_size_route_invariance(rt, su2_fmove):
    wl = su2_fmove
    c1 = wl.run().copy()
    # explicit F-moves: A' = A(0, 2 <- 1, 3), B' = B(2, 0 <- 1, 3), then C = A' B'
    Ap = wl.A.permute([0, 2], [1, 3])
    Bp = wl.B.permute([2, 0], [1, 3])
    c2 = Ap.compose(Bp)
    assert c2.space == c1.space
    del Ap, Bp
    err = dev_rel(c2, c1)
    assert err < 1e-12, err


@pytest.mark.parametrize("prec", ["split_bf16", "split_i8"])
def test_su2_fmove_full_size_split_within_1e10(rt, su2_fmove, prec):
    """The benchmark's split-precision contraction against the FP64 one
    (1.8 GFLOP per 565-leg sector job, K up to 7197: the 128 x 128 tiles)."""
    ref = su2_fmove.run().copy()
    got = su2_fmove.A.contract([1, 3], su2_fmove.B, [2, 0], [0, 1, 2, 3], 2, precision=prec)
    assert got.space == ref.space
    err = dev_rel(got, ref)
    del got
    assert err < 1e-10, err


def test_su2_permutation_roundtrip_full_size(rt, su2_fmove):
    A = su2_fmove.A
    P = A.permute([0, 2], [1, 3])
    Q = P.permute([0, 2], [1, 3])
    assert Q.space == A.space
    del P
    err = dev_rel(Q, A)
    assert err < 1e-14, err


@pytest.fixture(scope="module")
def su2_swap(rt):
    return SU2FMoveSwapContract(rt, mult=64, twoj_max=8, precision="f64")


def test_su2_swap_form_full_size_route_invariance(rt, su2_swap):
    """C = permute(A, (0,2,1), (3,))[2,3] . B[2,0] is A[1,3] . B[2,0]: the
    explicit F-move and the fused operand transform agree at 342.6 M
    elements per operand."""
    wl = su2_swap
    c1 = wl.run().copy()
    c2 = wl.A.contract([1, 3], wl.B, [2, 0], [0, 1, 2, 3], 2)
    assert c2.space == c1.space
    err = dev_rel(c2, c1)
    del c2
    assert err < 1e-12, err


def test_su2_swap_is_an_involution_full_size(rt, su2_swap):
    """Braiding codomain legs 1 and 2 twice is the identity (symmetric
    braiding), through two rounds of Multi groups."""
    Q = su2_swap.Ap.permute([0, 2, 1], [3])
    assert Q.space == su2_swap.A.space
    err = dev_rel(Q, su2_swap.A)
    del Q
    assert err < 1e-14, err


@pytest.mark.parametrize("prec", ["split_bf16", "split_i8"])
def test_su2_swap_form_full_size_split_within_1e10(rt, su2_swap, prec):
    ref = su2_swap.run().copy()
    got = su2_swap.Ap.contract([2, 3], su2_swap.B, [2, 0], [0, 1, 2, 3], 2, precision=prec)
    err = dev_rel(got, ref)
    del got, ref
    assert err < 1e-10, err


@pytest.fixture(scope="module")
def batched_ref(rt):
    """The FP64 CPU-port results of the 256 full-size merges (seeded inputs)."""
    from oracle import cpu_port as P
    from tenet_amd.workloads import BatchedMerge
    wl = BatchedMerge(rt, 256, 64, 1024, precision="f64")
    refs = [P.Contract(c.desc, c.dst_space.len)(a.to_numpy(), b.to_numpy()) for a, b, _, c in wl.pairs]
    del wl
    return refs


@pytest.mark.parametrize("prec,tol", [("f64", 1e-13), ("f32", 1e-5), ("split_bf16", 1e-10), ("split_i8", 1e-10)])
def test_batched_full_size_matches_cpu_port(rt, batched_ref, prec, tol):
    from tenet_amd.workloads import BatchedMerge
    wl = BatchedMerge(rt, 256, 64, 1024, precision=prec)
    assert wl.n_jobs > 6000
    wl.run()
    rt.synchronize()
    got = wl.results()
    num = sum(float(np.linalg.norm(g - r) ** 2) for g, r in zip(got, batched_ref))
    den = sum(float(np.linalg.norm(r) ** 2) for r in batched_ref)
    worst = max(float(np.linalg.norm(g - r) / np.linalg.norm(r)) for g, r in zip(got, batched_ref))
    assert math.sqrt(num / den) < tol and worst < 4 * tol, (prec, math.sqrt(num / den), worst)
    wl.capture()
    for _, _, Cm, _ in wl.pairs:
        Cm.scale_(0.0)
    rt.synchronize()
    wl.run()
    rt.synchronize()
    assert all(np.array_equal(x, y) for x, y in zip(wl.results(), got))
