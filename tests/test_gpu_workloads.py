"""GPU parity of the benchmark workloads at small sizes against the oracle.

The environment update is checked against the oracle's chain written with the
textbook index order (T2[a', w; s', b]); the workload keeps T2 in the core
GEMM's output order instead, which must not change L'.
Bars: FP64 within 1e-13 relative Frobenius error; split-BF16 within 1e-10.
"""
import numpy as np
import pytest

from descs import ospace
from oracle.tensor import OTensor
from tenet_amd.tensor import Runtime
from tenet_amd.workloads import MPSEnvUpdate, SU2FMoveContract

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rt():
    r = Runtime(0)
    yield r
    r.close()


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _o(t):
    return OTensor(ospace(t.space), t.to_numpy())


@pytest.mark.parametrize("rule_name,D", [("U1", 128), ("SU2", 256)])
@pytest.mark.parametrize("prec", ["f64", "split_bf16"])
def test_env_update_matches_oracle(rt, rule_name, D, prec):
    wl = MPSEnvUpdate(rt, rule_name, D, prec)
    got = wl.run().to_numpy()
    assert wl.out.space == wl.L.space
    L, A, W = _o(wl.L), _o(wl.A), _o(wl.W)
    t1 = L.contract([2], A, [0], [0, 1, 2, 3], 2)
    t2 = t1.contract([1, 2], W, [3, 1], [0, 2, 3, 1], 2)
    want = A.adjoint().contract([2, 0], t2, [0, 2], [0, 1, 2], 2).data
    tol = {"f64": 1e-13, "split_bf16": 1e-10}[prec]
    assert rel(got, want) < tol, (rule_name, prec, rel(got, want))


def test_fmove_contract_matches_oracle(rt):
    wl = SU2FMoveContract(rt, mult=6, twoj_max=4, precision="f64")
    got = wl.run().to_numpy()
    want = _o(wl.A).contract([1, 3], _o(wl.B), [2, 0], [0, 1, 2, 3], 2).data
    assert rel(got, want) < 1e-13
