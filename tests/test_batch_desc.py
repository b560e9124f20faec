"""Stacked (batched) contraction descriptors (tnth_contract_batch), CPU only.

A batched descriptor of c copies, replayed on stacked payloads, must give
exactly the stacked results of c single replays (the shared operand given
once): every job, layout, inactive block and uncovered region is the single
contraction's, moved by whole payload lengths.  The replay is the oracle's
(tests/descs.py), so this checks the descriptor arithmetic only; the device
runs the same descriptors in tests/test_gpu_heff.py.
"""
import numpy as np
import pytest

from descs import contract_desc_replay
from tenet_amd import _native as N
from tenet_amd.tensor import GradedSpace, HomSpace, Rule, _HostBatch, _HostContraction
from tenet_amd.workloads import mps_env_spaces

CASES = {
    # H_eff chain steps: environment . theta (lhs shared), theta-like . MPO (rhs shared)
    "L.theta": (lambda V, P, M: (HomSpace(V.rule, [V, M], [V]), HomSpace(V.rule, [V], [P, P, V])),
                [2], [0], [0, 1, 2, 3, 4], 2, 0),
    "x1.W": (lambda V, P, M: (HomSpace(V.rule, [V, M], [P, P, V]), HomSpace(V.rule, [M, P], [P, M])),
             [1, 2], [3, 1], [0, 3, 4, 1, 2], 2, 1),
    "x3.R": (lambda V, P, M: (HomSpace(V.rule, [V, M], [P, P, V]), HomSpace(V.rule, [V], [V, M])),
             [1, 4], [2, 0], [0, 1, 2, 3], 1, 1),
}


def _len(desc, side):
    d = desc.contents
    t = getattr(d, side)
    if t:
        return t.contents.src_len
    return d.lhs_len if side == "lhs" else d.rhs_len


@pytest.mark.parametrize("rule", ["U1", "SU2"])
@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("copies", [1, 3])
def test_batched_descriptor_replays_as_stacked_singles(rule, case, copies):
    _, V, P, M = mps_env_spaces(rule, 24)
    mk, ax, bx, perm, nout, shared = CASES[case]
    A, B = mk(V, P, M)
    one = _HostContraction(A, B, ax, bx, perm, nout, False, False)
    bat = _HostBatch(one, copies, shared)
    rng = np.random.default_rng(copies)
    la, lb = _len(one.desc, "lhs"), _len(one.desc, "rhs")
    core = one.desc.contents.core_len
    ins_a = [rng.normal(size=la) for _ in range(copies if shared != 0 else 1)]
    ins_b = [rng.normal(size=lb) for _ in range(copies if shared != 1 else 1)]
    singles = [contract_desc_replay(one.desc, ins_a[c if shared != 0 else 0], ins_b[c if shared != 1 else 0], core)
               for c in range(copies)]
    got = contract_desc_replay(bat.desc, np.concatenate(ins_a), np.concatenate(ins_b), core * copies)
    assert np.array_equal(got, np.concatenate(singles))
    d = bat.desc.contents
    assert d.n_jobs == one.desc.contents.n_jobs * copies
    assert bat.flops == one.flops * copies


def test_batch_rejects_bad_requests():
    _, V, P, M = mps_env_spaces("U1", 24)
    A, B = CASES["L.theta"][0](V, P, M)
    one = _HostContraction(A, B, [2], [0], [0, 1, 2, 3, 4], 2, False, False)
    for copies, shared in ((0, 0), (2, 7)):
        with pytest.raises(Exception):
            _HostBatch(one, copies, shared)
