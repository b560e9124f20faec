"""The U(1) spin-1/2 Heisenberg MPO used by the H_eff workloads.

Pinned by physics, not by restating the construction: contracted with the
boundary channels (left 4, right 0) the two-site MPO must equal
J (Sx Sx + Sy Sy + Sz Sz) built from the complex spin matrices, with spectrum
{-3/4 J, 1/4 J (x3)}; the three-site open chain must have ground energy -J
(S2 . (S1 + S3) at S13 = 1, S = 1/2).  The dense tensor comes from the
oracle's dense embedding of the block payload.
"""
import numpy as np
import pytest

from descs import ospace
from oracle.tensor import OTensor
from tenet_amd.workloads import heisenberg_mpo_payload

# dense M basis follows ascending sector ids: charge 0 (channels 0, 3, 4), -1 (1), +1 (2)
DENSE_OF_CHANNEL = {0: 0, 3: 1, 4: 2, 1: 3, 2: 4}


def dense_w(J):
    space, data = heisenberg_mpo_payload(J)
    return OTensor(ospace(space), data).to_dense()  # [w_r, s_in, s_out, w_l]


def spin_ops():
    sp = np.array([[0, 0], [1, 0]], dtype=complex)  # basis (down, up)
    sm = sp.T.copy()
    return (sp + sm) / 2, (sp - sm) / 2j, np.diag([-0.5, 0.5]).astype(complex)


def chain(W, n):
    """<o_1 .. o_n| H |i_1 .. i_n> of an open n-site chain as a (2^n, 2^n) matrix."""
    l, r = DENSE_OF_CHANNEL[4], DENSE_OF_CHANNEL[0]
    T = W[:, :, :, l].transpose(1, 2, 0)  # (i1, o1, w1)
    for _ in range(n - 1):
        T = np.einsum("...w,vstw->...stv", T, W)  # append (i_k, o_k), new right channel last
    T = T[..., r]  # (i1, o1, ..., in, on)
    order = list(range(1, 2 * n, 2)) + list(range(0, 2 * n, 2))
    return T.transpose(order).reshape(2 ** n, 2 ** n)


@pytest.mark.parametrize("J", [1.0, 0.37])
def test_two_site_mpo_is_exchange_coupling(J):
    H = chain(dense_w(J), 2)
    sx, sy, sz = spin_ops()
    want = J * sum(np.kron(a, a) for a in (sx, sy, sz))
    assert np.abs(want.imag).max() == 0
    assert np.allclose(H, want.real, atol=1e-15)
    assert np.allclose(np.linalg.eigvalsh(H), J * np.array([-0.75, 0.25, 0.25, 0.25]), atol=1e-14)


def test_three_site_chain_ground_energy():
    H = chain(dense_w(1.0), 3)
    sx, sy, sz = spin_ops()
    I = np.eye(2)
    bond = sum(np.kron(a, a) for a in (sx, sy, sz))
    want = np.kron(bond, I) + np.kron(I, bond)
    assert np.allclose(H, want.real, atol=1e-15)
    assert abs(np.linalg.eigvalsh(H)[0] + 1.0) < 1e-14


def test_mpo_blocks_conserve_charge():
    """Every stored element sits in a charge-allowed block, and the operator
    matrix is lower triangular in the channel order."""
    W = dense_w(1.0)
    q_ch = {0: 0, 3: 0, 4: 0, 1: -1, 2: 1}
    q_phys = [0, 1]
    for ch_r, dr in DENSE_OF_CHANNEL.items():
        for ch_l, dl in DENSE_OF_CHANNEL.items():
            blk = W[dr, :, :, dl]
            if ch_r > ch_l:
                assert not blk.any()
            for si in range(2):
                for so in range(2):
                    if blk[si, so]:
                        assert q_ch[ch_r] + q_phys[si] == q_phys[so] + q_ch[ch_l]
