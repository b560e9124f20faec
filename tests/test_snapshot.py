"""Snapshot codec (TeNeT v1 wire format) on the host library.

Pins: the framing offsets the reference's own tests parse
(tenet/tests/typed_serialization.rs:382-411: header end = 20 + key length at
byte 12; leg degeneracy offsets), the reference tests' round-trip and
determinism properties (:670-807), the rejection of malformed input
(:965-1090), and an independent parser written here from the format that
places every decoded block value at the oracle's layout position of its
(codomain tree, domain tree) key.
"""
import ctypes as C
import struct

import numpy as np
import pytest

from descs import spaces
from oracle.fusion import Tree
from tenet_amd import _native as N
from tenet_amd.tensor import GradedSpace, HomSpace, Rule

REAL_BITS = [0x8000000000000000, 0x3FF0000000000001, 0x7FF8000000000042]


def u64_at(b, off):
    return struct.unpack_from("<Q", b, off)[0]


def header_end(b):
    return 20 + u64_at(b, 12)


def leg_degeneracy_offsets(b):
    """Port of the reference test helper dense_leg_degeneracy_offsets."""
    pos = header_end(b) + 1
    rank = u64_at(b, pos) + u64_at(b, pos + 8)
    pos += 16
    offs = []
    for _ in range(rank):
        pos += 1
        n = u64_at(b, pos)
        pos += 8
        for _ in range(n):
            pos += 8 + u64_at(b, pos)
            offs.append(pos)
            pos += 8
    return offs


def encode_tensor(space, payload, adjoint=False):
    h = N.host()
    b = C.c_void_p()
    p = np.ascontiguousarray(payload, dtype=np.float64)
    N.check_host(h.tnth_snapshot_tensor(space._h, p.ctypes.data_as(C.POINTER(C.c_double)), space.len,
                                        int(adjoint), C.byref(b)))
    from tenet_amd.tensor import _take_bytes
    return _take_bytes(b)


def decode_tensor(data):
    h = N.host()
    sp, adj = C.c_void_p(), C.c_int()
    rc = h.tnth_snapshot_tensor_space(data, len(data), C.byref(sp), C.byref(adj))
    if rc != N.OK:
        raise N.TenetError(rc, h.tnth_last_error().decode())
    name = {v: k for k, v in N.RULES.items()}[h.tnth_space_rule(sp)]
    space = HomSpace._adopt(Rule(name), sp.value)
    out = np.full(space.len, 7.0)
    rc = h.tnth_snapshot_tensor_payload(data, len(data), space._h, out.ctypes.data_as(C.POINTER(C.c_double)),
                                        space.len)
    if rc != N.OK:
        raise N.TenetError(rc, h.tnth_last_error().decode())
    return space, out, bool(adj.value)


SU2_LEG = [(0, 1), (1, 2)]  # (2j, multiplicity) as in the reference test's su2_leg


def su2_spaces():
    return spaces("SU2", [(SU2_LEG, False)] * 3, [(SU2_LEG, True)])


def bit_fill(ospace):
    """The reference test's block function: real_bits[(2j_coupled + sum idx) % 3]."""
    data = np.zeros(ospace.required_len)
    for blk in ospace.blocks:
        coupled = blk.key[0].coupled
        for idx in np.ndindex(*blk.shape):
            off = blk.offset + sum(i * s for i, s in zip(idx, blk.strides))
            data[off] = struct.unpack("<d", struct.pack("<Q", REAL_BITS[(coupled + sum(idx)) % 3]))[0]
    return data


def test_leg_roundtrip_su2_dual_and_deterministic():
    g = GradedSpace(Rule("SU2"), SU2_LEG, dual=True)
    b = g.to_bytes()
    assert g.to_bytes() == b
    assert b[:8] == b"TENETTS\0" and struct.unpack_from("<H", b, 8)[0] == 1 and b[10] == 1 and b[11] == 0
    assert b[20:header_end(b)] == b"su2"
    r = GradedSpace.from_bytes(b)
    assert r == g and r.is_dual


def test_u1_dual_leg_does_not_dualize_twice():
    g = GradedSpace(Rule("U1"), [(3, 7)]).dual()
    assert [g.rule.label(s) for s in g.ids] == [-3]
    b = g.to_bytes()
    # label is the charge as little-endian i32 (the reference test codec)
    assert b[header_end(b) + 1 + 8 + 8:][:4] == struct.pack("<i", -3)
    r = GradedSpace.from_bytes(b)
    assert r.is_dual and [r.rule.label(s) for s in r.ids] == [-3] and list(r.degs) == [7]


def test_tensor_roundtrip_exact_bits_and_framing():
    ps, os_ = su2_spaces()
    data = bit_fill(os_)
    b = encode_tensor(ps, data)
    assert encode_tensor(ps, data) == b  # deterministic
    assert b[10] == 2 and b[11] == 1 and b[header_end(b)] == 1  # tensor, f64, dense
    # every leg degeneracy sits where the reference test helper finds it
    offs = leg_degeneracy_offsets(b)
    assert [u64_at(b, o) for o in offs] == [1, 2] * 4
    space, out, adj = decode_tensor(b)
    assert not adj and space == ps
    assert np.array_equal(out.view(np.uint64), data.view(np.uint64))
    assert any(len(blk.key[0].inner) for blk in os_.blocks)  # inner lines present
    assert encode_tensor(space, out) == b


def test_adjoint_flag_roundtrip():
    ps, os_ = su2_spaces()
    data = bit_fill(os_)
    b = encode_tensor(ps, data, adjoint=True)
    assert b[header_end(b)] == 3 and b[header_end(b) + 1] == 1
    _, out, adj = decode_tensor(b)
    assert adj and np.array_equal(out.view(np.uint64), data.view(np.uint64))


def _parse_blocks(b, rule):
    """Independent parser: [(codomain tree, domain tree, shape, values)]."""
    pos = header_end(b) + 1
    nc, nd = u64_at(b, pos), u64_at(b, pos + 8)
    pos += 16
    legs_dual = []
    for _ in range(nc + nd):
        legs_dual.append(bool(b[pos]))
        pos += 1
        n = u64_at(b, pos)
        pos += 8
        for _ in range(n):
            pos += 8 + u64_at(b, pos) + 8

    def label(raw):
        if rule == "SU2":
            return raw[0]
        q = struct.unpack("<i", raw)[0]
        return 2 * q if q >= 0 else -2 * q - 1

    def tree(kb, p, duals):
        ln = u64_at(kb, p)
        coupled = label(kb[p + 8:p + 8 + ln])
        p += 8 + ln
        lists = []
        for _ in range(2):
            n = u64_at(kb, p)
            p += 8
            xs = []
            for _ in range(n):
                ln = u64_at(kb, p)
                xs.append(label(kb[p + 8:p + 8 + ln]))
                p += 8 + ln
            lists.append(tuple(xs))
        nv = u64_at(kb, p)
        p += 8
        assert all(u64_at(kb, p + 8 * i) == 1 for i in range(nv)) and nv == max(0, len(lists[0]) - 1)
        p += 8 * nv
        return Tree(lists[0], tuple(duals), coupled, lists[1]), p

    nb = u64_at(b, pos)
    pos += 8
    out = []
    for _ in range(nb):
        kl = u64_at(b, pos)
        kb = b[pos + 8:pos + 8 + kl]
        pos += 8 + kl
        ct, p = tree(kb, 0, legs_dual[:nc])
        dt, p = tree(kb, p, legs_dual[nc:])
        assert p == kl
        rank = u64_at(b, pos)
        shape = tuple(u64_at(b, pos + 8 + 8 * i) for i in range(rank))
        pos += 8 + 8 * rank
        cnt = u64_at(b, pos)
        vals = np.frombuffer(b, dtype="<f8", count=cnt, offset=pos + 8)
        pos += 8 + 8 * cnt
        out.append((ct, dt, shape, vals))
    assert pos == len(b)
    return out


@pytest.mark.parametrize("rule_name", ["U1", "SU2"])
def test_blocks_land_at_oracle_layout(rule_name):
    leg = [(-1, 2), (0, 3), (1, 1)] if rule_name == "U1" else [(0, 2), (1, 2), (2, 1)]
    ps, os_ = spaces(rule_name, [(leg, False), (leg, True)], [(leg, False), (leg, False)])
    data = np.arange(1, os_.required_len + 1, dtype=np.float64)
    b = encode_tensor(ps, data)
    parsed = _parse_blocks(b, rule_name)
    assert len(parsed) == len(os_.blocks)
    keys = [(ct, dt) for ct, dt, _, _ in parsed]
    for ct, dt, shape, vals in parsed:
        blk = os_.blocks[os_.index[(ct, dt)]]
        assert shape == blk.shape
        want = []
        for lin in range(int(np.prod(blk.shape))):  # first axis fastest
            r, off = lin, blk.offset
            for d, st in zip(blk.shape, blk.strides):
                off += (r % d) * st
                r //= d
            want.append(data[off])
        assert np.array_equal(vals, np.array(want))
    assert len(set(keys)) == len(keys)


def test_malformed_snapshots_are_rejected_and_output_untouched():
    ps, os_ = su2_spaces()
    b = encode_tensor(ps, bit_fill(os_))
    for bad in (b[:-1], b + b"\0", b"XENETTS\0" + b[8:], b[:8] + b"\x02\x00" + b[10:], b[:10] + b"\x01" + b[11:]):
        with pytest.raises(N.TenetError):
            decode_tensor(bad)
    # block boundaries, then a missing and a duplicated block (counts kept consistent)
    pos = header_end(b) + 1
    nc, nd = u64_at(b, pos), u64_at(b, pos + 8)
    pos += 16
    for _ in range(nc + nd):
        pos += 1
        n = u64_at(b, pos)
        pos += 8
        for _ in range(n):
            pos += 8 + u64_at(b, pos) + 8
    nb = u64_at(b, pos)
    bounds, q = [], pos + 8
    for _ in range(nb):
        kl = u64_at(b, q)
        rank = u64_at(b, q + 8 + kl)
        cnt = u64_at(b, q + 16 + kl + 8 * rank)
        end = q + 8 + kl + 8 + 8 * rank + 8 + 8 * cnt
        bounds.append((q, end))
        q = end
    assert q == len(b) and nb >= 2
    blocks = [b[x:y] for x, y in bounds]
    head = b[:pos]
    missing = head + struct.pack("<Q", nb - 1) + b"".join(blocks[1:])
    duplicate = head + struct.pack("<Q", nb) + blocks[0] + blocks[0] + b"".join(blocks[2:])
    swapped = head + struct.pack("<Q", nb) + blocks[1] + blocks[0] + b"".join(blocks[2:])
    for bad in (missing, duplicate, swapped):
        with pytest.raises(N.TenetError):
            decode_tensor(bad)
    h = N.host()
    sp = C.c_void_p()
    # a failed payload decode leaves the caller's buffer untouched
    N.check_host(h.tnth_snapshot_tensor_space(b, len(b), C.byref(sp), None))
    space = HomSpace._adopt(Rule("SU2"), sp.value)
    out = np.full(space.len, 7.0)
    rc = h.tnth_snapshot_tensor_payload(b[:-1], len(b) - 1, space._h, out.ctypes.data_as(C.POINTER(C.c_double)),
                                        space.len)
    assert rc != N.OK and np.all(out == 7.0)
