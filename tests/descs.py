"""Shared test helpers: build matching oracle / product spaces and replay
compiled descriptors on the CPU with the oracle's replay (checker only)."""
import ctypes as C

import numpy as np

from oracle.fusion import HomSpace as OHomSpace, Leg as OLeg
from oracle.replay import replay
from oracle.symmetry import make_rule
from tenet_amd import _native as N
from tenet_amd.tensor import GradedSpace, HomSpace, Rule


def legs(rule_name, spec):
    """spec: list of (pairs, dual) with labels -> (product legs, oracle legs)."""
    pr = Rule(rule_name)
    orule = make_rule(rule_name)
    P, O = [], []
    for pairs, dual in spec:
        g = GradedSpace(pr, pairs, dual)
        P.append(g)
        O.append(OLeg([(orule.encode(l), d) for l, d in pairs], dual))
    return pr, orule, P, O


def spaces(rule_name, cod_spec, dom_spec):
    pr, orule, P, O = legs(rule_name, list(cod_spec) + list(dom_spec))
    n = len(cod_spec)
    return HomSpace(pr, P[:n], P[n:]), OHomSpace(orule, O[:n], O[n:])


def desc_to_specs(desc, coeffs=None):
    d = desc.contents
    lay = []
    for i in range(d.n_layouts):
        L = d.layouts[i]
        lay.append((L.offset, tuple(L.shape[:L.rank]), tuple(L.strides[:L.rank])))
    if coeffs is None:
        coeffs = [d.coeffs[i] for i in range(d.n_coeffs)]
    specs = []
    for i in range(d.n_blocks):
        b = d.blocks[i]
        if b.kind == 0:
            specs.append(("single", b.dst_layout, b.src_layout, coeffs[b.coeff]))
        else:
            n = b.src_count * b.dst_count
            specs.append(("multi", list(range(b.dst_layout, b.dst_layout + b.dst_count)),
                          list(range(b.src_layout, b.src_layout + b.src_count)),
                          coeffs[b.coeff:b.coeff + n]))
    inactive = [d.inactive[i] for i in range(d.n_inactive)]
    return lay, specs, inactive, d.dst_len, d.src_len


def replay_desc(desc, src, dst=None, alpha=1.0, beta=0.0, overwrite=True, coeffs=None):
    """CPU replay of a compiled tnt_transform_desc (oracle replay semantics);
    `coeffs` replaces the descriptor's coefficient array."""
    lay, specs, inactive, dlen, slen = desc_to_specs(desc, coeffs)
    # one layout table: replay() indexes dst/src layouts separately, so pass
    # the same table for both and renumber the scale-only layouts
    if dst is None:
        dst = np.zeros(dlen)
    touched_layouts = set()
    for s in specs:
        if s[0] == "single":
            touched_layouts.add(s[1])
        else:
            touched_layouts.update(s[1])
    # dst layout list for replay(): all layouts used as destinations
    dst_ids = sorted(touched_layouts | set(inactive))
    remap = {lid: i for i, lid in enumerate(dst_ids)}
    dst_layouts = [lay[i] for i in dst_ids]
    rspecs = []
    for s in specs:
        if s[0] == "single":
            rspecs.append(("single", remap[s[1]], s[2], s[3]))
        else:
            rspecs.append(("multi", [remap[x] for x in s[1]], s[2], s[3]))
    return replay(dst, src, dst_layouts, lay, rspecs, alpha, beta, overwrite=overwrite)


def contract_desc_replay(cdesc, lhs, rhs, core_len):
    """CPU replay of a tnt_contract_desc: transforms, per-job GEMM, out transform."""
    d = cdesc.contents
    L = replay_desc(d.lhs, lhs) if d.lhs else lhs
    R = replay_desc(d.rhs, rhs) if d.rhs else rhs
    core = np.zeros(core_len)
    for i in range(d.n_jobs):
        j = d.jobs[i]
        m, k, n = j.rows, j.contracted, j.cols
        if d.lhs_op != N.OP_IDENTITY:
            # stored X is k x m column-major; op(X) = X^T is its (m, k) row-major view
            A = L[j.lhs_offset:j.lhs_offset + m * k].reshape(m, k)
        else:
            A = L[j.lhs_offset:j.lhs_offset + m * k].reshape(k, m).T
        if d.rhs_op != N.OP_IDENTITY:
            B = R[j.rhs_offset:j.rhs_offset + k * n].reshape(k, n)
        else:
            B = R[j.rhs_offset:j.rhs_offset + k * n].reshape(n, k).T
        core[j.dst_offset:j.dst_offset + m * n] = (A @ B).T.reshape(-1)
    if d.out:
        return replay_desc(d.out, core)
    return core


def ospace(hs):
    """Oracle HomSpace with the same legs as a product HomSpace (ids are shared)."""
    orule = make_rule(hs.rule.name)
    mk = lambda g: OLeg(list(zip(g.ids, g.degs)), g.is_dual)
    return OHomSpace(orule, [mk(g) for g in hs.codomain], [mk(g) for g in hs.domain])
