"""Fusion trees, coupled-sector block layout and tree-pair recoupling (oracle).

Restates, in plain Python, the multiplicity-free parts of
tenet-core/src/fusion_tree.rs and fusion_space.rs that decide
(1) which blocks a TensorMap stores and where, and (2) the coefficients with
which a permutation of legs maps source fusion-tree pairs to destination
pairs.  TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).
"""
from collections import namedtuple
import itertools

from .symmetry import sqrtdim, b_symbol

Tree = namedtuple("Tree", "unc isdual coupled inner")


class Leg:
    """One leg: sectors sorted by id with degeneracies, plus the dual flag.

    Sectors are kept sorted by SectorId (tenet/src/typed.rs:7803-7823).
    """

    def __init__(self, pairs, is_dual=False):
        pairs = sorted((int(s), int(d)) for s, d in pairs)
        self.sectors = tuple(s for s, _ in pairs)
        self.degs = tuple(d for _, d in pairs)
        self.is_dual = bool(is_dual)
        self._deg = dict(pairs)

    def deg(self, s):
        return self._deg[s]

    def dim(self, rule):
        return sum(int(rule.dim(s)) * d for s, d in zip(self.sectors, self.degs))

    def dual(self, rule):
        return Leg([(rule.dual(s), d) for s, d in zip(self.sectors, self.degs)], not self.is_dual)

    def key(self):
        return (self.sectors, self.degs, self.is_dual)

    def __eq__(self, other):
        return self.key() == other.key()

    def __hash__(self):
        return hash(self.key())

    def __repr__(self):
        return f"Leg({list(zip(self.sectors, self.degs))}, dual={self.is_dual})"


def _visit_inner(rule, effective, coupled):
    """Inner-line lists (innermost first) of the trees of `effective` into `coupled`.

    Walks backward from the coupled sector like
    tenet-core/src/fusion_tree.rs:1076-1131 (visit_fusion_trees_where): the
    inner line adjacent to the last leg ranges over coupled (x) dual(last).
    """
    n = len(effective)
    if n == 0:
        return [[]] if coupled == rule.vacuum() else []
    if n == 1:
        return [[]] if effective[0] == coupled else []
    if n == 2:
        return [[]] if rule.nsym(effective[0], effective[1], coupled) else []
    last = effective[-1]
    out = []
    for a in rule.channels(coupled, rule.dual(last)):
        if not rule.nsym(a, last, coupled):
            continue
        for inner in _visit_inner(rule, effective[:-1], a):
            out.append(inner + [a] if n > 2 else inner)
    return out


def _reachable(rule, effective):
    """Coupled sectors reachable by fusing all legs (fusion_tree.rs:970-990)."""
    if not effective:
        return [rule.vacuum()]
    acc = [effective[0]]
    for s in effective[1:]:
        acc = sorted(set(c for f in acc for c in rule.channels(f, s)))
    return sorted(set(acc))


def trees_by_coupled(rule, legs):
    """Coupled sector -> ordered tree list for one product space.

    Leg tuples are visited with the first leg fastest
    (fusion_tree.rs:491-509), trees of each tuple are appended to their
    coupled group, and groups are sorted by coupled id (fusion_tree.rs:512-575).
    """
    groups = {}
    order = []
    sector_lists = [leg.sectors for leg in legs]
    isdual = tuple(leg.is_dual for leg in legs)
    # itertools.product varies the LAST factor fastest; reverse to make the
    # first leg fastest.
    for rev in itertools.product(*reversed(sector_lists)):
        tup = tuple(reversed(rev))
        for c in _reachable(rule, list(tup)):
            trees = [Tree(tup, isdual, c, tuple(inner)) for inner in _visit_inner(rule, list(tup), c)]
            if c not in groups:
                groups[c] = []
                order.append(c)
            groups[c].extend(trees)
    return {c: groups[c] for c in sorted(order) if groups[c]}


class Block:
    __slots__ = ("key", "offset", "shape", "strides", "sector_index")

    def __init__(self, key, offset, shape, strides, sector_index):
        self.key = key
        self.offset = offset
        self.shape = shape
        self.strides = strides
        self.sector_index = sector_index


class HomSpace:
    """Codomain <- domain space with its coupled-sector matrix layout.

    Keys: for each coupled sector present on both sides (ascending), domain
    trees outer and codomain trees inner (fusion_space.rs:2211-2261).  Block
    offsets/strides: column-major sector matrix, first leg fastest
    (fusion_space.rs:3953-4104).
    """

    def __init__(self, rule, codomain, domain):
        self.rule = rule
        self.codomain = tuple(codomain)
        self.domain = tuple(domain)
        self.nout = len(self.codomain)
        self.nin = len(self.domain)
        cod = trees_by_coupled(rule, self.codomain)
        dom = trees_by_coupled(rule, self.domain)
        self.blocks = []
        self.index = {}
        self.sectors = []  # (coupled, offset, rows, cols)
        offset = 0
        for c in sorted(set(cod) & set(dom)):
            ctrees, dtrees = cod[c], dom[c]
            rshapes = [tuple(self.codomain[i].deg(s) for i, s in enumerate(t.unc)) for t in ctrees]
            cshapes = [tuple(self.domain[i].deg(s) for i, s in enumerate(t.unc)) for t in dtrees]
            rdims = [_prod(s) for s in rshapes]
            cdims = [_prod(s) for s in cshapes]
            rstart = _cumsum(rdims)
            cstart = _cumsum(cdims)
            rows, cols = sum(rdims), sum(cdims)
            sidx = len(self.sectors)
            self.sectors.append((c, offset, rows, cols))
            for j, dt in enumerate(dtrees):
                for i, ct in enumerate(ctrees):
                    shape = rshapes[i] + cshapes[j]
                    strides = []
                    st = 1
                    for d in rshapes[i]:
                        strides.append(st)
                        st *= d
                    st = rows
                    for d in cshapes[j]:
                        strides.append(st)
                        st *= d
                    b = Block((ct, dt), offset + rstart[i] + rows * cstart[j], shape, tuple(strides), sidx)
                    self.index[(ct, dt)] = len(self.blocks)
                    self.blocks.append(b)
            offset += rows * cols
        self.required_len = offset

    @property
    def rank(self):
        return self.nout + self.nin

    def legs_linear(self):
        """All legs in codomain orientation: codomain legs then dual(domain legs)."""
        return list(self.codomain) + [leg.dual(self.rule) for leg in self.domain]

    def leg_as_codomain(self, axis):
        if axis < self.nout:
            return self.codomain[axis]
        return self.domain[axis - self.nout].dual(self.rule)

    def permuted(self, p1, p2):
        """Destination space of permute(t, p1, p2)."""
        legs = [self.leg_as_codomain(a) for a in range(self.rank)]
        return HomSpace(self.rule, [legs[a] for a in p1], [legs[a].dual(self.rule) for a in p2])

    def adjoint(self):
        return HomSpace(self.rule, self.domain, self.codomain)

    def same_as(self, other):
        return (self.rule.name == other.rule.name and self.codomain == other.codomain
                and self.domain == other.domain)


def _prod(xs):
    p = 1
    for x in xs:
        p *= x
    return p


def _cumsum(xs):
    out, s = [], 0
    for x in xs:
        out.append(s)
        s += x
    return out


# ---------------------------------------------------------------------------
# Tree-pair operations (multiplicity-free).
# ---------------------------------------------------------------------------

def _ext(tree, k):
    """Extended inner line k: uncoupled[0], inner..., coupled (fusion_tree.rs:8256-8273)."""
    n = len(tree.unc)
    if k == 0:
        return tree.unc[0]
    if k == n - 1:
        return tree.coupled
    return tree.inner[k - 1]


def artin_braid(rule, tree, index, inverse):
    """One Artin generator on a fusion tree (fusion_tree.rs:8088-8255)."""
    n = len(tree.unc)
    left, right = tree.unc[index], tree.unc[index + 1]
    unc = list(tree.unc)
    unc[index], unc[index + 1] = unc[index + 1], unc[index]
    isd = list(tree.isdual)
    isd[index], isd[index + 1] = isd[index + 1], isd[index]
    unc, isd = tuple(unc), tuple(isd)
    vac = rule.vacuum()
    if left == vac or right == vac:
        inner = list(tree.inner)
        if index > 0:
            out_tree = Tree(unc, isd, tree.coupled, tree.inner)
            src = _ext(out_tree, index + 1) if left == vac else _ext(out_tree, index - 1)
            inner[index - 1] = src
        return [(Tree(unc, isd, tree.coupled, tuple(inner)), 1.0)]
    if index == 0:
        coupled = tree.inner[0] if n > 2 else tree.coupled
        coeff = rule.R(right, left, coupled) if inverse else rule.R(left, right, coupled)
        return [(Tree(unc, isd, tree.coupled, tree.inner), coeff)]
    a = _ext(tree, index - 1)
    b = left
    c = _ext(tree, index)
    d = right
    e = _ext(tree, index + 1)
    terms = []
    for cp in rule.channels(a, d):
        if not rule.nsym(cp, b, e):
            continue
        inner = list(tree.inner)
        inner[index - 1] = cp
        f = rule.F(d, a, b, e, cp, c)
        if inverse:
            coeff = rule.R(d, c, e) * f * rule.R(d, a, cp)
        else:
            coeff = rule.R(c, d, e) * f * rule.R(a, d, cp)
        terms.append((Tree(unc, isd, tree.coupled, tuple(inner)), coeff))
    return terms


def _accumulate(terms):
    """Sum coefficients of equal keys, first-seen order (FusionTermAccumulator)."""
    acc, order = {}, []
    for k, v in terms:
        if k in acc:
            acc[k] += v
        else:
            acc[k] = v
            order.append(k)
    return [(k, acc[k]) for k in order]


def braid_steps(perm, levels):
    """Artin steps of a permutation (fusion_tree.rs:2477-2505, PreparedTreeBraid::new)."""
    rank = len(perm)
    work = list(perm)
    lev = list(levels)
    steps = []
    for target in range(rank - 1):
        source = work[target]
        for index in range(source - 1, target - 1, -1):
            steps.append((index, lev[index] > lev[index + 1]))
            lev[index], lev[index + 1] = lev[index + 1], lev[index]
        for i in range(target + 1, rank):
            if work[i] < source:
                work[i] += 1
        work[target] = target
    return steps


def braid_tree(rule, tree, steps):
    current = [(tree, 1.0)]
    for index, inverse in steps:
        nxt = []
        for t, c in current:
            for t2, c2 in artin_braid(rule, t, index, inverse):
                nxt.append((t2, c * c2))
        current = _accumulate(nxt)
    return current


def bendright(rule, pair):
    """Move the last codomain leg to the end of the domain (fusion_tree.rs:8749-8880, 9000-9020)."""
    f1, f2 = pair
    n1, n2 = len(f1.unc), len(f2.unc)
    b = f1.unc[-1]
    bdual = f1.isdual[-1]
    c = f1.coupled
    if n1 == 1:
        a = rule.vacuum()
    elif n1 == 2:
        a = f1.unc[0]
    else:
        a = f1.inner[-1]
    nf1 = Tree(f1.unc[:-1], f1.isdual[:-1], a, f1.inner[:-1] if n1 > 2 else ())
    nf2 = Tree(f2.unc + (rule.dual(b),), f2.isdual + (not bdual,), a,
               f2.inner + ((c,) if n2 > 1 else ()))
    coeff = sqrtdim(rule, c) / sqrtdim(rule, a) * b_symbol(rule, a, b, c)
    if bdual:
        coeff *= rule.fs(rule.dual(b))
    return (nf1, nf2), coeff


def bendleft(rule, pair):
    """Move the last domain leg to the end of the codomain (fusion_tree.rs:8927-8995, 9024-9040)."""
    (g1, g2), coeff = bendright(rule, (pair[1], pair[0]))
    return (g2, g1), coeff


def repartition(rule, pair, target):
    """Bend legs until the codomain has `target` legs (fusion_tree.rs:3785-3812)."""
    coeff = 1.0
    n1 = len(pair[0].unc)
    while n1 < target:
        pair, c = bendleft(rule, pair)
        coeff *= c
        n1 += 1
    while n1 > target:
        pair, c = bendright(rule, pair)
        coeff *= c
        n1 -= 1
    return pair, coeff


def linearized_permutation(p1, p2, n1, n2):
    """fusion_tree.rs:7097-7130 and :12537-12543."""
    def lin(axis):
        return axis if axis < n1 else n2 + 2 * n1 - 1 - axis
    out = []
    for pos in range(n1 + n2):
        if pos < len(p1):
            axis = p1[pos]
        else:
            axis = p2[len(p2) - 1 - (pos - len(p1))]
        out.append(lin(axis))
    return out


def permute_pair(rule, pair, p1, p2):
    """Coefficient expansion of permute(f1, f2; p1, p2) (fusion_tree.rs:3278-3420).

    Repartition everything into the codomain, braid with levels
    (0..n1, reversed domain), repartition to len(p1).
    """
    n1, n2 = len(pair[0].unc), len(pair[1].unc)
    perm = linearized_permutation(p1, p2, n1, n2)
    if list(p1) == list(range(n1)) and list(p2) == list(range(n1, n1 + n2)):
        return [(pair, 1.0)]
    if perm == list(range(n1 + n2)):
        q, c = repartition(rule, pair, len(p1))
        return [(q, c)]
    allc, c0 = repartition(rule, pair, n1 + n2)
    levels = list(range(n1)) + list(range(n1, n1 + n2))[::-1]
    steps = braid_steps(perm, levels)
    out = []
    for t, c in braid_tree(rule, allc[0], steps):
        q, c2 = repartition(rule, (t, allc[1]), len(p1))
        out.append((q, c0 * c * c2))
    return _accumulate(out)


# ---------------------------------------------------------------------------
# Transform group planning (tenet-tensors/src/tree_transform/plan.rs:1489-1540,
# 2107-2200): blocks sharing uncoupled sectors form one group; injective
# 1:1 groups become Single specs, everything else a Multi spec with U[dst,src].
# ---------------------------------------------------------------------------

def group_key(pair):
    return (pair[0].unc, pair[1].unc, pair[0].isdual, pair[1].isdual)


def transform_specs(src_space, dst_space, p1, p2):
    """List of ('single', dst_block, src_block, coeff) / ('multi', dst_blocks, src_blocks, U)."""
    rule = src_space.rule
    groups, order = {}, []
    for bi, blk in enumerate(src_space.blocks):
        g = group_key(blk.key)
        if g not in groups:
            groups[g] = []
            order.append(g)
        groups[g].append(bi)
    specs = []
    for g in order:
        src_idx = groups[g]
        rows = [permute_pair(rule, src_space.blocks[i].key, p1, p2) for i in src_idx]
        dst_keys, dpos = [], {}
        direct = True
        seen = set()
        for r in rows:
            if len(r) != 1 or r[0][0] in seen:
                direct = False
            if len(r) == 1:
                seen.add(r[0][0])
            for k, _ in r:
                if k not in dpos:
                    dpos[k] = len(dst_keys)
                    dst_keys.append(k)
        for k in dst_keys:
            if k not in dst_space.index:
                raise KeyError(f"destination key {k} missing from destination space")
        if direct:
            for i, r in zip(src_idx, rows):
                specs.append(("single", dst_space.index[r[0][0]], i, r[0][1]))
        else:
            U = [[0.0] * len(src_idx) for _ in dst_keys]
            for col, r in enumerate(rows):
                for k, c in r:
                    U[dpos[k]][col] += c
            specs.append(("multi", [dst_space.index[k] for k in dst_keys], list(src_idx), U))
    return specs
