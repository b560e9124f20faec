"""Block-sparse tensors of the oracle: fill, permute (replay), compose, contract.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  Everything is FP64
numpy; block views are taken with explicit offsets/strides so the storage
order is exactly the reference coupled-sector layout.
"""
import numpy as np

from .fusion import HomSpace, Leg, transform_specs
from .symmetry import make_rule

MASK64 = (1 << 64) - 1


def splitmix_raw(seed, n):
    """The first n 64-bit outputs of splitmix64 started at `seed` (the stream of
    tenet/src/typed.rs:2086-2092): state_k = seed + k * 0x9E3779B97F4A7C15,
    output = mix(state_k), k = 1..n."""
    k = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & MASK64) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def splitmix_unit(seed, n):
    """n draws of the reference seeded fill (tenet/src/typed.rs:2086-2092), vectorised:
    unit = (output >> 11) / 2^52 - 1   in [-1, 1).
    """
    z = splitmix_raw(seed, n)
    return (z >> np.uint64(11)).astype(np.float64) / float(1 << 52) - 1.0


# Constants of the normal fill's own log / sin / cos (mathematics only):
# ln 2 split so that e * LN2_HI is exact for |e| <= 1074 (low 21 bits of the
# double nearest ln 2 cleared), the atanh series coefficients 1/(2k+1), and the
# Taylor coefficients (-1)^k / (2k+1)! and (-1)^k / (2k)! of sin and cos.
LN2_HI = 0.6931471803691238        # 0x3FE62E42FEE00000
LN2_LO = 1.9082149292705877e-10    # ln 2 - LN2_HI
SQRT_HALF = 0.7071067811865476
PI_HALF = 1.5707963267948966
ATANH_C = [1.0 / (2 * k + 1) for k in range(1, 12)]   # z^1 .. z^11 of atanh(s)/s - 1


def _fact(n):
    f = 1
    for i in range(2, n + 1):
        f *= i
    return f


SIN_C = [(-1) ** k / _fact(2 * k + 1) for k in range(1, 10)]   # z^1 .. z^9 of sin(t)/t - 1
COS_C = [(-1) ** k / _fact(2 * k) for k in range(1, 10)]       # z^1 .. z^9 of cos(t) - 1


def _horner(z, coeffs):
    """sum_k coeffs[k] z^(k+1) evaluated as z*(c0 + z*(c1 + ...)), one rounding per op."""
    p = np.full_like(z, coeffs[-1])
    for c in reversed(coeffs[:-1]):
        p = p * z + c
    return p * z


def log_unit(u):
    """ln(u) for u in (0, 1] from +, -, *, / only (no fused multiply-add).

    u = m 2^e with m in [sqrt(1/2), sqrt(2)); ln m = 2 atanh(s), s = (m-1)/(m+1),
    |s| < 0.172, series to s^23 (truncation < 2^-60 relative); then
    ln u = e LN2_HI + (ln m + e LN2_LO).  Every operation is a correctly rounded
    IEEE-754 double operation, so the device fill evaluating the same sequence
    (with contraction off) gives the same bits.  Accuracy: within 2 ulp of ln.
    """
    m, e = np.frexp(u)
    low = m < SQRT_HALF
    m = np.where(low, m * 2.0, m)
    e = np.where(low, e - 1, e).astype(np.float64)
    f = m - 1.0
    s = f / (m + 1.0)
    s2 = s * 2.0
    lnm = s2 + s2 * _horner(s * s, ATANH_C)
    return e * LN2_HI + (lnm + e * LN2_LO)


def sincos_2pi(t):
    """(sin 2 pi t, cos 2 pi t) for t in (0, 1] from +, -, *, rint only.

    q = rint(4t), x = 4t - q in [-1/2, 1/2] (exact), a = x pi/2 in [-pi/4, pi/4];
    Taylor series of sin and cos to a^19 / a^18 (truncation < 2^-60); the
    quadrant q mod 4 rotates (sin a, cos a) by exact sign/swap.
    """
    t4 = t * 4.0
    q = np.rint(t4)
    x = t4 - q
    a = x * PI_HALF
    z = a * a
    sa = a + a * _horner(z, SIN_C)
    ca = 1.0 + _horner(z, COS_C)
    qm = np.mod(q, 4.0)
    sin = np.where(qm == 0, sa, np.where(qm == 1, ca, np.where(qm == 2, -sa, -ca)))
    cos = np.where(qm == 0, ca, np.where(qm == 1, -sa, np.where(qm == 2, -ca, sa)))
    return sin, cos


def splitmix_normal(seed, n):
    """normal(0,1) by Box-Muller on consecutive draw pairs of the same stream.

    Elements 2p and 2p+1 use draws (2p+1, 2p+2) of the seeded splitmix64
    stream (tenet/src/typed.rs:2086-2092) mapped to u = ((z >> 11) + 1) 2^-53
    in (0, 1]:  r = sqrt(-2 ln u1),  (r cos 2 pi u2, r sin 2 pi u2).  ln, sin and
    cos are this module's log_unit / sincos_2pi, so that the device fill
    (tnt_buf_fill_random with normal != 0) is bit-identical.  The reference has
    no normal fill; this definition serves the synthetic workloads.
    """
    m = 2 * ((n + 1) // 2)
    z = splitmix_raw(seed, m)
    u = ((z >> np.uint64(11)).astype(np.float64) + 1.0) * 2.0 ** -53
    u1, u2 = u[0::2], u[1::2]
    r = np.sqrt(log_unit(u1) * -2.0)
    sn, cs = sincos_2pi(u2)
    out = np.empty(m)
    out[0::2] = r * cs
    out[1::2] = r * sn
    return out[:n]


class OTensor:
    def __init__(self, space: HomSpace, data=None):
        self.space = space
        self.rule = space.rule
        if data is None:
            data = np.zeros(space.required_len)
        self.data = np.asarray(data, dtype=np.float64)
        assert self.data.shape == (space.required_len,)

    # -- construction ------------------------------------------------------
    @classmethod
    def zeros(cls, space):
        return cls(space)

    @classmethod
    def rand(cls, space, seed):
        return cls(space, splitmix_unit(seed, space.required_len))

    @classmethod
    def randn(cls, space, seed):
        return cls(space, splitmix_normal(seed, space.required_len))

    @classmethod
    def from_block_fn(cls, space, fn):
        """fn(key, idx) per stored element, idx = degeneracy index per axis."""
        t = cls(space)
        for blk in space.blocks:
            view = t.block(blk)
            for idx in np.ndindex(*blk.shape):
                view[idx] = fn(blk.key, idx)
        return t

    def block(self, blk):
        """Strided numpy view of one fusion-tree block (column-major strides)."""
        return np.lib.stride_tricks.as_strided(
            self.data[blk.offset:], shape=blk.shape,
            strides=tuple(8 * s for s in blk.strides), writeable=True)

    def sector_matrix(self, si):
        c, off, rows, cols = self.space.sectors[si]
        return self.data[off:off + rows * cols].reshape(cols, rows).T

    # -- reductions (tenet/src/typed.rs:699-716, 11210-11262) -------------
    def inner(self, other):
        total = 0.0
        for si, (c, off, rows, cols) in enumerate(self.space.sectors):
            a = self.data[off:off + rows * cols]
            b = other.data[off:off + rows * cols]
            total += self.rule.dim(c) * float(np.dot(a, b))
        return total

    def norm(self):
        return float(np.sqrt(self.inner(self)))

    def tr(self):
        total = 0.0
        for si, (c, off, rows, cols) in enumerate(self.space.sectors):
            total += self.rule.dim(c) * float(np.trace(self.sector_matrix(si)))
        return total

    # -- linear algebra ----------------------------------------------------
    def add(self, other, alpha=1.0, beta=1.0):
        """alpha * self + beta * other on identical spaces."""
        assert self.space.same_as(other.space)
        return OTensor(self.space, alpha * self.data + beta * other.data)

    def adjoint(self):
        """Real adjoint: sector matrices transposed, codomain and domain swapped."""
        sp = self.space.adjoint()
        out = OTensor(sp)
        for si, (c, off, rows, cols) in enumerate(sp.sectors):
            src_si = [s[0] for s in self.space.sectors].index(c)
            out.sector_matrix(si)[:, :] = self.sector_matrix(src_si).T
        return out

    def compose(self, other):
        """Per coupled sector C_c = A_c B_c (fusion_replay.rs:457-555, 628-734)."""
        assert self.space.domain == other.space.codomain, "compose: spaces do not match"
        sp = HomSpace(self.rule, self.space.codomain, other.space.domain)
        out = OTensor(sp)
        a_idx = {s[0]: i for i, s in enumerate(self.space.sectors)}
        b_idx = {s[0]: i for i, s in enumerate(other.space.sectors)}
        for si, (c, off, rows, cols) in enumerate(sp.sectors):
            if c in a_idx and c in b_idx:
                out.sector_matrix(si)[:, :] = self.sector_matrix(a_idx[c]) @ other.sector_matrix(b_idx[c])
        return out

    def permute(self, p1, p2):
        """TensorKit permute via the group plan and Single/Multi replay.

        Replay semantics of tenet-operations/src/transform_replay.rs:3413-3606:
        Single blocks are a scaled strided copy; Multi groups pack the source
        blocks (in destination axis order) as columns, multiply by U^T, and
        scatter columns to the destination blocks.
        """
        dst_space = self.space.permuted(p1, p2)
        out = OTensor(dst_space)
        self.permute_into(out, p1, p2, 1.0, 0.0)
        return out

    def permute_into(self, out, p1, p2, alpha, beta):
        perm = list(p1) + list(p2)
        specs = transform_specs(self.space, out.space, p1, p2)
        touched = np.zeros(len(out.space.blocks), dtype=bool)
        new = np.zeros_like(out.data)
        newt = OTensor(out.space, new)
        for spec in specs:
            if spec[0] == "single":
                _, di, si, coeff = spec
                touched[di] = True
                dst = newt.block(out.space.blocks[di])
                src = self.block(self.space.blocks[si])
                dst[...] = coeff * np.transpose(src, perm)
            else:
                _, dis, sis, U = spec
                U = np.asarray(U)
                cols = [np.transpose(self.block(self.space.blocks[s]), perm) for s in sis]
                for r, di in enumerate(dis):
                    touched[di] = True
                    acc = np.zeros(out.space.blocks[di].shape)
                    for c, col in enumerate(cols):
                        acc += U[r, c] * col
                    newt.block(out.space.blocks[di])[...] = acc
        # Touched blocks: alpha * new + beta * old; inactive blocks: beta * old
        # (transform_replay.rs:288-330).
        result = beta * out.data if beta != 0.0 else np.zeros_like(out.data)
        for bi, blk in enumerate(out.space.blocks):
            if touched[bi]:
                view_r = OTensor(out.space, result).block(blk)
                view_r[...] += alpha * newt.block(blk)
        out.data[:] = result

    def contract(self, axes_a, other, axes_b, out_axes=None, out_nout=None):
        """TensorKit contraction: permute both operands, compose, permute output.

        Route of tenet-tensors/src/contract/dynamic.rs:1241-1530.
        """
        ra, rb = self.space.rank, other.space.rank
        open_a = [i for i in range(ra) if i not in axes_a]
        open_b = [i for i in range(rb) if i not in axes_b]
        a2 = self.permute(open_a, list(axes_a))
        b2 = other.permute(list(axes_b), open_b)
        c = a2.compose(b2)
        n = len(open_a) + len(open_b)
        if out_axes is None:
            out_axes = list(range(n))
        if out_nout is None:
            out_nout = len(open_a)
        p1, p2 = list(out_axes[:out_nout]), list(out_axes[out_nout:])
        if p1 == list(range(len(open_a))) and p2 == list(range(len(open_a), n)):
            return c
        return c.permute(p1, p2)

    def to_dense(self):
        """Dense array in the uncoupled basis -- only meaningful for abelian rules."""
        legs = self.space.legs_linear()
        offs = []
        for leg in legs:
            # Embed every axis in the basis of its non-dual underlying space so
            # that a V* axis (sector dual(s)) lines up with the V axis (sector s).
            base = leg.dual(self.rule) if leg.is_dual else leg
            o, acc = {}, 0
            for s, d in zip(base.sectors, base.degs):
                o[self.rule.dual(s) if leg.is_dual else s] = acc
                acc += d
            offs.append((o, acc))
        dense = np.zeros([a for _, a in offs])
        nout = self.space.nout
        for blk in self.space.blocks:
            ct, dt = blk.key
            secs = list(ct.unc) + [self.rule.dual(s) for s in dt.unc]
            sl = tuple(slice(offs[i][0][s], offs[i][0][s] + blk.shape[i]) for i, s in enumerate(secs))
            dense[sl] = self.block(blk)
        return dense


def space(rule, cod, dom):
    return HomSpace(rule, cod, dom)


def leg(rule, labels_degs, is_dual=False):
    return Leg([(rule.encode(l), d) for l, d in labels_degs], is_dual)


__all__ = ["OTensor", "splitmix_unit", "space", "leg", "make_rule", "HomSpace", "Leg"]
