"""Symmetry providers of the oracle (U(1), SU(2), bosonic Z2 and their products).

Sector ids follow the reference encodings exactly, because block order in
the coupled-sector layout is the order of sector ids:

* U(1): zigzag id  ((q << 1) ^ (q >> 31))  -- tenet-sectors/src/abelian.rs:570-576
* SU(2): id = 2j                           -- tenet-sectors/src/su2.rs:20-60
* Z2:    id = parity
* products: Cantor pairing of the component ids (TensorKitProductCodec)
                                           -- tenet-sectors/src/algebra.rs:1434-1479

Symbols follow the reference providers:

* U(1): F = R = dim = FS = 1               -- abelian.rs:723-780
* SU(2): F in Racah's Condon-Shortley gauge, TensorKit convention
  F(a,b,c,d,e,f) = sqrt((2e+1)(2f+1)) (-1)^(a+b+c+d) {a b e; c d f}
  (racah::su2_f_symbol called at su2.rs:300-317); R(a,b,c) = (-1)^(a+b-c)
  (su2.rs:319-344); dim = 2j+1 (su2.rs:346-349); FS = (-1)^(2j) (su2.rs:367-369).
* A/B symbols derived from F, dim and FS   -- tenet-sectors/src/algebra.rs:602-625
"""
from fractions import Fraction
from functools import lru_cache
import math


def zigzag(q: int) -> int:
    return 2 * q if q >= 0 else -2 * q - 1


def unzigzag(z: int) -> int:
    return (z >> 1) ^ -(z & 1)


class U1:
    """U(1) provider (abelian.rs:517-780)."""
    name = "U1"
    unique = True

    def vacuum(self):
        return 0

    def encode(self, q):
        return zigzag(int(q))

    def label(self, sid):
        return unzigzag(sid)

    def dual(self, a):
        return zigzag(-unzigzag(a))

    def channels(self, a, b):
        return [zigzag(unzigzag(a) + unzigzag(b))]

    def nsym(self, a, b, c):
        return 1 if unzigzag(a) + unzigzag(b) == unzigzag(c) else 0

    def F(self, a, b, c, d, e, f):
        return 1.0

    def R(self, a, b, c):
        return 1.0

    def dim(self, a):
        return 1.0

    def fs(self, a):
        return 1.0


class Z2:
    """Bosonic Z2 provider (trivial symbols)."""
    name = "Z2"
    unique = True

    def vacuum(self):
        return 0

    def encode(self, p):
        return int(p) & 1

    def label(self, sid):
        return sid

    def dual(self, a):
        return a

    def channels(self, a, b):
        return [a ^ b]

    def nsym(self, a, b, c):
        return 1 if (a ^ b) == c else 0

    def F(self, a, b, c, d, e, f):
        return 1.0

    def R(self, a, b, c):
        return 1.0

    def dim(self, a):
        return 1.0

    def fs(self, a):
        return 1.0


_FACT = [1]


def _fact(n):
    while len(_FACT) <= n:
        _FACT.append(_FACT[-1] * len(_FACT))
    return _FACT[n]


def _delta2(a, b, c):
    """Triangle coefficient squared, doubled-spin arguments, exact rational."""
    return Fraction(_fact((a + b - c) // 2) * _fact((a - b + c) // 2) * _fact((-a + b + c) // 2),
                    _fact((a + b + c) // 2 + 1))


def _triangle(a, b, c):
    return (a + b + c) % 2 == 0 and abs(a - b) <= c <= a + b


@lru_cache(maxsize=None)
def sixj_doubled(j1, j2, j3, j4, j5, j6):
    """Wigner 6j {j1 j2 j3; j4 j5 j6} for doubled spins, Racah's formula.

    The alternating sum is exact (integer/rational); only the final square
    root of the triangle coefficients is rounded.
    """
    if not (_triangle(j1, j2, j3) and _triangle(j1, j5, j6) and _triangle(j4, j2, j6)
            and _triangle(j4, j5, j3)):
        return 0.0
    a1 = (j1 + j2 + j3) // 2
    a2 = (j1 + j5 + j6) // 2
    a3 = (j4 + j2 + j6) // 2
    a4 = (j4 + j5 + j3) // 2
    b1 = (j1 + j2 + j4 + j5) // 2
    b2 = (j2 + j3 + j5 + j6) // 2
    b3 = (j3 + j1 + j6 + j4) // 2
    total = Fraction(0)
    for t in range(max(a1, a2, a3, a4), min(b1, b2, b3) + 1):
        num = _fact(t + 1)
        den = (_fact(t - a1) * _fact(t - a2) * _fact(t - a3) * _fact(t - a4)
               * _fact(b1 - t) * _fact(b2 - t) * _fact(b3 - t))
        total += Fraction((-1) ** t * num, den)
    if total == 0:
        return 0.0
    d2 = _delta2(j1, j2, j3) * _delta2(j1, j5, j6) * _delta2(j4, j2, j6) * _delta2(j4, j5, j3)
    sq = total * total * d2  # exact square of the 6j value
    mag = math.sqrt(sq.numerator) / math.sqrt(sq.denominator)
    return mag if total > 0 else -mag


class SU2:
    """SU(2) provider in the Condon-Shortley gauge (su2.rs:60-370)."""
    name = "SU2"
    unique = False

    def vacuum(self):
        return 0

    def encode(self, twoj):
        return int(twoj)

    def label(self, sid):
        return sid

    def dual(self, a):
        return a

    def channels(self, a, b):
        return list(range(abs(a - b), a + b + 1, 2))

    def nsym(self, a, b, c):
        return 1 if _triangle(a, b, c) else 0

    @lru_cache(maxsize=None)
    def F(self, a, b, c, d, e, f):
        if not (_triangle(a, b, e) and _triangle(e, c, d) and _triangle(b, c, f)
                and _triangle(a, f, d)):
            return 0.0
        if a == b == c == d == e == f == 0:
            return 1.0
        sign = -1.0 if ((a + b + c + d) // 2) % 2 else 1.0
        return math.sqrt((e + 1) * (f + 1)) * sign * sixj_doubled(a, b, e, c, d, f)

    def R(self, a, b, c):
        if not _triangle(a, b, c):
            return 0.0
        return -1.0 if ((a + b - c) // 2) % 2 else 1.0

    def dim(self, a):
        return float(a + 1)

    def fs(self, a):
        return -1.0 if a % 2 else 1.0


def cantor_encode(left, right):
    s = left + right
    return s * (s + 1) // 2 + left


def cantor_decode(sid):
    s = int((math.isqrt(8 * sid + 1) - 1) // 2)
    while s * (s + 1) // 2 > sid:
        s -= 1
    while (s + 1) * (s + 2) // 2 <= sid:
        s += 1
    left = sid - s * (s + 1) // 2
    return left, s - left


class Product:
    """Deligne product provider with the Cantor-pairing codec (product_rule.rs)."""

    def __init__(self, left, right):
        self.l = left
        self.r = right
        self.name = f"{left.name}x{right.name}"
        self.unique = left.unique and right.unique

    def vacuum(self):
        return cantor_encode(self.l.vacuum(), self.r.vacuum())

    def encode(self, label):
        return cantor_encode(self.l.encode(label[0]), self.r.encode(label[1]))

    def label(self, sid):
        a, b = cantor_decode(sid)
        return (self.l.label(a), self.r.label(b))

    def _split(self, sid):
        return cantor_decode(sid)

    def dual(self, a):
        x, y = self._split(a)
        return cantor_encode(self.l.dual(x), self.r.dual(y))

    def channels(self, a, b):
        (ax, ay), (bx, by) = self._split(a), self._split(b)
        out = [cantor_encode(x, y) for x in self.l.channels(ax, bx) for y in self.r.channels(ay, by)]
        return sorted(out)

    def nsym(self, a, b, c):
        (ax, ay), (bx, by), (cx, cy) = self._split(a), self._split(b), self._split(c)
        return self.l.nsym(ax, bx, cx) * self.r.nsym(ay, by, cy)

    def F(self, a, b, c, d, e, f):
        s = [self._split(x) for x in (a, b, c, d, e, f)]
        return (self.l.F(*[x[0] for x in s]) * self.r.F(*[x[1] for x in s]))

    def R(self, a, b, c):
        s = [self._split(x) for x in (a, b, c)]
        return self.l.R(*[x[0] for x in s]) * self.r.R(*[x[1] for x in s])

    def dim(self, a):
        x, y = self._split(a)
        return self.l.dim(x) * self.r.dim(y)

    def fs(self, a):
        x, y = self._split(a)
        return self.l.fs(x) * self.r.fs(y)


def sqrtdim(rule, a):
    return math.sqrt(rule.dim(a))


def a_symbol(rule, a, b, c):
    """A symbol (tenet-sectors/src/algebra.rs:602-612)."""
    factor = sqrtdim(rule, a) * sqrtdim(rule, b) / sqrtdim(rule, c)
    return factor * rule.fs(a) * rule.F(rule.dual(a), a, b, b, rule.vacuum(), c)


def b_symbol(rule, a, b, c):
    """B symbol (tenet-sectors/src/algebra.rs:614-624)."""
    return (sqrtdim(rule, a) * sqrtdim(rule, b) / sqrtdim(rule, c)
            * rule.F(a, b, rule.dual(b), a, c, rule.vacuum()))


def make_rule(name):
    if name == "U1":
        return U1()
    if name == "SU2":
        return SU2()
    if name == "Z2":
        return Z2()
    if name in ("U1xZ2", "U1Z2"):
        return Product(U1(), Z2())
    raise ValueError(name)
