"""Python mirror of the reference's typed facade on top of the gfx950 backend.

Names follow tenet/src/typed.rs and tenet/src/runtime.rs: ``Runtime``
(one backend context = one leased executor, runtime.rs:802-840),
``GradedSpace`` (a leg), ``TensorMap`` with ``permute`` (typed.rs:13163),
``contract`` (typed.rs:13653), ``compose`` (typed.rs:13842), ``adjoint``,
``add``, ``inner``/``norm`` (typed.rs:699-716), ``tr``.  Storage is a flat FP64
payload in HBM in the reference's coupled-sector layout; every operation on
it runs through libtenet_hip.so.  Structure (trees, offsets, coefficients)
comes from libtenet_host.so and is compiled once per (space, operation) and
cached on the Runtime, like the reference's structure caches.
"""
import ctypes as C
import math

import numpy as np

from . import _native as N
from ._native import TenetError


# ---------------------------------------------------------------------------
# Sectors and legs
# ---------------------------------------------------------------------------
class Rule:
    """A fusion rule of the reference's built-in providers: U1, SU2, Z2, U1xZ2."""

    def __init__(self, name):
        if name not in N.RULES:
            raise ValueError(f"unknown rule {name}")
        self.name = name
        self.id = N.RULES[name]

    def encode(self, label):
        h = N.host()
        if self.name == "U1":
            return int(h.tnth_u1_id(int(label)))
        if self.name == "SU2":
            return int(label)  # doubled spin
        if self.name == "Z2":
            return int(label) & 1
        q, p = label
        return int(h.tnth_product_id(h.tnth_u1_id(int(q)), int(p) & 1))

    def label(self, sid):
        h = N.host()
        if self.name == "U1":
            return int(h.tnth_u1_charge(sid))
        if self.name in ("SU2", "Z2"):
            return int(sid)
        l, r = C.c_uint64(), C.c_uint64()
        h.tnth_product_split(sid, C.byref(l), C.byref(r))
        return (int(h.tnth_u1_charge(l.value)), int(r.value))

    def dim(self, sid):
        return float(N.host().tnth_symbol_dim(self.id, sid))

    def dual(self, sid):
        return int(N.host().tnth_symbol_dual(self.id, sid))

    def __eq__(self, o):
        return isinstance(o, Rule) and o.name == self.name

    def __hash__(self):
        return hash(self.name)

    def __repr__(self):
        return f"Rule({self.name})"


def _take_bytes(handle):
    """Copy and free a tnth_bytes buffer."""
    h = N.host()
    try:
        n = h.tnth_bytes_size(handle)
        return C.string_at(h.tnth_bytes_data(handle), n) if n else b""
    finally:
        h.tnth_bytes_free(handle)


class GradedSpace:
    """One leg: (sector, degeneracy) pairs and a dual flag (typed.rs:7795-7830)."""

    def __init__(self, rule, pairs, dual=False, _ids=False):
        self.rule = rule if isinstance(rule, Rule) else Rule(rule)
        enc = [(int(s) if _ids else self.rule.encode(s), int(d)) for s, d in pairs]
        enc.sort()
        ids = [s for s, _ in enc]
        if len(set(ids)) != len(ids):
            raise ValueError("sector declared more than once")
        self.ids = tuple(ids)
        self.degs = tuple(d for _, d in enc)
        self.is_dual = bool(dual)

    def dual(self):
        return GradedSpace(self.rule, [(self.rule.dual(s), d) for s, d in zip(self.ids, self.degs)],
                           not self.is_dual, _ids=True)

    def dim(self):
        return int(sum(self.rule.dim(s) * d for s, d in zip(self.ids, self.degs)))

    def key(self):
        return (self.rule.name, self.ids, self.degs, self.is_dual)

    def __eq__(self, o):
        return isinstance(o, GradedSpace) and o.key() == self.key()

    def __hash__(self):
        return hash(self.key())

    def to_bytes(self):
        """Space snapshot in TeNeT's v1 wire format (GradedSpace::to_bytes_with,
        tenet/src/typed/serialization.rs:1700-1712)."""
        h = N.host()
        b = C.c_void_p()
        n = len(self.ids)
        N.check_host(h.tnth_snapshot_leg(self.rule.id, n, (C.c_uint64 * max(1, n))(*self.ids),
                                         (C.c_int64 * max(1, n))(*self.degs), int(self.is_dual), C.byref(b)))
        return _take_bytes(b)

    @classmethod
    def from_bytes(cls, data):
        """Inverse of to_bytes (GradedSpace::from_bytes_with)."""
        h = N.host()
        data = bytes(data)
        rule, n, dual = C.c_int(), C.c_size_t(), C.c_int()
        N.check_host(h.tnth_snapshot_leg_decode(data, len(data), C.byref(rule), C.byref(n), None, None, None))
        ids, degs = (C.c_uint64 * max(1, n.value))(), (C.c_int64 * max(1, n.value))()
        N.check_host(h.tnth_snapshot_leg_decode(data, len(data), None, None, ids, degs, C.byref(dual)))
        name = {v: k for k, v in N.RULES.items()}[rule.value]
        return cls(Rule(name), [(ids[i], degs[i]) for i in range(n.value)], bool(dual.value), _ids=True)

    def __repr__(self):
        pairs = [(self.rule.label(s), d) for s, d in zip(self.ids, self.degs)]
        return f"GradedSpace({self.rule.name}, {pairs}{', dual' if self.is_dual else ''})"


class HomSpace:
    """codomain <- domain with its coupled-sector block layout (host structure)."""

    def __init__(self, rule, codomain, domain, _handle=None, _owner=None):
        self.rule = rule if isinstance(rule, Rule) else Rule(rule)
        self._owner = _owner
        h = N.host()
        if _handle is None:
            legs = list(codomain) + list(domain)
            ns = [len(l.ids) for l in legs]
            ids = [s for l in legs for s in l.ids]
            degs = [d for l in legs for d in l.degs]
            dual = [1 if l.is_dual else 0 for l in legs]
            out = C.c_void_p()
            N.check_host(h.tnth_space_create(
                self.rule.id, len(codomain), len(domain), N.sz_array(ns),
                (C.c_uint64 * max(1, len(ids)))(*ids), (C.c_int64 * max(1, len(degs)))(*degs),
                (C.c_uint8 * max(1, len(dual)))(*dual), C.byref(out)))
            self._h = out
            self._own = True
            self.codomain = tuple(codomain)
            self.domain = tuple(domain)
        else:
            self._h = C.c_void_p(_handle)
            self._own = False
            nout, nin = h.tnth_space_nout(self._h), h.tnth_space_nin(self._h)
            legs = [self._leg(a) for a in range(nout + nin)]
            self.codomain = tuple(legs[:nout])
            self.domain = tuple(legs[nout:])
        self.nout = len(self.codomain)
        self.nin = len(self.domain)
        self.len = int(h.tnth_space_len(self._h))
        self._sectors = None
        self._key = None
        self._adj = None

    def _leg(self, axis):
        h = N.host()
        n = C.c_size_t()
        N.check_host(h.tnth_space_leg(self._h, axis, C.byref(n), None, None, None))
        ids = (C.c_uint64 * max(1, n.value))()
        degs = (C.c_int64 * max(1, n.value))()
        dual = C.c_uint8()
        N.check_host(h.tnth_space_leg(self._h, axis, None, ids, degs, C.byref(dual)))
        return GradedSpace(self.rule, [(ids[i], degs[i]) for i in range(n.value)], bool(dual.value),
                           _ids=True)

    @classmethod
    def _adopt(cls, rule, handle):
        """Wrap a space handle returned by the host library, taking ownership."""
        sp = cls(rule, None, None, _handle=handle)
        sp._own = True
        return sp

    def __del__(self):
        if getattr(self, "_own", False) and N._host is not None:
            N._host.tnth_space_destroy(self._h)
            self._own = False

    @property
    def rank(self):
        return self.nout + self.nin

    def key(self):
        # spaces are immutable: the key (and the adjoint below) are built once
        if self._key is None:
            self._key = (self.rule.name, tuple(l.key() for l in self.codomain), tuple(l.key() for l in self.domain))
        return self._key

    def __eq__(self, o):
        return isinstance(o, HomSpace) and o.key() == self.key()

    def __hash__(self):
        return hash(self.key())

    @property
    def sectors(self):
        """[(coupled id, offset, rows, cols, dim)] in storage order."""
        if self._sectors is None:
            h = N.host()
            out = []
            for i in range(h.tnth_space_n_sectors(self._h)):
                c, o, r, k, d = C.c_uint64(), C.c_int64(), C.c_int64(), C.c_int64(), C.c_double()
                N.check_host(h.tnth_space_sector(self._h, i, C.byref(c), C.byref(o), C.byref(r),
                                                 C.byref(k), C.byref(d)))
                out.append((c.value, o.value, r.value, k.value, d.value))
            self._sectors = out
        return self._sectors

    def blocks(self):
        """[(offset, shape, strides, coupled, uncoupled ids)] in storage order."""
        h = N.host()
        out = []
        R = N.MAX_RANK
        for i in range(h.tnth_space_n_blocks(self._h)):
            off, rank = C.c_int64(), C.c_int32()
            shape, strides = (C.c_int64 * R)(), (C.c_int64 * R)()
            c = C.c_uint64()
            N.check_host(h.tnth_space_block(self._h, i, C.byref(off), C.byref(rank), shape, strides,
                                            C.byref(c)))
            unc = (C.c_uint64 * R)()
            N.check_host(h.tnth_space_block_trees(self._h, i, unc, None, None))
            r = rank.value
            out.append((off.value, tuple(shape[:r]), tuple(strides[:r]), c.value, tuple(unc[:r])))
        return out

    def permuted(self, p1, p2):
        out = C.c_void_p()
        N.check_host(N.host().tnth_space_permuted(self._h, N.sz_array(p1), len(p1), N.sz_array(p2),
                                                  len(p2), C.byref(out)))
        sp = HomSpace(self.rule, None, None, _handle=out.value)
        sp._own = True
        return sp

    def adjoint(self):
        if self._adj is None:
            self._adj = HomSpace(self.rule, self.domain, self.codomain)
        return self._adj

    def __repr__(self):
        return f"HomSpace({self.rule.name}, {self.nout}<-{self.nin}, len={self.len})"


# ---------------------------------------------------------------------------
# Runtime (backend context + structure caches)
# ---------------------------------------------------------------------------
class _Transform:
    def __init__(self, rt, space, p1, p2):
        h = N.host()
        out = C.c_void_p()
        N.check_host(h.tnth_transform_compile(space._h, N.sz_array(p1), len(p1), N.sz_array(p2),
                                              len(p2), C.byref(out)))
        self._h = out
        self.dst_space = HomSpace(space.rule, None, None, _handle=h.tnth_transform_dst_space(out),
                                  _owner=self)
        ns, nm = C.c_size_t(), C.c_size_t()
        h.tnth_transform_counts(out, C.byref(ns), C.byref(nm))
        self.n_single, self.n_multi = ns.value, nm.value
        self.desc = h.tnth_transform_desc(out)
        plan = C.c_void_p()
        rt._check(N.hip().tnt_plan_upload(rt._ctx, self.desc, C.byref(plan)))
        self.plan = plan
        self.rt = rt

    def __del__(self):
        try:
            N.hip().tnt_plan_destroy(self.rt._ctx, self.plan)
            N.host().tnth_transform_destroy(self._h)
        except Exception:
            pass


class _HostContraction:
    """Host-compiled contraction structure; shared by every Runtime (device-independent)."""

    def __init__(self, a, b, axes_a, axes_b, out_perm, out_nout, adj_a, adj_b):
        h = N.host()
        out = C.c_void_p()
        N.check_host(h.tnth_contract_compile(
            a._h, b._h, N.sz_array(axes_a), N.sz_array(axes_b), len(axes_a), N.sz_array(out_perm),
            out_nout, int(adj_a), int(adj_b), C.byref(out)))
        self._h = out
        self.dst_space = HomSpace(a.rule, None, None, _handle=h.tnth_contract_dst_space(out), _owner=self)
        self.flops = float(h.tnth_contract_flops(out))
        self.desc = h.tnth_contract_desc(out)

    def __del__(self):
        try:
            N.host().tnth_contract_destroy(self._h)
        except Exception:
            pass


class _HostBatch:
    """`copies` stacked instances of a host contraction (tnth_contract_batch);
    shared: 0 = the lhs is shared by every copy, 1 = the rhs, 2 = neither."""

    def __init__(self, host, copies, shared):
        out = C.c_void_p()
        N.check_host(N.host().tnth_contract_batch(host.desc, copies, shared, C.byref(out)))
        self._h = out
        self.host = host
        self.copies, self.shared = copies, shared
        self.desc = N.host().tnth_batch_desc(out)
        self.flops = host.flops * copies

    def __del__(self):
        try:
            N.host().tnth_batch_destroy(self._h)
        except Exception:
            pass


_HOST_CONTRACTIONS = {}


def _host_contraction(key, *args):
    hc = _HOST_CONTRACTIONS.get(key)
    if hc is None:
        hc = _HostContraction(*args)
        _HOST_CONTRACTIONS[key] = hc
    return hc


class _Contraction:
    def __init__(self, rt, host):
        self.host = host
        self.dst_space = getattr(host, "dst_space", None)
        self.flops = host.flops
        self.desc = host.desc
        art = C.c_void_p()
        rt._check(N.hip().tnt_contract_upload(rt._ctx, self.desc, C.byref(art)))
        self.art = art
        self.rt = rt

    @property
    def route(self):
        """"fused_narrow" (one recoupling pass, coefficients gathered from the
        rhs), "tiny_gathered" (FP64 calls: one launch, transforms folded into
        the GEMM's loads and stores) or "transform_gemm" (operand transforms,
        grouped GEMM, output transform)."""
        r = C.c_int()
        self.rt._check(N.hip().tnt_contract_route(self.art, C.byref(r)))
        return {1: "fused_narrow", 2: "tiny_gathered"}.get(r.value, "transform_gemm")

    def __del__(self):
        try:
            N.hip().tnt_contract_destroy(self.rt._ctx, self.art)
        except Exception:
            pass


class Runtime:
    """One gfx950 backend context (HIP stream + pools) plus structure caches.

    Mirrors Runtime::builder().cuda(device).build() (tenet/src/runtime.rs:1037,
    :1175-1262) with the new dense backend selected.
    """

    def __init__(self, device=0, precision="f64"):
        self.device = device
        ctx = C.c_void_p()
        rc = N.hip().tnt_ctx_create(device, C.byref(ctx))
        if rc != N.OK:
            raise TenetError(rc, f"cannot create a backend context on device {device}")
        self._ctx = ctx
        self.precision = precision
        self._transforms = {}
        self._contractions = {}

    def _check(self, rc):
        if rc != N.OK:
            raise TenetError(rc, N.hip().tnt_last_error(self._ctx).decode())

    def synchronize(self):
        self._check(N.hip().tnt_ctx_synchronize(self._ctx))

    def stream(self):
        s = C.c_void_p()
        self._check(N.hip().tnt_ctx_stream(self._ctx, C.byref(s)))
        return s.value

    def transform(self, space, p1, p2):
        key = (space.key(), tuple(p1), tuple(p2))
        t = self._transforms.get(key)
        if t is None:
            t = _Transform(self, space, list(p1), list(p2))
            self._transforms[key] = t
        return t

    def contraction(self, a, b, axes_a, axes_b, out_perm, out_nout, adj_a=False, adj_b=False):
        key = (a.key(), b.key(), tuple(axes_a), tuple(axes_b), tuple(out_perm), out_nout, adj_a, adj_b)
        c = self._contractions.get(key)
        if c is None:
            host = _host_contraction(key, a, b, list(axes_a), list(axes_b), list(out_perm), out_nout, adj_a, adj_b)
            c = _Contraction(self, host)
            self._contractions[key] = c
        return c

    def contraction_batch(self, a, b, axes_a, axes_b, out_perm, out_nout, copies, shared, adj_a=False, adj_b=False):
        """Device artifact of `copies` stacked instances of a contraction (one
        tnt_contract call runs them all); see _HostBatch for `shared`."""
        key = (a.key(), b.key(), tuple(axes_a), tuple(axes_b), tuple(out_perm), out_nout, adj_a, adj_b,
               "batch", copies, shared)
        c = self._contractions.get(key)
        if c is None:
            one = self.contraction(a, b, axes_a, axes_b, out_perm, out_nout, adj_a, adj_b)
            c = _Contraction(self, _HostBatch(one.host, copies, shared))
            c.dst_space = one.dst_space  # the space of one copy
            self._contractions[key] = c
        return c

    def clear_caches(self):
        self.synchronize()
        self._transforms.clear()
        self._contractions.clear()

    def close(self):
        if getattr(self, "_ctx", None):
            self.clear_caches()
            N.hip().tnt_ctx_destroy(self._ctx)
            self._ctx = None


PRECISIONS = {"f64": N.PREC_F64, "f32": N.PREC_F32, "split_bf16": N.PREC_SPLIT_BF16, "split_i8": N.PREC_SPLIT_I8}


# ---------------------------------------------------------------------------
# TensorMap
# ---------------------------------------------------------------------------
class TensorMap:
    """Block-sparse FP64 tensor map codomain <- domain, payload resident in HBM."""

    def __init__(self, rt, space, _buf=None, _adjoint=False, _base=None):
        self.rt = rt
        self.space = space
        self._adjoint = _adjoint
        self._base = _base  # lazy adjoint keeps its parent's storage alive
        if _buf is None:
            b = C.c_void_p()
            rt._check(N.hip().tnt_buf_alloc(rt._ctx, max(space.len, 1), 0, C.byref(b)))
            _buf = b
            self._own = True
        else:
            self._own = False
        self._buf = _buf

    def __del__(self):
        if getattr(self, "_own", False) and getattr(self.rt, "_ctx", None):
            try:
                N.hip().tnt_buf_free(self.rt._ctx, self._buf)
            except Exception:
                pass

    # -- construction ------------------------------------------------------
    @staticmethod
    def _space(codomain, domain):
        legs = list(codomain) + list(domain)
        rule = legs[0].rule
        return HomSpace(rule, codomain, domain)

    @classmethod
    def zeros(cls, rt, codomain, domain=()):
        t = cls(rt, codomain if isinstance(codomain, HomSpace) else cls._space(codomain, domain))
        rt._check(N.hip().tnt_buf_zero(rt._ctx, t._buf))
        return t

    @classmethod
    def rand(cls, rt, codomain, domain=(), seed=0, normal=False):
        """Seeded fill over the flat payload (Fill::Rand, typed.rs:2086-2092)."""
        t = cls(rt, codomain if isinstance(codomain, HomSpace) else cls._space(codomain, domain))
        rt._check(N.hip().tnt_buf_fill_random(rt._ctx, t._buf, seed & ((1 << 64) - 1), int(normal)))
        return t

    @classmethod
    def from_numpy(cls, rt, space, data):
        data = np.ascontiguousarray(data, dtype=np.float64)
        if data.shape != (space.len,):
            raise ValueError(f"payload length {data.shape} != {space.len}")
        t = cls(rt, space)
        rt._check(N.hip().tnt_buf_upload(rt._ctx, t._buf, data.ctypes.data, space.len))
        return t

    @classmethod
    def from_block_fn(cls, rt, codomain, domain, fn):
        """fn(uncoupled ids, coupled id, index tuple) per stored element (host fill)."""
        space = cls._space(codomain, domain)
        data = np.zeros(space.len)
        for off, shape, strides, coupled, unc in space.blocks():
            view = np.lib.stride_tricks.as_strided(data[off:], shape=shape,
                                                   strides=tuple(8 * s for s in strides))
            for idx in np.ndindex(*shape):
                view[idx] = fn(unc, coupled, idx)
        return cls.from_numpy(rt, space, data)

    def to_numpy(self):
        if self._adjoint:
            raise TenetError(-2, "materialise a lazy adjoint with contract/compose first")
        out = np.empty(self.space.len)
        if self.space.len:
            self.rt._check(N.hip().tnt_buf_download(self.rt._ctx, self._buf, out.ctypes.data, self.space.len))
        return out


    def to_bytes(self):
        """Tensor snapshot in TeNeT's v1 wire format (TensorMap::to_bytes_with,
        tenet/src/typed/serialization.rs:912-957): the payload is downloaded and
        encoded on the host; a lazy adjoint is stored as the adjoint of its parent."""
        base = self._base if self._adjoint else self
        data = base.to_numpy()
        b = C.c_void_p()
        N.check_host(N.host().tnth_snapshot_tensor(base.space._h, data.ctypes.data_as(C.POINTER(C.c_double)),
                                                   base.space.len, int(self._adjoint), C.byref(b)))
        return _take_bytes(b)

    @classmethod
    def from_bytes(cls, rt, data):
        """Inverse of to_bytes (TensorMap::from_bytes_with, serialization.rs:1569-1640):
        validates the whole snapshot on the host, then uploads the payload."""
        h = N.host()
        data = bytes(data)
        sp, adj = C.c_void_p(), C.c_int()
        N.check_host(h.tnth_snapshot_tensor_space(data, len(data), C.byref(sp), C.byref(adj)))
        rule_name = {v: k for k, v in N.RULES.items()}[h.tnth_space_rule(sp)]
        space = HomSpace._adopt(Rule(rule_name), sp.value)
        payload = np.empty(space.len)
        N.check_host(h.tnth_snapshot_tensor_payload(data, len(data), space._h,
                                                    payload.ctypes.data_as(C.POINTER(C.c_double)), space.len))
        t = cls.from_numpy(rt, space, payload)
        return t.adjoint() if adj.value else t
    @property
    def codomain_rank(self):
        return self.space.nout

    @property
    def domain_rank(self):
        return self.space.nin

    # -- structure ops -----------------------------------------------------
    def adjoint(self):
        """Lazy adjoint (MatrixOp::Adjoint); consumed by contract/compose."""
        if self._adjoint:
            return self._base
        return TensorMap(self.rt, self.space.adjoint(), _buf=self._buf, _adjoint=True, _base=self)

    def permute(self, p1, p2, out=None, alpha=1.0, beta=0.0):
        """TensorKit permute: codomain <- p1 axes, domain <- p2 axes (typed.rs:13163)."""
        if self._adjoint:
            raise TenetError(-2, "permute of a lazy adjoint: compose it first")
        t = self.rt.transform(self.space, p1, p2)
        if out is None:
            out = TensorMap(self.rt, t.dst_space)
            beta = 0.0
        elif out.space != t.dst_space:
            raise TenetError(-3, "destination space does not match permute(p1, p2)")
        self.rt._check(N.hip().tnt_tree_transform(self.rt._ctx, t.plan, out._buf, self._buf, alpha, beta,
                                                 1 if beta == 0.0 else 0))
        return out

    def contract(self, axes, other, other_axes, out_perm=None, out_nout=None, alpha=1.0, beta=0.0,
                 out=None, precision=None):
        """Contract self[axes] with other[other_axes] (typed.rs:13653, TensorKit @tensor order)."""
        a = self._base if self._adjoint else self
        b = other._base if other._adjoint else other
        n_open = self.space.rank + other.space.rank - 2 * len(axes)
        if out_perm is None:
            out_perm = list(range(n_open))
        if out_nout is None:
            out_nout = self.space.rank - len(axes)
        c = self.rt.contraction(a.space, b.space, list(axes), list(other_axes), list(out_perm), out_nout,
                                self._adjoint, other._adjoint)
        if out is None:
            out = TensorMap(self.rt, c.dst_space)
            beta = 0.0
        elif out.space is not c.dst_space and out.space != c.dst_space:
            raise TenetError(-3, "destination space does not match the contraction")
        prec = PRECISIONS[precision or self.rt.precision]
        self.rt._check(N.hip().tnt_contract(self.rt._ctx, c.art, out._buf, a._buf, b._buf, alpha, beta, prec))
        return out

    def compose(self, other, precision=None):
        """self * other over the shared legs (typed.rs:13842)."""
        n1 = self.space.nout
        nin = self.space.nin
        axes = list(range(n1, n1 + nin))
        oaxes = list(range(nin))
        return self.contract(axes, other, oaxes, precision=precision)

    def add(self, other, alpha=1.0, beta=1.0):
        """alpha * self + beta * other on identical spaces."""
        if self.space != other.space:
            raise TenetError(-3, "add: spaces differ")
        out = TensorMap(self.rt, self.space)
        h = N.hip()
        self.rt._check(h.tnt_axpby(self.rt._ctx, out._buf, self._buf, alpha, 0.0))
        self.rt._check(h.tnt_axpby(self.rt._ctx, out._buf, other._buf, beta, 1.0))
        return out

    def axpy_(self, other, alpha):
        """self += alpha * other (in place)."""
        self.rt._check(N.hip().tnt_axpby(self.rt._ctx, self._buf, other._buf, alpha, 1.0))
        return self

    def scale_(self, alpha):
        self.rt._check(N.hip().tnt_axpby(self.rt._ctx, self._buf, self._buf, alpha, 0.0))
        return self

    def copy(self):
        out = TensorMap(self.rt, self.space)
        self.rt._check(N.hip().tnt_axpby(self.rt._ctx, out._buf, self._buf, 1.0, 0.0))
        return out

    # -- reductions --------------------------------------------------------
    def inner(self, other):
        """sum_c dim(c) <A_c, B_c> in FP64 on the device (typed.rs:699-716, 11210-11262)."""
        if self.space != other.space:
            raise TenetError(-3, "inner: spaces differ")
        secs = self.space.sectors
        regs = (N.Region * max(1, len(secs)))(*[N.Region(o, r * k) for _, o, r, k, _ in secs])
        ws = (C.c_double * max(1, len(secs)))(*[d for *_, d in secs])
        out = C.c_double()
        self.rt._check(N.hip().tnt_weighted_dot(self.rt._ctx, self._buf, other._buf, regs, ws, len(secs),
                                               C.byref(out)))
        return out.value

    def norm(self):
        return math.sqrt(max(self.inner(self), 0.0))

    def _dot_tables(self):
        """Region and weight arrays of the weighted inner product, per space."""
        t = getattr(self.space, "_dot_tables", None)
        if t is None:
            secs = self.space.sectors
            regs = (N.Region * max(1, len(secs)))(*[N.Region(o, r * k) for _, o, r, k, _ in secs])
            ws = (C.c_double * max(1, len(secs)))(*[d for *_, d in secs])
            t = self.space._dot_tables = (regs, ws, len(secs))
        return t

    def inner_async(self, other, scalars, index):
        """Queue sum_c dim(c) <A_c, B_c> into scalars[index] on the device (no
        host synchronisation; KrylovVector::dot_real, tenet-krylov/src/vector.rs:13)."""
        if self.space != other.space:
            raise TenetError(-3, "inner: spaces differ")
        regs, ws, n = self._dot_tables()
        self.rt._check(N.hip().tnt_weighted_dot_async(self.rt._ctx, self._buf, other._buf, regs, ws, n,
                                                      scalars, index))

    def tr(self):
        """sum_c dim(c) tr(block_c) for codomain == domain (host reduction of a download)."""
        data = self.to_numpy()
        total = 0.0
        for _, o, r, k, d in self.space.sectors:
            m = data[o:o + r * k].reshape(k, r).T
            total += d * float(np.trace(m))
        return total


__all__ = ["Rule", "GradedSpace", "HomSpace", "Runtime", "TensorMap", "TenetError"]
