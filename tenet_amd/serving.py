"""Continuous-batching server for two-site H_eff matvecs (one MI355X).

A request is (arrival time, bond dimension D, seed): build theta (seeded
normal fill on the device), apply H_eff of its D bucket (four block-sparse
contractions, split-INT8 GEMMs by default) and take the FP64 Rayleigh
numerator <theta, H theta> on the device.  Environments (L, R, W) of every D
bucket stay resident in HBM for the life of the server.

Scheduling (host thread):
  * arrivals are admitted when a slot is free (at most `max_inflight` in
    flight) and the HBM footprint of the in-flight requests plus the new
    one fits the budget -- admission by HBM footprint;
  * admitted requests of one bond-dimension bucket that are waiting together
    are batched: up to `max_batch` of them run as one stacked contraction
    chain (tnth_contract_batch: every copy's sector GEMMs in the same grouped
    launches, every copy's recoupling in the same passes, the environments
    read once), so a batch costs the launches of one request;
  * batches (and single requests) are issued round-robin over `n_streams`
    backend contexts (one HIP stream each, default priority), so work that
    arrives while other work runs overlaps on the device;
  * completion is polled with stream events, freeing slot and footprint,
    after which waiting arrivals are admitted.
Device buffers: a batch holds one buffer set (theta, the three H_eff
intermediates and the result, each stacked over the batch's copies) of its
(stream, bucket, batch size) from the moment it is admitted until its
completion event fires; sets are reused by later batches of that key, and
the footprint admission counts the bytes of the sets in flight.  The four tnt_contract calls of a buffer set are captured once, when
the server warms up, as a hipGraph on its stream, so a request issues one fill
(its seed), one graph launch, one stream-ordered dot (its result slot) and
one event record (`graphs=False`, and sets built later while serving, replay
the prebuilt argument lists of the four calls instead).  Every request
computes on its own buffers with deterministic kernels, so its result is
bit-for-bit the same alone or in any mix, batched or not, and with or without
graphs (the stacked chain runs each copy's elements through the same kernels
in the same order; tests in tests/test_gpu_heff.py).
Latency = completion seen by the scheduler - arrival.
"""
import ctypes as C
import math
import time
from collections import deque

import numpy as np

from . import _native as N
from .heff import TwoSiteHeff
from .tensor import Runtime, TensorMap


def poisson_trace(n, rate=2000.0, median_d=512, sigma=0.7, dmin=32, dmax=4096, seed=5, buckets=None):
    """n requests: exponential inter-arrivals at `rate`/s, D ~ lognormal clipped, snapped to buckets."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n))
    d = np.clip(np.exp(rng.normal(math.log(median_d), sigma, size=n)), dmin, dmax)
    if buckets is None:
        buckets = default_buckets(dmin, dmax)
    b = np.array(buckets)
    snapped = [int(b[np.argmin(np.abs(np.log(b) - math.log(x)))]) for x in d]
    return [(float(ti), di, 10_000 + i) for i, (ti, di) in enumerate(zip(t, snapped))]


def default_buckets(dmin=32, dmax=4096, n=15):
    """Log-spaced bond dimensions (multiples of 32) the environments are built for."""
    g = np.exp(np.linspace(math.log(dmin), math.log(dmax), n))
    return sorted(set(int(max(dmin, 32 * round(x / 32))) for x in g))


class _Req:
    __slots__ = ("rid", "arrival", "D", "seed", "ctx_index", "event", "bufs", "y_copy",
                 "footprint", "t_issue", "t_done", "slot", "batch")


class _Batch:
    __slots__ = ("reqs", "ctx_index", "D", "event", "bufs", "footprint")


class _Raw:
    """A device payload buffer without a tensor space (stacked copies)."""

    def __init__(self, rt, n, base=None, offset=0):
        self.rt, self.n, self._base = rt, n, base
        b = C.c_void_p()
        if base is None:
            rt._check(N.hip().tnt_buf_alloc(rt._ctx, max(1, n), N.F64, C.byref(b)))
        else:
            p = C.c_void_p()
            rt._check(N.hip().tnt_buf_info(base._buf, C.byref(p), None, None, None))
            rt._check(N.hip().tnt_buf_wrap(rt._ctx, C.c_void_p(p.value + 8 * offset), max(1, n), N.F64, C.byref(b)))
        self._buf = b

    def __del__(self):
        try:
            N.hip().tnt_buf_free(self.rt._ctx, self._buf)
        except Exception:
            pass


class _BufSet:
    """One batch's device buffers on one (stream, bucket, batch size) and the
    prebuilt tnt_contract argument tuples of the H_eff chain over them;
    thetas / ys hold copy c's views."""
    __slots__ = ("B", "theta", "scratch", "y", "thetas", "ys", "regions", "calls", "bytes", "graph", "keep")


class HeffServer:
    def __init__(self, rule="U1", buckets=None, n_streams=4, max_inflight=16,
                 hbm_budget_bytes=48 << 30, precision="split_i8", env_seed=7, mpo="random",
                 graphs=True, max_batch=8, batch_max_d=1024):
        """max_batch: largest batch (a power of two; 1 disables batching);
        batch_max_d: largest bucket that batches (larger requests fill the
        device alone)."""
        self.rule = rule
        self.graphs = graphs
        self.batch_sizes = [b for b in (1, 2, 4, 8, 16) if b <= max(1, max_batch)]
        self.batch_max_d = batch_max_d
        self.precision = precision
        self.max_inflight = max_inflight
        self.budget = hbm_budget_bytes
        self.rts = [Runtime(0, precision=precision) for _ in range(n_streams)]
        self.buckets = list(buckets or default_buckets())
        self.env = {}       # D -> TwoSiteHeff on context 0
        self.views = {}     # (D, ctx index) -> environment views on that context
        for D in self.buckets:
            self.env[D] = TwoSiteHeff(self.rts[0], rule, D, seed=env_seed + D, precision=precision, mpo=mpo)
        self.rts[0].synchronize()  # the other streams read the environments
        self.flops = {}
        self.fp = {}
        self.results = None
        self._prec = {"f64": N.PREC_F64, "f32": N.PREC_F32, "split_bf16": N.PREC_SPLIT_BF16,
                      "split_i8": N.PREC_SPLIT_I8}[precision]
        self._free = {}                       # (stream, D, batch size) -> idle buffer sets
        self._events = [[] for _ in self.rts]  # idle events per stream
        self._dot = {}                        # D -> ((offset, len) per sector of theta, weights, count)

    # -- per-bucket structures (cold: compiled per context once) -------------
    def warm(self):
        """Compile every (bucket, stream) artifact, build one buffer set per pair
        and capture its graph.  The graphs are captured after every bucket has
        run once on every stream: a stream's split GEMM workspace grows (and is
        reallocated) while larger buckets warm up, and a graph keeps the
        addresses it was captured with."""
        for D in self.buckets:
            for k, rt in enumerate(self.rts):
                for B in self._sizes(D):
                    bs = self._new_set(D, k, B)
                    if B == 1:
                        self.flops[D] = self._view(D, k)._flops
                        self.fp[D] = bs.bytes
                    self._free.setdefault((k, D, B), []).append(bs)
                rt.synchronize()
        if self.graphs:
            for (k, _, _), sets in self._free.items():
                for bs in sets:
                    if bs.graph is None:
                        bs.graph = self._capture(k, bs.calls)
                self.rts[k].synchronize()

    def _sizes(self, D):
        return self.batch_sizes if D <= self.batch_max_d else [1]

    def _new_set(self, D, k, B=1):
        v = self._view(D, k)
        rt = self.rts[k]
        bs = _BufSet()
        bs.B = B
        bs.graph = None  # captured by warm(); sets built later, while serving, issue eagerly
        if D not in self._dot:
            secs = v.theta_space.sectors
            ws = (C.c_double * max(1, len(secs)))(*[d for *_, d in secs])
            self._dot[D] = ([(o, r * c) for _, o, r, c, _ in secs], ws, len(secs))
        offs, ws, ns = self._dot[D]
        n_theta = v.theta_space.len
        # the dots run on copy c's views (theta and H theta of one request)
        regs = (N.Region * max(1, ns))(*[N.Region(o, ln) for o, ln in offs])
        bs.regions = [(regs, ws, ns)] * B
        if B == 1:
            bs.theta = TensorMap.zeros(rt, v.theta_space)
            bs.scratch = {}
            bs.y = v.apply(bs.theta, bs.scratch)
            bs.calls = v.launch_list(bs.theta, bs.scratch)
            bs.thetas, bs.ys = [bs.theta._buf], [bs.y._buf]
            bs.bytes = 8 * (bs.theta.space.len + sum(t.space.len for t in bs.scratch.values()))
            return bs
        # stacked chain: environments shared, theta and every intermediate stacked
        env = {"L": v.L, "W1": v.W1, "W2": v.W2, "R": v.R}
        spaces = {"theta": v.theta_space}
        bufs = {"theta": _Raw(rt, B * n_theta)}
        bs.calls = []
        for lhs, rhs, ax, bx, perm, nout, res in TwoSiteHeff.CHAIN:
            ls = env[lhs].space if lhs in env else spaces[lhs]
            rsp = env[rhs].space if rhs in env else spaces[rhs]
            shared = 0 if lhs in env else 1
            c = rt.contraction_batch(ls, rsp, ax, bx, perm, nout, B, shared)
            spaces[res] = c.dst_space
            bufs[res] = _Raw(rt, B * c.dst_space.len)
            lb = env[lhs]._buf if lhs in env else bufs[lhs]._buf
            rb = env[rhs]._buf if rhs in env else bufs[rhs]._buf
            bs.calls.append((c.art, bufs[res]._buf, lb, rb))
        bs.theta, bs.y = bufs["theta"], bufs["y"]
        bs.scratch = {k_: b_ for k_, b_ in bufs.items() if k_ not in ("theta", "y")}
        n_y = spaces["y"].len
        bs.keep = [_Raw(rt, n_theta, bs.theta, c * n_theta) for c in range(B)] + \
                  [_Raw(rt, n_y, bs.y, c * n_y) for c in range(B)]
        bs.thetas = [x._buf for x in bs.keep[:B]]
        bs.ys = [x._buf for x in bs.keep[B:]]
        bs.bytes = 8 * sum(b_.n for b_ in bufs.values())
        h = N.hip()
        rt._check(h.tnt_buf_zero(rt._ctx, bs.theta._buf))
        for art, d, a, b in bs.calls:  # run once: every table and workspace exists before capture
            rt._check(h.tnt_contract(rt._ctx, art, d, a, b, 1.0, 0.0, self._prec))
        return bs

    def _capture(self, k, calls):
        """hipGraph of the four contractions of one buffer set on stream k (the
        chain ran once eagerly on these buffers, so every table and workspace
        it uses exists; the capture allocates and uploads nothing)."""
        h = N.hip()
        rt = self.rts[k]
        ctx = rt._ctx
        rt._check(h.tnt_graph_begin(ctx))
        rc = 0
        for art, d, a, b in calls:
            rc = h.tnt_contract(ctx, art, d, a, b, 1.0, 0.0, self._prec)
            if rc:
                break
        g = C.c_void_p()
        rc_end = h.tnt_graph_end(ctx, C.byref(g))
        rt._check(rc)
        rt._check(rc_end)
        return g

    def _issue(self, bt, out_buf, keep):
        """Queue one batch on its stream: fill each copy's theta, the four
        (stacked) contractions, each request's dot into its result slot, then
        one completion event."""
        h = N.hip()
        k, D, B = bt.ctx_index, bt.D, len(bt.reqs)
        rt = self.rts[k]
        free = self._free.setdefault((k, D, B), [])
        bs = free.pop() if free else self._new_set(D, k, B)
        ctx = rt._ctx
        for c, r in enumerate(bt.reqs):
            rt._check(h.tnt_buf_fill_random(ctx, bs.thetas[c], r.seed & ((1 << 64) - 1), 1))
        if bs.graph is not None:
            rt._check(h.tnt_graph_launch(ctx, bs.graph))
        else:
            for art, d, a, b in bs.calls:
                rt._check(h.tnt_contract(ctx, art, d, a, b, 1.0, 0.0, self._prec))
        for c, r in enumerate(bt.reqs):
            regs, ws, ns = bs.regions[c]
            rt._check(h.tnt_weighted_dot_async(ctx, bs.thetas[c], bs.ys[c], regs, ws, ns, out_buf._buf, r.rid))
            if keep:  # a copy of the result vector, read at completion (tests)
                r.y_copy = TensorMap.zeros(rt, self._view(D, k).theta_space)
                src = bs.y._buf if B == 1 else bs.ys[c]
                rt._check(h.tnt_axpby(ctx, r.y_copy._buf, src, 1.0, 0.0))
        ev = self._events[k].pop() if self._events[k] else None
        if ev is None:
            ev = C.c_void_p()
            rt._check(h.tnt_event_create(ctx, C.byref(ev)))
        rt._check(h.tnt_event_record(ctx, ev))
        bt.bufs, bt.event = bs, ev

    def _view(self, D, k):
        key = (D, k)
        if key not in self.views:
            self.views[key] = self.env[D] if k == 0 else self.env[D].on(self.rts[k])
        return self.views[key]

    # -- serving loop ----------------------------------------------------------
    def serve(self, trace, keep_results=False, log=None):
        """Run a trace [(arrival_s, D, seed)]; returns a stats dict."""
        h = N.hip()
        n = len(trace)
        # one FP64 result per request, written on the device by the request's own stream
        out_buf = [TensorMap.zeros(rt, self._vector_space(n)) for rt in self.rts]
        slots = deque(range(self.max_inflight))
        pending = deque(sorted(enumerate(trace), key=lambda x: x[1][0]))
        waiting = {}    # bucket -> arrived, not yet admitted requests (arrival order)
        inflight = []   # batches
        used = 0
        done = []
        next_ctx = 0
        results = {}
        issue_s = 0.0
        max_inflight_seen = 0
        batch_hist = {}
        max_used = 0
        for rt in self.rts:
            rt.synchronize()
        t0 = time.perf_counter()
        while pending or waiting or inflight:
            now = time.perf_counter() - t0
            # admit the arrivals waiting now (slot + HBM footprint), batching
            # requests of one bucket: a batch is the largest batch size not
            # above the bucket's waiting requests and the free slots
            while pending and pending[0][1][0] <= now:
                item = pending.popleft()
                waiting.setdefault(item[1][1], deque()).append(item)
            blocked = False
            for D in sorted(waiting, key=lambda d: waiting[d][0][1][0]):  # oldest bucket first
                items = waiting[D]
                while items and slots and not blocked:
                    limit = min(len(items), len(slots))
                    # the largest batch whose buffers fit the HBM budget; one
                    # request is admitted over budget only when nothing runs
                    fits = [b for b in self._sizes(D) if b <= limit and used + b * self.fp[D] <= self.budget]
                    if not fits and not inflight:
                        fits = [1]
                    if not fits:
                        blocked = True  # wait for memory to free up
                        break
                    B = max(fits)
                    fp = self.fp[D] * B
                    bt = _Batch()
                    bt.reqs, bt.D, bt.footprint = [], D, fp
                    for _ in range(B):
                        i, (arr, _, seed) = items.popleft()
                        r = _Req()
                        r.rid, r.arrival, r.D, r.seed = i, arr, D, seed
                        r.footprint = self.fp[D]
                        r.slot = slots.popleft()
                        r.batch = bt
                        bt.reqs.append(r)
                    bt.ctx_index = next_ctx
                    next_ctx = (next_ctx + 1) % len(self.rts)
                    t_i = time.perf_counter()
                    for r in bt.reqs:
                        r.ctx_index, r.t_issue = bt.ctx_index, t_i
                    self._issue(bt, out_buf[bt.ctx_index], keep_results)
                    issue_s += time.perf_counter() - t_i
                    used += fp
                    max_used = max(max_used, used)
                    inflight.append(bt)
                    batch_hist[B] = batch_hist.get(B, 0) + 1
                if not items:
                    del waiting[D]
            max_inflight_seen = max(max_inflight_seen, sum(len(b.reqs) for b in inflight))
            # retire finished batches
            still = []
            for bt in inflight:
                rt = self.rts[bt.ctx_index]
                q = h.tnt_event_query(rt._ctx, bt.event)
                if q < 0:
                    rt._check(q)
                if q == 1:
                    t_done = time.perf_counter() - t0
                    for r in bt.reqs:
                        r.t_done = t_done
                        if keep_results:
                            results[(r.rid, "y")] = r.y_copy.to_numpy()
                            r.y_copy = None
                        slots.append(r.slot)
                        done.append(r)
                    self._events[bt.ctx_index].append(bt.event)
                    self._free[(bt.ctx_index, bt.D, len(bt.reqs))].append(bt.bufs)
                    used -= bt.footprint
                    bt.bufs = bt.event = None
                else:
                    still.append(bt)
            inflight = still
            # idle until the next arrival: sleep through long gaps, spin the last
            # 0.2 ms (a sleep can overshoot by tens of microseconds)
            if not inflight and not waiting and pending and pending[0][1][0] - now > 4e-4:
                time.sleep(pending[0][1][0] - now - 2e-4)
        wall = time.perf_counter() - t0
        for k, rt in enumerate(self.rts):
            vals = out_buf[k].to_numpy()
            for r in done:
                if r.ctx_index == k:
                    results[r.rid] = float(vals[r.rid])
        lat = np.array([r.t_done - r.arrival for r in done])
        flops = sum(self.flops[r.D] for r in done)
        self.results = results
        return {
            "requests": n,
            "wall_s": wall,
            "matvecs_per_s": n / wall,
            "tflops": flops / wall / 1e12,
            "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_ms": float(lat.mean() * 1e3),
            "offered_rate": n / max(trace[-1][0], 1e-9),
            "host_issue_ms_per_request": issue_s / max(1, n) * 1e3,
            "max_inflight": max_inflight_seen,
            "batches": dict(sorted(batch_hist.items())),
            "max_footprint_bytes": max_used,
        }

    # -- helpers -----------------------------------------------------------------
    @staticmethod
    def _vector_space(n):
        from .tensor import GradedSpace, HomSpace, Rule
        rule = Rule("U1")
        return HomSpace(rule, [GradedSpace(rule, [(0, max(1, n))])], [GradedSpace(rule, [(0, 1)])])

    def close(self):
        h = N.hip()
        for k, evs in enumerate(self._events):
            for ev in evs:
                h.tnt_event_destroy(self.rts[k]._ctx, ev)
        self._events = [[] for _ in self.rts]
        for (k, _, _), sets in self._free.items():
            for bs in sets:
                if bs.graph is not None:
                    h.tnt_graph_destroy(self.rts[k]._ctx, bs.graph)
                    bs.graph = None
        self._free.clear()
        self.env.clear()
        self.views.clear()
        for rt in self.rts:
            rt.close()
