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
  * admitted requests are issued round-robin over `n_streams` backend
    contexts (one HIP stream each, default priority), so requests that arrive
    while others run overlap on the device;
  * completion is polled with stream events, freeing slot and footprint,
    after which waiting arrivals are admitted.
Device buffers: a request holds one buffer set (theta, the three H_eff
intermediates, the result) of its (stream, bucket) pair from the moment it is
admitted until its completion event fires; sets are reused by later requests
on that pair, and the footprint admission counts the bytes of the sets in
flight.  The four tnt_contract calls of a buffer set are captured once, when
the server warms up, as a hipGraph on its stream, so a request issues one fill
(its seed), one graph launch, one stream-ordered dot (its result slot) and
one event record (`graphs=False`, and sets built later while serving, replay
the prebuilt argument lists of the four calls instead).  Every request
computes on its own buffers with deterministic kernels, so its result is
bit-for-bit the same alone or in any mix, and with or without graphs (tests
in tests/test_gpu_heff.py).
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
                 "footprint", "t_issue", "t_done", "slot")


class _BufSet:
    """One request's device buffers on one (stream, bucket) pair and the
    prebuilt tnt_contract argument tuples of the H_eff chain over them."""
    __slots__ = ("theta", "scratch", "y", "calls", "bytes", "graph")


class HeffServer:
    def __init__(self, rule="U1", buckets=None, n_streams=4, max_inflight=16,
                 hbm_budget_bytes=48 << 30, precision="split_i8", env_seed=7, mpo="random",
                 graphs=True):
        self.rule = rule
        self.graphs = graphs
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
        self._free = {}                       # (stream, D) -> idle buffer sets
        self._events = [[] for _ in self.rts]  # idle events per stream
        self._dot = {}                        # D -> (regions, weights, count) of theta's sectors

    # -- per-bucket structures (cold: compiled per context once) -------------
    def warm(self):
        """Compile every (bucket, stream) artifact, build one buffer set per pair
        and capture its graph.  The graphs are captured after every bucket has
        run once on every stream: a stream's split GEMM workspace grows (and is
        reallocated) while larger buckets warm up, and a graph keeps the
        addresses it was captured with."""
        for D in self.buckets:
            for k, rt in enumerate(self.rts):
                bs = self._new_set(D, k)
                self.flops[D] = self._view(D, k)._flops
                self.fp[D] = bs.bytes
                self._free.setdefault((k, D), []).append(bs)
                rt.synchronize()
        if self.graphs:
            for (k, _), sets in self._free.items():
                for bs in sets:
                    if bs.graph is None:
                        bs.graph = self._capture(k, bs.calls)
                self.rts[k].synchronize()

    def _new_set(self, D, k):
        v = self._view(D, k)
        bs = _BufSet()
        bs.theta = TensorMap.zeros(self.rts[k], v.theta_space)
        bs.scratch = {}
        bs.y = v.apply(bs.theta, bs.scratch)
        bs.calls = v.launch_list(bs.theta, bs.scratch)
        bs.bytes = 8 * (bs.theta.space.len + sum(t.space.len for t in bs.scratch.values()))
        bs.graph = None  # captured by warm(); sets built later, while serving, issue eagerly
        if D not in self._dot:
            secs = v.theta_space.sectors
            regs = (N.Region * max(1, len(secs)))(*[N.Region(o, r * c) for _, o, r, c, _ in secs])
            ws = (C.c_double * max(1, len(secs)))(*[d for *_, d in secs])
            self._dot[D] = (regs, ws, len(secs))
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

    def _issue(self, r, out_buf, keep):
        """Queue one request on its stream: fill theta, the four contractions,
        the dot into the request's result slot, then a completion event."""
        h = N.hip()
        k, D = r.ctx_index, r.D
        rt = self.rts[k]
        free = self._free.setdefault((k, D), [])
        bs = free.pop() if free else self._new_set(D, k)
        ctx = rt._ctx
        rt._check(h.tnt_buf_fill_random(ctx, bs.theta._buf, r.seed & ((1 << 64) - 1), 1))
        if bs.graph is not None:
            rt._check(h.tnt_graph_launch(ctx, bs.graph))
        else:
            for art, d, a, b in bs.calls:
                rt._check(h.tnt_contract(ctx, art, d, a, b, 1.0, 0.0, self._prec))
        regs, ws, ns = self._dot[D]
        rt._check(h.tnt_weighted_dot_async(ctx, bs.theta._buf, bs.y._buf, regs, ws, ns, out_buf._buf, r.rid))
        if keep:  # a copy of the result vector, read at completion (tests)
            r.y_copy = TensorMap.zeros(rt, bs.y.space)
            rt._check(h.tnt_axpby(ctx, r.y_copy._buf, bs.y._buf, 1.0, 0.0))
        ev = self._events[k].pop() if self._events[k] else None
        if ev is None:
            ev = C.c_void_p()
            rt._check(h.tnt_event_create(ctx, C.byref(ev)))
        rt._check(h.tnt_event_record(ctx, ev))
        r.bufs, r.event = bs, ev

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
        inflight = []
        used = 0
        done = []
        next_ctx = 0
        results = {}
        issue_s = 0.0
        max_inflight_seen = 0
        for rt in self.rts:
            rt.synchronize()
        t0 = time.perf_counter()
        while pending or inflight:
            now = time.perf_counter() - t0
            # admit arrivals (slot + HBM footprint)
            while pending and pending[0][1][0] <= now and slots:
                i, (arr, D, seed) = pending[0]
                fp = self.fp[D]
                if used + fp > self.budget and inflight:
                    break
                pending.popleft()
                r = _Req()
                r.rid, r.arrival, r.D, r.seed = i, arr, D, seed
                r.footprint = fp
                r.slot = slots.popleft()
                r.ctx_index = next_ctx
                next_ctx = (next_ctx + 1) % len(self.rts)
                r.t_issue = time.perf_counter()
                self._issue(r, out_buf[r.ctx_index], keep_results)
                issue_s += time.perf_counter() - r.t_issue
                used += fp
                inflight.append(r)
            max_inflight_seen = max(max_inflight_seen, len(inflight))
            # retire finished requests
            still = []
            for r in inflight:
                rt = self.rts[r.ctx_index]
                q = h.tnt_event_query(rt._ctx, r.event)
                if q < 0:
                    rt._check(q)
                if q == 1:
                    r.t_done = time.perf_counter() - t0
                    if keep_results:
                        results[(r.rid, "y")] = r.y_copy.to_numpy()
                        r.y_copy = None
                    self._events[r.ctx_index].append(r.event)
                    self._free[(r.ctx_index, r.D)].append(r.bufs)
                    used -= r.footprint
                    slots.append(r.slot)
                    done.append(r)
                    r.bufs = r.event = None
                else:
                    still.append(r)
            inflight = still
            # idle until the next arrival: sleep through long gaps, spin the last
            # 0.2 ms (a sleep can overshoot by tens of microseconds)
            if not inflight and pending and pending[0][1][0] - now > 4e-4:
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
        for (k, _), sets in self._free.items():
            for bs in sets:
                if bs.graph is not None:
                    h.tnt_graph_destroy(self.rts[k]._ctx, bs.graph)
                    bs.graph = None
        self._free.clear()
        self.env.clear()
        self.views.clear()
        for rt in self.rts:
            rt.close()
