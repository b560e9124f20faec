#!/usr/bin/env python3
"""Benchmark of the block-sparse symmetric contraction path on one MI355X.

Default workload (the headline metric's configuration): one U(1) MPS
left-environment update at bond dimension D=4096 (charges -10..10, Gaussian
degeneracies sigma 4), physical d=2, MPO bond 5, FP64:
    T1 = L . A,   T2 = T1 . W,   L' = A^dagger . T2
-- three contractions, five recoupling replays, 15.96 GFLOP of block GEMM per
step.  Inputs are resident in HBM before timing (seeded normal fill on device).

Prints ONE JSON line.  `value` = whole-step contract TFLOP/s (algorithmic
FP64 GEMM flops / wall time); `roofline` = the dominant kernel (grouped block
GEMM) timed with HIP events on the backend stream inside the timed region;
`cpu_baseline` = the CPU oracle (numpy, FP64) on the same workload.

Other workloads: --workload su2_env (SU(2) bond of dimension 4096),
--workload su2_fmove (SU(2) 2j<=8 F-move + contraction, 4.27 TFLOP/step),
--workload batched (256 U(1)xZ2 merges in one grouped launch, hipGraph) and
--workload serve (Poisson stream of two-site H_eff matvecs, continuous
batching over HIP streams; reports matvecs/s and p50/p99 latency).
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAKS = {"f64": 78.6, "f32": 157.3, "split_bf16": 2500.0, "split_i8": 5000.0}  # dense MFMA TFLOP/s (TOP/s for int8)
HBM_PEAK_GBS = 8000.0


GEMM_KERNELS = ["gemm_kernel<{}, 128, 128>", "gemm_kernel<{}, 64, 64>", "gemm_kernel<{}, 128, 64>",
                "gemm_kernel<{}, 64, 128>", "gemm_skinny_kernel<{}>", "gemm_kernel<{}, 128, 128, 2, 4>",
                "gemm_small_kernel<{}>"]
SPLIT_PRODUCTS = 15  # bf16 MFMA products per FP64 product at five digit planes


TRAFFIC_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic.txt")


def measured_traffic(workload, precision, kernel):
    """HBM bytes per launch of a dominant kernel from the committed counter
    table (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of the same bench command,
    KiB per dispatch scaled by the copy calibration), or (None, None) when no
    row matches the workload, precision and kernel."""
    try:
        with open(TRAFFIC_TABLE) as fh:
            lines = fh.read().splitlines()
    except OSError:
        return None, None
    scale = {}
    for ln in lines:
        if ln.startswith("# scale"):
            _, _, name, val = ln.split()
            scale[name] = float(val)
        if not ln.strip() or ln.startswith("#"):
            continue
        f = [x.strip() for x in ln.split("|")]
        if len(f) == 6 and f[0] == workload and f[1] == precision and f[2] == kernel:
            fetch = float(f[3]) * 1024 * scale.get("FETCH_SIZE", 1.0)
            write = float(f[4]) * 1024 * scale.get("WRITE_SIZE", 1.0)
            return fetch + write, f[5]
    return None, None


def roofline(precision, dom, workload=None):
    """Roofline record of the dominant GEMM kernel: algorithmic flops per launch
    over its average launch time (HIP events on the backend stream)."""
    avg_s = dom["ms"] * 1e-3 / max(1, dom["launches"])
    per_launch = dom["flops"] / max(1, dom["launches"])
    f64_eq = per_launch / avg_s / 1e12 if avg_s else 0.0
    traffic, source = measured_traffic(workload, precision, dom["kernel"])
    rec = {"bound": "mfma", "unit": "TFLOP/s", "traffic": traffic, "kernel": dom["kernel"],
           "launches": dom["launches"], "avg_launch_ms": avg_s * 1e3,
           "flops_per_launch": per_launch}
    if traffic is not None:
        rec["traffic_source"] = source
    if dom.get("bytes"):
        # memory-bound batch of tiny GEMMs: operands read once, results written once
        gbs = dom["bytes"] / max(1, dom["launches"]) / avg_s / 1e9 if avg_s else 0.0
        rec.update(bound="hbm", unit="GB/s", achieved=gbs, peak=HBM_PEAK_GBS, frac=gbs / HBM_PEAK_GBS,
                   bytes_per_launch=dom["bytes"] / max(1, dom["launches"]), tflops=f64_eq)
        return rec
    if precision in ("split_bf16", "split_i8"):
        # bf16 / int8 MFMA work actually issued: SPLIT_PRODUCTS digit-plane products per FP64 product
        ach = f64_eq * SPLIT_PRODUCTS
        rec.update(achieved=ach, peak=PEAKS[precision], frac=ach / PEAKS[precision], f64_equivalent=f64_eq)
        if precision == "split_i8":
            rec["unit"] = "TOP/s"
    else:
        rec.update(achieved=f64_eq, peak=PEAKS[precision], frac=f64_eq / PEAKS[precision])
    return rec


def settle(step, sync, ms):
    """Run `step` back to back for `ms` milliseconds of wall time (the queue
    is drained every few steps so it never runs far ahead of the clock);
    returns the number of steps run."""
    n = 0
    if ms > 0:
        sync()
        t = time.perf_counter()
        while (time.perf_counter() - t) * 1e3 < ms:
            for _ in range(4):
                step()
            n += 4
            sync()
    sync()
    return n


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_port_baseline(workload, wl, budget_s=30.0):
    """The C++/OpenMP CPU port of the host route (oracle/cpu_port.cpp) on the
    device workload's own compiled descriptors and downloaded payloads, at the
    full workload size: warm (descriptors compiled and workspaces allocated
    before the clock starts), the median of the runs that fit in budget_s (at
    least one, at most five).  Returns (seconds, flops, threads, sample, runs,
    result payload(s))."""
    import numpy as np

    from oracle import cpu_port as P
    if workload in ("u1_env", "su2_env"):
        fs = [P.Contract(c.desc, c.dst_space.len) for c in wl.contractions]
        L, A, W = wl.L.to_numpy(), wl.A.to_numpy(), wl.W.to_numpy()

        def step():
            return fs[2](A, fs[1](fs[0](L, A), W))
        sample = "full environment update (L.A, .W, A^dagger.) at the bench size"
    elif workload == "batched":
        ops = [(P.Contract(c.desc, c.dst_space.len), a.to_numpy(), b.to_numpy()) for a, b, _, c in wl.pairs]

        def step():
            return [f(a, b) for f, a, b in ops]
        sample = f"all {len(ops)} merges in FP64"
    elif hasattr(wl, "transform"):
        xf = wl.transform
        A, B = wl.A.to_numpy(), wl.B.to_numpy()
        ap = np.empty(xf.dst_space.len)
        f = P.Contract(wl.contraction.desc, wl.contraction.dst_space.len)

        def step():
            P.transform(xf.desc, A, dst=ap)
            return f(ap, B)
        sample = "full F-move (3<-1 codomain swap) + contraction"
    else:
        A, B = wl.A.to_numpy(), wl.B.to_numpy()
        f = P.Contract(wl.contraction.desc, wl.contraction.dst_space.len)

        def step():
            return f(A, B)
        sample = "full contraction A[1,3].B[2,0]"
    times, out = [], None
    t0 = time.perf_counter()
    while len(times) < 5:
        t = time.perf_counter()
        out = step()
        times.append(time.perf_counter() - t)
        if time.perf_counter() - t0 + times[-1] > budget_s:
            break
    times.sort()
    return times[len(times) // 2], wl.flops, P.threads(), sample, len(times), out


def cpu_baseline(workload, D, budget_s=20.0, wl=None, fmove="swap", min_runs=3, su2_count="dimension"):
    """Time the oracle (numpy FP64, host cores) on the same workload.  With the
    device workload `wl` given, the oracle's operands are its downloaded
    payloads, so the parity check compares results of identical inputs."""
    import numpy as np

    from oracle.fusion import HomSpace as OH, Leg as OL
    from oracle.symmetry import make_rule
    from oracle.tensor import OTensor
    from tenet_amd.workloads import mps_env_spaces, su2_fmove_space
    try:
        from threadpoolctl import threadpool_info
        cores = max([i.get("num_threads", 1) for i in threadpool_info() if i.get("user_api") == "blas"] or [1])
    except Exception:  # pragma: no cover
        cores = os.cpu_count() or 1
    if workload in ("u1_env", "su2_env"):
        rname = "U1" if workload == "u1_env" else "SU2"
        orule = make_rule(rname)
        _, V, P, M = mps_env_spaces(rname, D, su2_count)

        def leg(g):
            return OL(list(zip(g.ids, g.degs)), g.is_dual)
        oV, oP, oM = leg(V), leg(P), leg(M)
        sL, sA, sW = OH(orule, [oV, oM], [oV]), OH(orule, [oV], [oP, oV]), OH(orule, [oM, oP], [oP, oM])
        if wl is not None:
            L, A, W = OTensor(sL, wl.L.to_numpy()), OTensor(sA, wl.A.to_numpy()), OTensor(sW, wl.W.to_numpy())
        else:
            L, A, W = OTensor.randn(sL, 3), OTensor.randn(sA, 4), OTensor.randn(sW, 5)
        Ad = A.adjoint()

        def step():
            t1 = L.contract([2], A, [0], [0, 1, 2, 3], 2)
            t2 = t1.contract([1, 2], W, [3, 1], [0, 2, 3, 1], 2)
            return Ad.contract([2, 0], t2, [0, 2], [0, 1, 2], 2)
        sample = f"full {rname} environment update at D={D}" + (" multiplets" if su2_count == "multiplets" else "")
    elif workload == "batched":
        from tenet_amd.workloads import batched_specs
        orule = make_rule("U1xZ2")
        oP = OL([(orule.encode((0, 0)), 1), (orule.encode((1, 1)), 1)])
        ops = []
        nsample = 16
        for i, (pl, pm, pr) in enumerate(batched_specs(256)[:nsample]):
            Vl, Vm, Vr = (OL([(orule.encode(l), d) for l, d in x]) for x in (pl, pm, pr))
            ops.append((OTensor.rand(OH(orule, [Vl, oP], [Vm]), 1000 + 2 * i),
                        OTensor.rand(OH(orule, [Vm], [oP, Vr]), 1001 + 2 * i)))

        def step():
            return [a.compose(b) for a, b in ops]
        sample = f"first {nsample} of the 256 merges (FP64 numpy oracle)"
    else:
        orule = make_rule("SU2")
        _, V = su2_fmove_space(8, 8)  # bounded sample: multiplicity floor(8/(2j+1))
        oV = OL(list(zip(V.ids, V.degs)))
        sp = OH(orule, [oV, oV], [oV, oV])
        B = OTensor.rand(sp, 2)
        if fmove == "swap":
            A = OTensor.rand(OH(orule, [oV, oV, oV], [oV]), 1)

            def step():
                return A.permute([0, 2, 1], [3]).contract([2, 3], B, [2, 0], [0, 1, 2, 3], 2)
        else:
            A = OTensor.rand(sp, 1)

            def step():
                return A.contract([1, 3], B, [2, 0], [0, 1, 2, 3], 2)
        sample = f"SU(2) F-move ({fmove} form) + contraction at multiplicity floor(8/(2j+1)) (bounded sample)"
    times = []
    out = None
    t_start = time.perf_counter()
    # at least min_runs runs and about 10 s of CPU work, at most budget_s
    while len(times) < min_runs or (time.perf_counter() - t_start) < min(10.0, budget_s):
        if len(times) >= min_runs and (time.perf_counter() - t_start) >= budget_s:
            break
        t = time.perf_counter()
        out = step()
        times.append(time.perf_counter() - t)
    times.sort()
    return times[len(times) // 2], cores, sample, out, len(times)


def cpu_serve_baseline(rule, trace, budget_s=20.0):
    """Oracle (numpy FP64) H_eff matvecs over the head of the same trace."""
    from oracle.fusion import HomSpace as OH, Leg as OL
    from oracle.symmetry import make_rule
    from oracle.tensor import OTensor
    from tenet_amd.workloads import mps_env_spaces
    orule = make_rule(rule)
    env = {}
    t_work = 0.0
    done = 0
    t_start = time.perf_counter()
    for _, D, seed in trace:
        if time.perf_counter() - t_start > budget_s:
            break
        if D not in env:
            _, V, P, M = mps_env_spaces(rule, D)
            lg = lambda g: OL(list(zip(g.ids, g.degs)), g.is_dual)
            oV, oP, oM = lg(V), lg(P), lg(M)
            env[D] = (OTensor.randn(OH(orule, [oV, oM], [oV]), 1), OTensor.randn(OH(orule, [oM, oP], [oP, oM]), 2),
                      OTensor.randn(OH(orule, [oM, oP], [oP, oM]), 3), OTensor.randn(OH(orule, [oV], [oV, oM]), 4),
                      OH(orule, [oV], [oP, oP, oV]))
        L, W1, W2, R, tsp = env[D]
        th = OTensor.randn(tsp, seed)
        t = time.perf_counter()
        x1 = L.contract([2], th, [0], [0, 1, 2, 3, 4], 2)
        x2 = x1.contract([1, 2], W1, [3, 1], [0, 3, 4, 1, 2], 2)
        x3 = x2.contract([1, 3], W2, [3, 1], [0, 3, 1, 4, 2], 2)
        y = x3.contract([1, 4], R, [2, 0], [0, 1, 2, 3], 1)
        th.inner(y)
        t_work += time.perf_counter() - t
        done += 1
    return done / t_work, done


def run_serve(args):
    """Continuous-batching H_eff server on a Poisson trace; one JSON line."""
    import numpy as np

    from tenet_amd.serving import HeffServer, poisson_trace
    t = time.perf_counter()
    mpo = "heisenberg" if args.rule == "U1" else "random"
    srv = HeffServer(args.rule, n_streams=args.streams, max_inflight=args.inflight,
                     precision=args.precision, mpo=mpo, graphs=args.serve_graphs == "on",
                     max_batch=args.max_batch)
    srv.warm()
    cold = time.perf_counter() - t
    log(f"[bench] serve: {len(srv.buckets)} bond-dimension buckets {srv.buckets}, environments resident, "
        f"compiled on {args.streams} streams in {cold:.1f}s")
    # warm-up pass over a short trace (pools, tables), then the timed trace
    warm = poisson_trace(max(64, args.warmup * 32), rate=args.rate, seed=6, buckets=srv.buckets)
    srv.serve(warm)
    trace = poisson_trace(args.requests, rate=args.rate, seed=5, buckets=srv.buckets)
    stats = srv.serve(trace)
    log(f"[bench] serve: {stats['requests']} requests in {stats['wall_s']:.2f}s "
        f"(offered {stats['offered_rate']:.0f}/s): {stats['matvecs_per_s']:.0f} matvecs/s, "
        f"p50 {stats['p50_ms']:.3f} ms, p99 {stats['p99_ms']:.3f} ms, {stats['tflops']:.2f} TFLOP/s; "
        f"host issue {stats['host_issue_ms_per_request']:.3f} ms/request, peak in flight {stats['max_inflight']}")
    # capacity: the same request mix offered far above what the server sustains
    sat = None
    if args.saturate > 0:
        sat_trace = poisson_trace(args.requests, rate=args.saturate, seed=8, buckets=srv.buckets)
        sat = srv.serve(sat_trace)
        log(f"[bench] serve capacity: {sat['matvecs_per_s']:.0f} matvecs/s at offered {sat['offered_rate']:.0f}/s, "
            f"{sat['tflops']:.2f} TFLOP/s, p50 {sat['p50_ms']:.3f} ms (queueing included), "
            f"host issue {sat['host_issue_ms_per_request']:.3f} ms/request")
    cpu = None
    if args.cpu_baseline == "on":
        rate, n = cpu_serve_baseline(args.rule, trace)
        try:
            from threadpoolctl import threadpool_info
            cores = max([i.get("num_threads", 1) for i in threadpool_info() if i.get("user_api") == "blas"] or [1])
        except Exception:  # pragma: no cover
            cores = os.cpu_count() or 1
        cpu = {"value": rate, "unit": "matvecs/s", "cores": int(cores), "kind": "port",
               "sample": f"first {n} requests of the same trace, served back to back by the FP64 numpy oracle "
                         f"(matvec + Rayleigh numerator, no queueing)"}
        log(f"[bench] cpu baseline {rate:.2f} matvecs/s over {n} requests")
    ds = np.array([d for _, d, _ in trace])
    rec = {
        "metric": "H_eff matvecs/s (continuous batching, p50/p99 latency)",
        "value": stats["matvecs_per_s"],
        "unit": "matvecs/s",
        "n_gpus": 1,
        "steps": stats["requests"],
        "warmup": len(warm),
        "ms_per_step": 1e3 / stats["matvecs_per_s"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f64": "f64", "f32": "f32 (f64 in/out)", "split_bf16": "split-bf16 (f64 in/out)",
                  "split_i8": "split-int8 (f64 in/out)"}[args.precision]
                 + ", f64 dot",
        "data": "synthetic (seeded normal(0,1) environments and states on device; "
                + ("spin-1/2 Heisenberg MPO)" if mpo == "heisenberg" else "seeded normal(0,1) MPO tensors)"),
        "config": {"workload": f"serve_two_site_heff_{args.rule.lower()}", "mpo": mpo, "arrivals": f"poisson {args.rate:.0f}/s",
                   "max_batch": args.max_batch,
                   "bond_dims": "lognormal median 512 sigma 0.7 clipped 32..4096, snapped to buckets",
                   "buckets": srv.buckets, "streams": args.streams, "slots": args.inflight,
                   "mean_D": float(ds.mean())},
        "latency_ms": {"p50": stats["p50_ms"], "p99": stats["p99_ms"], "mean": stats["mean_ms"]},
        "offered_rate": stats["offered_rate"],
        "tflops": stats["tflops"],
        "host_issue_ms_per_request": stats["host_issue_ms_per_request"],
        "batches": stats["batches"],
        "issue": "hipGraph replay of the four contractions per request" if args.serve_graphs == "on"
                 else "prebuilt launch lists",
        "capacity": None if sat is None else {
            "matvecs_per_s": sat["matvecs_per_s"], "offered_rate": sat["offered_rate"], "tflops": sat["tflops"],
            "latency_ms_incl_queueing": {"p50": sat["p50_ms"], "p99": sat["p99_ms"]},
            "host_issue_ms_per_request": sat["host_issue_ms_per_request"], "batches": sat["batches"]},
        "cpu_baseline": cpu,
        "cold_compile_s": cold,
    }
    print(json.dumps(rec), flush=True)
    srv.close()


# The reference's published warm microbench (µs / iteration; TeNeT column,
# Apple M4 Max, Accelerate, one thread), benchmarks/history/
# contraction-microbench-2026-07-11.md:51-59 (d = 4) and :68-76 (d = 16).
MICROBENCH_PUBLISHED = {("U1", 4): (4.10, 9.72, 12.43), ("SU2", 4): (7.19, 30.02, 44.63),
                        ("U1", 16): (1886.1, 2645.3, 3281.0), ("SU2", 16): (7741.3, 11915.0, 15075.6)}
# microbench_fusion.rs:24-50: (name, lhs axes, rhs axes, output axes)
MICROBENCH_WORKLOADS = (("compose", [2, 3], [0, 1], [0, 1, 2, 3]), ("swap", [3, 2], [0, 1], [0, 1, 2, 3]),
                        ("swap+out", [3, 2], [0, 1], [1, 0, 2, 3]))


def run_microbench(args):
    """The reference's contraction microbench (tenet-tensors/examples/
    microbench_fusion.rs): A, B in V (x) V <- V (x) V, uniform degeneracy d,
    U(1) {-1, 0, 1} and SU(2) {0, 1/2, 1}, fills (i % 17) 0.25 - 2 and
    (i % 13) 0.5 - 3, compose / swap / swap+out, warm.  Per cell: µs per
    iteration of back-to-back calls (the device stream keeps the calls in
    order; timed to a device synchronisation) and of calls each followed by
    a synchronisation, the checksum (sum of the destination), and the
    reference's published TeNeT figure beside it.  One JSON line."""
    import math as _m

    import numpy as np

    from tenet_amd import _native as N
    from tenet_amd.tensor import GradedSpace, HomSpace, Rule, Runtime, TensorMap
    rt = Runtime(0)
    h = N.hip()
    kn, kms, kw = C.c_size_t(), C.c_double(), C.c_double()
    cells = []
    for d in (4, 16):
        for sym, pairs in (("U1", [(-1, d), (0, d), (1, d)]), ("SU2", [(0, d), (1, d), (2, d)])):
            rule = Rule(sym)
            V = GradedSpace(rule, pairs)
            sp = HomSpace(rule, [V, V], [V, V])
            n = sp.len
            A = TensorMap.from_numpy(rt, sp, (np.arange(n) % 17) * 0.25 - 2.0)
            B = TensorMap.from_numpy(rt, sp, (np.arange(n) % 13) * 0.5 - 3.0)
            for (name, la, lb, perm), ref_us in zip(MICROBENCH_WORKLOADS, MICROBENCH_PUBLISHED[(sym, d)]):
                t = time.perf_counter()
                out = A.contract(la, B, lb, perm, 2)
                rt.synchronize()
                cold = time.perf_counter() - t
                for _ in range(3):
                    A.contract(la, B, lb, perm, 2, out=out)
                rt.synchronize()
                iters, t0 = 0, time.perf_counter()
                while time.perf_counter() - t0 < args.min_ms * 1e-3 or iters < 10:
                    A.contract(la, B, lb, perm, 2, out=out)
                    iters += 1
                rt.synchronize()
                us = (time.perf_counter() - t0) * 1e6 / iters
                ns, t1 = 0, time.perf_counter()
                while time.perf_counter() - t1 < args.min_ms * 1e-3 or ns < 10:
                    A.contract(la, B, lb, perm, 2, out=out)
                    rt.synchronize()
                    ns += 1
                us_sync = (time.perf_counter() - t1) * 1e6 / ns
                # the same calls straight through the C ABI (what a compiled
                # caller pays: no Python frontend), and the device time of the
                # kernels of one call (HIP events around each kernel)
                c = rt.contraction(A.space, B.space, la, lb, perm, 2)
                fn, args_ = h.tnt_contract, (rt._ctx, c.art, out._buf, A._buf, B._buf, 1.0, 0.0, 0)
                nn, t3 = 0, time.perf_counter()
                while time.perf_counter() - t3 < args.min_ms * 1e-3 or nn < 10:
                    fn(*args_)
                    nn += 1
                rt.synchronize()
                us_native = (time.perf_counter() - t3) * 1e6 / nn
                rt._check(h.tnt_ctx_reset_stats(rt._ctx))
                rt._check(h.tnt_ctx_set_timing(rt._ctx, 1))
                for _ in range(20):
                    fn(*args_)
                rt.synchronize()
                rt._check(h.tnt_ctx_set_timing(rt._ctx, 0))
                dev_ms = 0.0
                for kind in (0, 1):  # recoupling passes + FP64 GEMM table runs
                    rt._check(h.tnt_ctx_kernel_stats(rt._ctx, kind, C.byref(kn), C.byref(kms), C.byref(kw)))
                    dev_ms += kms.value
                checksum = float(out.to_numpy().sum())
                cells.append({"symmetry": sym, "d": d, "workload": name, "us_per_iter": us,
                              "us_per_call_synchronised": us_sync, "us_per_call_c_abi": us_native,
                              "device_us_per_call": dev_ms * 1e3 / 20, "iters": iters, "checksum": checksum,
                              "cold_ms": cold * 1e3, "reference_us": ref_us, "ratio_vs_reference": us / ref_us})
                log(f"[bench] microbench {sym} d={d} {name}: {us:.2f} us/iter ({us_sync:.2f} synchronised, "
                    f"{us_native:.2f} through the C ABI, {dev_ms * 1e3 / 20:.2f} on the device), "
                    f"checksum {checksum:.6f}; reference {ref_us} us")
    rt.close()
    gm = _m.exp(sum(_m.log(c["us_per_iter"]) for c in cells) / len(cells))
    gm_ref = _m.exp(sum(_m.log(c["reference_us"]) for c in cells) / len(cells))
    rec = {"metric": "contraction microbench us/iteration (warm), geometric mean over cells",
           "value": gm, "unit": "us", "n_gpus": 1, "steps": len(cells), "warmup": 3,
           "ms_per_step": gm * 1e-3, "higher_is_better": False, "scaling": "strong",
           "vs_baseline": gm / gm_ref, "dtype": "f64",
           "data": "synthetic (the reference's fills (i%17)*0.25-2 and (i%13)*0.5-3)",
           "config": {"workload": "microbench_fusion", "symmetries": "U1 {-1,0,1}, SU2 {0,1/2,1}",
                      "degeneracies": [4, 16], "workloads": [w[0] for w in MICROBENCH_WORKLOADS],
                      "min_ms_per_cell": args.min_ms},
           "baseline_source": "benchmarks/history/contraction-microbench-2026-07-11.md (TeNeT column, "
                              "Apple M4 Max, 1 thread)",
           "cells": cells}
    print(json.dumps(rec), flush=True)


def run_gemm_vs_vendor(args):
    """The core GEMM jobs of the U(1) environment update (L.A and Adag.T2 at
    bond dimension D) on two FP64 routes over the same operand sizes: this
    build's grouped MFMA kernel (one tnt_gemm_grouped call per contraction,
    tile shape picked by measurement as the autotuner does) and the vendor
    library behind torch.mm (one call per job).  One JSON line."""
    import torch

    from tenet_amd import _native as N
    from tenet_amd.tensor import HomSpace, Runtime, _HostContraction
    from tenet_amd.workloads import MPSEnvUpdate, mps_env_spaces
    r, V, P, M = mps_env_spaces("U1", args.D)
    L, A, W = HomSpace(r, [V, M], [V]), HomSpace(r, [V], [P, V]), HomSpace(r, [M, P], [P, M])
    c1 = _HostContraction(L, A, [2], [0], [0, 1, 2, 3], 2, False, False)
    c2 = _HostContraction(c1.dst_space, W, *MPSEnvUpdate.T2_AXES, False, False)
    c3 = _HostContraction(A, c2.dst_space, *MPSEnvUpdate.LP_AXES, True, False)
    groups = []
    for c in (c1, c3):
        d = c.desc.contents
        groups.append([(d.jobs[i].rows, d.jobs[i].contracted, d.jobs[i].cols) for i in range(d.n_jobs)])
    shapes = [s_ for g_ in groups for s_ in g_]
    flops = float(sum(2 * m * k * n for m, k, n in shapes))
    rt = Runtime(0)
    dev = torch.device("cuda", 0)
    g = torch.Generator(device=dev).manual_seed(1)
    # column-major (m x k) A, (k x n) B, (m x n) C are row-major (k, m), (n, k),
    # (n, m) tensors: the grouped kernel runs the jobs as the environment update
    # does, and torch.mm(B^T, A^T) = C^T computes the same products
    ops = [[(torch.randn(k, m, dtype=torch.float64, device=dev, generator=g),
             torch.randn(n, k, dtype=torch.float64, device=dev, generator=g),
             torch.empty(n, m, dtype=torch.float64, device=dev)) for m, k, n in grp] for grp in groups]
    jobs = [(N.GemmPtrJob * len(grp))(*[
        N.GemmPtrJob(cm.data_ptr(), a.data_ptr(), b.data_ptr(), m, k, n, m, m, k, 0, 0)
        for (m, k, n), (a, b, cm) in zip(grp, o)]) for grp, o in zip(groups, ops)]
    h = N.hip()

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / args.steps

    t_ours, cfgs = 0.0, []
    for jb, grp in zip(jobs, groups):
        best = None
        for cfg in (-1, 0, 1, 2, 3, 5, 6):
            h.tnt_set_tile_config(rt._ctx, cfg)
            t = timed(lambda: rt._check(h.tnt_gemm_grouped(rt._ctx, jb, len(grp), 1.0, 0.0, 0)))
            if best is None or t < best[0]:
                best = (t, cfg)
        t_ours += best[0]
        cfgs.append(GEMM_KERNELS[best[1]].format("double") if best[1] >= 0 else "size rule")
    h.tnt_set_tile_config(rt._ctx, -1)

    def vendor():
        for o in ops:
            for a, b, cm in o:
                torch.mm(b, a, out=cm)

    t_vendor = timed(vendor)
    ref = [cm.clone() for o in ops for _, _, cm in o]
    torch.cuda.synchronize()  # the copies (torch stream) before the kernel (backend stream)
    for jb, grp in zip(jobs, groups):
        rt._check(h.tnt_gemm_grouped(rt._ctx, jb, len(grp), 1.0, 0.0, 0))
    torch.cuda.synchronize()
    outs = [cm for o in ops for _, _, cm in o]
    err = max(float((cm - r_).abs().max() / r_.abs().max()) for cm, r_ in zip(outs, ref))
    log(f"[bench] {len(shapes)} FP64 sector GEMMs, {flops / 1e9:.2f} GFLOP: grouped MFMA kernel "
        f"{t_ours * 1e3:.3f} ms ({flops / t_ours / 1e12:.1f} TFLOP/s; tiles {cfgs}), torch.mm per job "
        f"{t_vendor * 1e3:.3f} ms ({flops / t_vendor / 1e12:.1f} TFLOP/s); max rel diff {err:.1e}")
    rec = {"metric": "FP64 sector GEMM TFLOP/s (grouped MFMA kernel vs vendor torch.mm)",
           "value": flops / t_ours / 1e12, "unit": "TFLOP/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": t_ours * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic (torch normal(0,1) operands)",
           "config": {"workload": "u1_env_core_gemms", "bond_dim": args.D, "jobs": len(shapes),
                      "tiles": cfgs},
           "vendor": {"route": "torch.mm per job (hipBLASLt / rocBLAS)", "tflops": flops / t_vendor / 1e12,
                      "ms_per_step": t_vendor * 1e3},
           "max_rel_diff": err}
    print(json.dumps(rec), flush=True)
    rt.close()


# The other configurations, measured after the headline one in the default
# run (each in a child process with its own time limit, so a failure there
# leaves the headline record intact): the SU(2) environment update at
# D = 4096 (D counted as the leg dimension, and as multiplets), the SU(2)
# F-move (FP64 and split-INT8), the batched merges, the
# served H_eff trace (U(1) and SU(2)) and the reference's microbench.
SECONDARY = (  # (name, arguments, time limit in seconds; each takes 5-25 s)
    ("su2_env_f64", ["--workload", "su2_env"], 90),
    ("su2_env_multiplets_f64", ["--workload", "su2_env", "--su2-count", "multiplets"], 90),
    ("su2_fmove_f64", ["--workload", "su2_fmove", "--steps", "10", "--warmup", "2"], 150),
    ("su2_fmove_split_i8", ["--workload", "su2_fmove", "--precision", "split_i8", "--steps", "10",
                            "--warmup", "2", "--cpu-baseline", "off"], 90),
    ("su2_fmove_cross_split_i8", ["--workload", "su2_fmove", "--fmove", "cross", "--precision", "split_i8",
                                  "--steps", "10", "--warmup", "2", "--cpu-baseline", "off"], 90),
    ("batched_f32", ["--workload", "batched", "--precision", "f32"], 90),
    ("serve_u1", ["--workload", "serve", "--rule", "U1", "--cpu-baseline", "off"], 90),
    ("serve_su2", ["--workload", "serve", "--rule", "SU2", "--cpu-baseline", "off"], 90),
    ("microbench", ["--workload", "microbench", "--min-ms", "100"], 90),
)
SECONDARY_BUDGET_S = 300.0  # configurations still waiting when this is spent are skipped


def _summary(rec):
    """The figures of a child bench record kept in the headline record."""
    keep = ("metric", "value", "unit", "ms_per_step", "dtype", "parity_rel_err_vs_oracle",
            "parity_rel_err_vs_cpu_port_full_size", "rel_err_vs_f64", "latency_ms", "offered_rate", "tflops")
    out = {k: rec[k] for k in keep if k in rec and rec[k] is not None}
    out["workload"] = rec.get("config", {}).get("workload")
    r = rec.get("roofline")
    if r:
        out["roofline"] = {k: r.get(k) for k in ("kernel", "achieved", "peak", "unit", "frac", "traffic")}
    if rec.get("recoupling"):
        out["recoupling_gbs"] = rec["recoupling"]["achieved"]
    c = rec.get("cpu_baseline")
    if c:
        out["cpu_baseline"] = {k: c.get(k) for k in ("value", "unit", "cores", "kind", "cpu_model")}
    if rec.get("capacity"):
        out["capacity_matvecs_per_s"] = rec["capacity"]["matvecs_per_s"]
    if rec.get("r_only_swap"):
        out["r_only_swap_gbs"] = rec["r_only_swap"]["achieved"]
    if rec.get("cells"):
        out["cells"] = [{"case": f"{x['symmetry']} d={x['d']} {x['workload']}", "us": round(x["us_per_iter"], 2),
                         "us_c_abi": round(x["us_per_call_c_abi"], 2), "reference_us": x["reference_us"],
                         "checksum": x["checksum"]} for x in rec["cells"]]
    return out


def run_secondary():
    """Run the SECONDARY configurations one after another; {name: summary or error}."""
    import subprocess
    out = {}
    here = os.path.dirname(os.path.abspath(__file__))
    t0 = time.perf_counter()
    for name, extra, limit in SECONDARY:
        left = SECONDARY_BUDGET_S - (time.perf_counter() - t0)
        if left < 10.0:
            out[name] = {"error": "skipped: the secondary time budget is spent"}
            continue
        limit = min(limit, left)
        log(f"[bench] secondary: {name}")
        cmd = [sys.executable, os.path.join(here, "bench.py"), "--secondary", "off"] + extra
        try:
            p = subprocess.run(cmd, capture_output=True, text=True, timeout=limit)
        except subprocess.TimeoutExpired:
            out[name] = {"error": f"timed out after {limit}s"}
            log(f"[bench] secondary {name}: timed out")
            continue
        lines = [ln for ln in p.stdout.splitlines() if ln.startswith("{")]
        if p.returncode != 0 or not lines:
            out[name] = {"error": f"exit {p.returncode}: {p.stderr.strip().splitlines()[-1:]}"}
            log(f"[bench] secondary {name}: failed ({p.returncode})")
            continue
        out[name] = _summary(json.loads(lines[-1]))
        log(f"[bench] secondary {name}: {out[name].get('value')} {out[name].get('unit')}")
    return out


def run_complex_gemm(args):
    """The core GEMM jobs of the U(1) environment update (as gemm_vs_vendor)
    on complex (C64) payloads: one tnt_gemm_batch_complex call per
    contraction (expand pass, real m x 2k x 2n grouped MFMA GEMM, combine pass)
    against the vendor complex GEMM behind torch.mm (one call per job), on the
    same device buffers.  Complex flops = 8 mkn.  One JSON line."""
    import torch

    from tenet_amd import _native as N
    from tenet_amd.tensor import HomSpace, Runtime, _HostContraction
    from tenet_amd.workloads import MPSEnvUpdate, mps_env_spaces
    r, V, P, M = mps_env_spaces("U1", args.D)
    L, A, W = HomSpace(r, [V, M], [V]), HomSpace(r, [V], [P, V]), HomSpace(r, [M, P], [P, M])
    c1 = _HostContraction(L, A, [2], [0], [0, 1, 2, 3], 2, False, False)
    c2 = _HostContraction(c1.dst_space, W, *MPSEnvUpdate.T2_AXES, False, False)
    c3 = _HostContraction(A, c2.dst_space, *MPSEnvUpdate.LP_AXES, True, False)
    groups = []
    for c in (c1, c3):
        d = c.desc.contents
        groups.append([(d.jobs[i].rows, d.jobs[i].contracted, d.jobs[i].cols) for i in range(d.n_jobs)])
    shapes = [s_ for g_ in groups for s_ in g_]
    flops = float(sum(8 * m * k * n for m, k, n in shapes))
    precision = args.precision or "f64"
    prec = {"f64": 0, "f32": 1, "split_bf16": 2, "split_i8": 3}[precision]
    rt = Runtime(0)
    h = N.hip()
    dev = torch.device("cuda", 0)
    g = torch.Generator(device=dev).manual_seed(1)
    calls = []
    for grp in groups:
        nl = sum(m * k for m, k, n in grp)
        nr = sum(k * n for m, k, n in grp)
        nd = sum(m * n for m, k, n in grp)
        tl = torch.randn(nl, dtype=torch.complex128, device=dev, generator=g)
        tr = torch.randn(nr, dtype=torch.complex128, device=dev, generator=g)
        td = torch.zeros(nd, dtype=torch.complex128, device=dev)
        js, views, lo, ro, do = [], [], 0, 0, 0
        for m, k, n in grp:
            js.append(N.GemmJob(do, lo, ro, m, k, n))
            # column-major (m x k) A, (k x n) B, (m x n) C are row-major (k, m),
            # (n, k), (n, m) tensors, and torch.mm(B^T, A^T) = C^T
            views.append((tl[lo:lo + m * k].view(k, m), tr[ro:ro + k * n].view(n, k),
                          td[do:do + m * n].view(n, m)))
            lo, ro, do = lo + m * k, ro + k * n, do + m * n
        bufs = []
        for t_ in (td, tl, tr):
            b = C.c_void_p()
            rt._check(h.tnt_buf_wrap(rt._ctx, C.c_void_p(t_.data_ptr()), t_.numel(), N.C64, C.byref(b)))
            bufs.append(b)
        calls.append(((N.GemmJob * len(js))(*js), len(js), bufs, views, td))
    torch.cuda.synchronize()

    def one(c):
        arr, nj, (bd, bl, br), _, _ = c
        rt._check(h.tnt_gemm_batch_complex(rt._ctx, bd, bl, br, arr, nj, None, 0, 0, 0, 1.0, 0.0, 0.0, 0.0, prec))

    def vendor():
        for _, _, _, views, _ in calls:
            for a, b, cm in views:
                torch.mm(b, a, out=cm)

    def timed(fn, sync):
        for _ in range(args.warmup):
            fn()
        sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        sync()
        return (time.perf_counter() - t0) / args.steps

    t_vendor = timed(vendor, torch.cuda.synchronize)
    ref = [c[4].clone() for c in calls]
    for c in calls:
        c[4].zero_()
    torch.cuda.synchronize()  # torch's stream before the backend's
    # tile shape per contraction picked by measurement, as the autotuner does
    t_ours, cfgs = 0.0, []
    for c in calls:
        best = None
        for cfg in (-1, 0, 1, 2, 3, 5, 6):
            h.tnt_set_tile_config(rt._ctx, cfg)
            t = timed(lambda: one(c), rt.synchronize)
            if best is None or t < best[0]:
                best = (t, cfg)
        t_ours += best[0]
        cfgs.append(GEMM_KERNELS[best[1]].format("double") if best[1] >= 0 else "size rule")
    h.tnt_set_tile_config(rt._ctx, -1)
    err = max(float(torch.linalg.norm(c[4] - r_) / torch.linalg.norm(r_)) for c, r_ in zip(calls, ref))
    log(f"[bench] {len(shapes)} C64 sector GEMMs, {flops / 1e9:.2f} GFLOP (8mkn): tnt_gemm_batch_complex "
        f"({precision}) {t_ours * 1e3:.3f} ms ({flops / t_ours / 1e12:.1f} TFLOP/s; tiles {cfgs}), torch.mm per "
        f"job {t_vendor * 1e3:.3f} ms ({flops / t_vendor / 1e12:.1f} TFLOP/s); rel Frobenius diff {err:.1e}")
    rec = {"metric": "C64 sector GEMM TFLOP/s (8mkn; tnt_gemm_batch_complex vs vendor torch.mm)",
           "value": flops / t_ours / 1e12, "unit": "TFLOP/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": t_ours * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "c64", "data": "synthetic (torch complex normal operands)",
           "config": {"workload": "u1_env_core_gemms_c64", "bond_dim": args.D, "jobs": len(shapes),
                      "precision": precision, "tiles": cfgs},
           "vendor": {"route": "torch.mm per job (complex128, hipBLASLt / rocBLAS)",
                      "tflops": flops / t_vendor / 1e12, "ms_per_step": t_vendor * 1e3},
           "rel_frobenius_diff_vs_vendor": err}
    print(json.dumps(rec), flush=True)
    for c in calls:
        for b in c[2]:
            h.tnt_buf_free(rt._ctx, b)
    rt.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--settle-ms", type=float, default=100.0,
                    help="back-to-back steps before each timed region so the GPU clock settles (0: none)")
    ap.add_argument("--workload", default="u1_env",
                    choices=["u1_env", "su2_env", "su2_fmove", "batched", "serve", "copy_calibration",
                             "gemm_vs_vendor", "complex_gemm", "microbench"])
    ap.add_argument("--min-ms", type=float, default=200.0, help="microbench: timed milliseconds per cell")
    ap.add_argument("--count", type=int, default=256, help="batched: number of contractions")
    ap.add_argument("--precision", default=None, choices=["f64", "f32", "split_bf16", "split_i8"],
                    help="GEMM precision (default: split_i8 for serve, f64 otherwise)")
    ap.add_argument("--D", type=int, default=4096)
    ap.add_argument("--su2-count", default="dimension", choices=["dimension", "multiplets"],
                    help="su2_env: D is the leg dimension, or the number of SU(2) multiplets")
    ap.add_argument("--mult", type=int, default=64, help="su2_fmove multiplicity scale")
    ap.add_argument("--fmove", default="swap", choices=["swap", "cross"],
                    help="su2_fmove form: codomain swap of a 3<-1 tensor then contract, or A[1,3].B[2,0]")
    ap.add_argument("--cpu-baseline", default="on", choices=["on", "off"])
    ap.add_argument("--rule", default="U1", choices=["U1", "SU2"], help="serve: symmetry of the chain")
    ap.add_argument("--requests", type=int, default=4000, help="serve: requests in the timed trace")
    ap.add_argument("--rate", type=float, default=2000.0, help="serve: mean arrival rate (req/s)")
    ap.add_argument("--streams", type=int, default=4, help="serve: backend contexts (HIP streams)")
    ap.add_argument("--inflight", type=int, default=16, help="serve: request slots")
    ap.add_argument("--max-batch", type=int, default=8, help="serve: largest batch of same-bucket requests")
    ap.add_argument("--serve-graphs", default="on", choices=["on", "off"],
                    help="serve: issue each request's contractions as one hipGraph replay")
    ap.add_argument("--saturate", type=float, default=64000.0,
                    help="serve: arrival rate of a second trace that measures capacity (0: skip)")
    ap.add_argument("--graph", default="off", choices=["on", "off"],
                    help="replay the step from a hipGraph in the throughput region (env / F-move workloads)")
    ap.add_argument("--autotune", default="on", choices=["on", "off"],
                    help="pick GEMM tile shapes per contraction by measurement before timing")
    ap.add_argument("--secondary", default=None, choices=["on", "off"],
                    help="after the headline record, measure the other configurations in child processes "
                         "and add their figures under 'secondary' (default: on for the default workload)")
    args = ap.parse_args()
    if args.secondary is None:
        args.secondary = "on" if args.workload == "u1_env" else "off"
    if args.precision is None:
        # the serving configuration runs split-INT8 GEMMs with FP64 dot products
        args.precision = "split_i8" if args.workload == "serve" else "f64"
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (the contraction would shard by "
              "coupled sector, with one all-to-all per leg-moving recoupling). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "serve":
        run_serve(args)
        return
    if args.workload == "gemm_vs_vendor":
        run_gemm_vs_vendor(args)
        return
    if args.workload == "complex_gemm":
        run_complex_gemm(args)
        return
    if args.workload == "microbench":
        run_microbench(args)
        return

    from tenet_amd import _native as N
    from tenet_amd.tensor import Runtime
    from tenet_amd.workloads import BatchedMerge, MPSEnvUpdate, SU2FMoveContract

    rt = Runtime(0)
    batched_err = None
    batched_cfg = None
    if args.workload == "copy_calibration":
        # y = x over 2^27 doubles (1 GiB read + 1 GiB written per launch): the
        # reference point that scales FETCH_SIZE / WRITE_SIZE to bytes
        from tenet_amd.tensor import GradedSpace, HomSpace, Rule, TensorMap
        rule = Rule("U1")
        V = GradedSpace(rule, [(0, 1 << 27)])
        W = GradedSpace(rule, [(0, 1)])
        sp = HomSpace(rule, [V], [W])
        x = TensorMap.rand(rt, sp, seed=1)
        y = TensorMap.zeros(rt, sp)
        for _ in range(args.warmup):
            N.hip().tnt_axpby(rt._ctx, y._buf, x._buf, 1.0, 0.0)
        rt.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            N.hip().tnt_axpby(rt._ctx, y._buf, x._buf, 1.0, 0.0)
        rt.synchronize()
        dt = (time.perf_counter() - t) / max(1, args.steps)
        log(f"[bench] copy calibration: {args.steps} launches of {8 << 27} bytes read + written, "
            f"{2 * (8 << 27) / dt / 1e9:.0f} GB/s")
        rt.close()
        return
    t = time.perf_counter()
    if args.workload == "u1_env":
        wl = MPSEnvUpdate(rt, "U1", args.D, args.precision)
        cfg = {"workload": "u1_mps_env_update", "bond_dim": args.D, "charges": "-10..10",
               "degeneracies": "gaussian sigma=4 (largest remainder)", "phys": "U1 {0,1}",
               "mpo_bond": "U1 {0:3,+1:1,-1:1}", "steps_per_update": "L.A, (.)W, Adag.(.)"}
    elif args.workload == "su2_env":
        wl = MPSEnvUpdate(rt, "SU2", args.D, args.precision, su2_count=args.su2_count)
        cfg = {"workload": "su2_mps_env_update" + ("_multiplets" if args.su2_count == "multiplets" else ""),
               "bond_dim": args.D,
               "bond_dim_counts": ("multiplets: sum_j m_j = D (leg dimension %d)" % wl.V.dim()
                                   if args.su2_count == "multiplets" else "leg dimension: sum_j m_j (2j+1) = D"),
               "spins": "2j=0..20", "multiplicities": "gaussian in j, sigma=3 (largest remainder)",
               "phys": "spin-1/2", "mpo_bond": "SU2 {0:2, 1:1}"}
    elif args.workload == "batched":
        import numpy as np
        wl = BatchedMerge(rt, args.count, precision=args.precision)
        cfg = {"workload": "batched_u1xz2_two_site_merge", "contractions": args.count,
               "bond_dims": "uniform 64..1024 (seed 4)", "charges": "-6..6 x parity",
               "phys": "U1xZ2 {(0,0),(1,1)}", "launch": "one grouped GEMM launch per tile shape, hipGraph replay",
               "gemm_jobs": wl.n_jobs}
        # FP64 reference of the same batch for the error of reduced-precision payloads
        wl.run("f64")
        rt.synchronize()
        ref = [r.copy() for r in wl.results()]
        wl.run()
        rt.synchronize()
        got = wl.results()
        num = sum(float(np.linalg.norm(g - r) ** 2) for g, r in zip(got, ref))
        den = sum(float(np.linalg.norm(r) ** 2) for r in ref)
        batched_err = (num / den) ** 0.5
        if args.autotune == "on":
            # tile shape of the grouped launch by measurement of graph replays: the
            # size rule, all 64 x 64 workgroup tiles, or all one-wave 32 x 32 tiles
            hh = N.hip()
            trials = []
            for c in (-1, 1, 6):
                hh.tnt_set_tile_config(rt._ctx, c)
                wl.graph = None
                wl.capture()
                for _ in range(3):
                    wl.run()
                rt.synchronize()
                t_ = time.perf_counter()
                for _ in range(10):
                    wl.run()
                rt.synchronize()
                trials.append(((time.perf_counter() - t_) / 10, c, wl.graph))
                log(f"[autotune] batched cfg {c}: {trials[-1][0] * 1e3:.3f} ms per graph replay")
            best = min(trials, key=lambda x: x[0])
            for _, c, g in trials:
                if g is not best[2]:
                    rt._check(hh.tnt_graph_destroy(rt._ctx, g))
            batched_cfg = best[1]
            wl.graph = best[2]
            hh.tnt_set_tile_config(rt._ctx, batched_cfg)  # kept for the event-timed eager region
        else:
            wl.capture()
        log(f"[bench] batched: {wl.n_jobs} sector GEMMs, {args.precision} rel. Frobenius error vs FP64 {batched_err:.3e}")
    elif args.fmove == "swap":
        from tenet_amd.workloads import SU2FMoveSwapContract
        wl = SU2FMoveSwapContract(rt, args.mult, 8, args.precision)
        cfg = {"workload": "su2_fmove_swap_contract", "spins": "2j=0..8", "multiplicity": f"floor({args.mult}/(2j+1))",
               "contraction": "A' = permute(A, (0,2,1), (3,)) [V(x)V(x)V <- V, inner line recoupled], "
                              "C = A'[2,3] . B[2,0]",
               "fmove_groups": {"single": wl.transform.n_single, "multi": wl.transform.n_multi}}
    else:
        wl = SU2FMoveContract(rt, args.mult, 8, args.precision)
        cfg = {"workload": "su2_fmove_contract", "spins": "2j=0..8", "multiplicity": f"floor({args.mult}/(2j+1))",
               "contraction": "A[1,3] . B[2,0] (both operands recoupled)"}
    cold = time.perf_counter() - t
    log(f"[bench] {cfg['workload']} built + compiled in {cold:.2f}s, {wl.flops / 1e9:.2f} GFLOP/step")
    for _ in range(args.warmup):
        wl.run()
    rt.synchronize()
    h = N.hip()
    tuned = {"batched": batched_cfg} if batched_cfg is not None else None
    if args.autotune == "on" and args.workload != "batched":
        from tenet_amd.autotune import autotune
        settle(wl.run, rt.synchronize, args.settle_ms / 2)
        conts = list(wl.contractions) if hasattr(wl, "contractions") else [wl.contraction]
        tuned = autotune(rt, conts, wl.run, reps=max(2, min(5, args.steps // 4)), log=log,
                         graph_step=wl if hasattr(wl, "capture") else None)
    # Timed region 1: K steps as a caller runs them -> value, ms_per_step.  The
    # step's launches are replayed from a hipGraph (captured after tuning) where
    # the workload supports it; each replay runs every kernel of the step.
    use_graph = args.graph == "on" and hasattr(wl, "capture")
    step = wl.run
    if use_graph:
        wl.capture()
        step = wl.replay
        for _ in range(2):
            step()
    # The GPU's clock follows its load over tens of milliseconds: after the
    # host-side gaps of set-up and tuning, short steps would otherwise be timed
    # while the clock is still ramping (measured: ~2.0 GHz in the first
    # milliseconds of load, 2.38 GHz once settled).  Steps run back to back
    # for --settle-ms first, as they do inside a sweep.
    settle_steps = settle(step, rt.synchronize, args.settle_ms)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    rt.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    # Timed region 2: the same K steps with every kernel class bracketed by HIP
    # events on the backend stream -> roofline / recoupling (the events widen the
    # gaps between kernels, so this region's wall time is reported beside only).
    settle(wl.run, rt.synchronize, args.settle_ms / 4)
    rt._check(h.tnt_ctx_reset_stats(rt._ctx))
    rt._check(h.tnt_ctx_set_timing(rt._ctx, 1))
    rt.synchronize()
    t2 = time.perf_counter()
    for _ in range(args.steps):
        wl.run()  # eager: graph replays are not bracketed per kernel
    rt.synchronize()
    ms_events = (time.perf_counter() - t2) * 1e3 / args.steps
    rt._check(h.tnt_ctx_set_timing(rt._ctx, 0))
    value = wl.flops / (ms * 1e-3) / 1e12
    prec_id = {"f64": 0, "f32": 1, "split_bf16": 2, "split_i8": 3}[args.precision]
    n, tms, work = C.c_size_t(), C.c_double(), C.c_double()

    def stats(kind):
        rt._check(h.tnt_ctx_kernel_stats(rt._ctx, kind, C.byref(n), C.byref(tms), C.byref(work)))
        return n.value, tms.value, work.value

    # all block-GEMM table runs (every launch, split pre-passes included)
    gemm_launches, gemm_ms, gemm_work = stats(1 + prec_id)
    gemm_tflops = gemm_work / (gemm_ms * 1e-3) / 1e12 if gemm_ms else 0.0
    # dominant kernel: the GEMM kernel (one tile shape) with the most time
    if args.workload == "batched":
        # the step is one hipGraph replay of the grouped launch(es); graph replays are
        # not bracketed per kernel, so the step time (device-synchronised) stands in
        dom = {"kernel": f"grouped GEMM hipGraph replay ({args.precision})", "launches": args.steps,
               "ms": ms * args.steps, "flops": wl.flops * args.steps, "bytes": wl.bytes * args.steps}
    elif args.precision in ("split_bf16", "split_i8"):
        kind = "split-BF16 table run (digit pre-pass + bf16 MFMA GEMM" if args.precision == "split_bf16" else \
            "split-INT8 table run (exponent + digit pre-passes + int8 MFMA GEMM"
        dom = {"kernel": kind + "; FP64 kernels for narrow jobs)",
               "launches": gemm_launches, "ms": gemm_ms, "flops": gemm_work}
    else:
        per = [(stats(8 + 8 * prec_id + c), c) for c in range(len(GEMM_KERNELS))]
        (dl, dms, dw), dc = max(per, key=lambda x: x[0][1])
        dom = {"kernel": GEMM_KERNELS[dc].format("double" if prec_id == 0 else "float"),
               "launches": dl, "ms": dms, "flops": dw}
    xf_launches, xf_ms, xf_bytes = stats(0)
    xf_gbs = xf_bytes / (xf_ms * 1e-3) / 1e9 if xf_ms else 0.0
    log(f"[bench] {ms:.3f} ms/step ({ms_events:.3f} with kernel events), {value:.2f} TFLOP/s; GEMM {gemm_tflops:.2f} TFLOP/s over "
        f"{gemm_launches} table runs ({gemm_ms / args.steps:.3f} ms/step); recoupling {xf_gbs:.0f} GB/s over "
        f"{xf_launches} launches ({xf_ms / args.steps:.3f} ms/step); dominant kernel {dom['kernel']}: "
        f"{dom['launches']} launches, {dom['ms'] / max(1, dom['launches']):.4f} ms avg")

    r_swap = None
    if args.workload == "su2_fmove":
        # the plain 2 <- 2 codomain swap: an R-move only (Single groups), timed
        # apart from the step, on a tensor of the same leg space
        from tenet_amd.tensor import TensorMap as _TM
        X = _TM.rand(rt, [wl.V, wl.V], [wl.V, wl.V], seed=7)
        tr = rt.transform(X.space, (1, 0), (2, 3))
        Y = X.permute((1, 0), (2, 3))
        rt.synchronize()
        rt._check(h.tnt_ctx_reset_stats(rt._ctx))
        rt._check(h.tnt_ctx_set_timing(rt._ctx, 1))
        for _ in range(args.steps):
            X.permute((1, 0), (2, 3), out=Y)
        rt.synchronize()
        rt._check(h.tnt_ctx_set_timing(rt._ctx, 0))
        rl, rms, rbytes = stats(0)
        r_gbs = rbytes / (rms * 1e-3) / 1e9 if rms else 0.0
        r_swap = {"transform": "permute(A, (1,0), (2,3)) on V(x)V <- V(x)V", "single": tr.n_single,
                  "multi": tr.n_multi, "launches": rl, "ms_per_call": rms / args.steps,
                  "bytes_per_call": rbytes / args.steps, "achieved": r_gbs, "peak": HBM_PEAK_GBS,
                  "unit": "GB/s", "frac": r_gbs / HBM_PEAK_GBS}
        log(f"[bench] R-only 2<-2 swap: {tr.n_single} single / {tr.n_multi} multi groups, "
            f"{rms / args.steps:.3f} ms per call, {r_gbs:.0f} GB/s")
        del X, Y

    cpu = None
    parity = None
    parity_port = None
    if args.cpu_baseline == "on":
        import numpy as np

        from oracle import cpu_port as P
        # the C++/OpenMP port at full size: the baseline value, and a full-size
        # parity check of the device result (same descriptors, same payloads)
        pt, pflops, pthreads, psample, pruns, pout = cpu_port_baseline(args.workload, wl)
        if args.workload == "batched":
            got = wl.results()
            parity_port = (sum(float(np.linalg.norm(g - o) ** 2) for g, o in zip(got, pout))
                           / sum(float(np.linalg.norm(o) ** 2) for o in pout)) ** 0.5
        else:
            got = wl.run().to_numpy()
            parity_port = float(np.linalg.norm(got - pout) / np.linalg.norm(pout))
        del pout, got
        log(f"[bench] cpu port: {pflops / pt / 1e9:.1f} GFLOP/s on {pthreads} threads ({P.cpu_model()}), "
            f"median of {pruns} run(s) of {pt:.2f}s; GPU vs port rel err {parity_port:.2e}")
        # the numpy oracle (an independent restatement of structure and replay):
        # the parity checker; timed once beside the port
        t_cpu, cores, sample, out, reps = cpu_baseline(args.workload, args.D,
                                                       wl=wl if args.workload in ("u1_env", "su2_env") else None,
                                                       fmove=args.fmove, budget_s=0.0, min_runs=1,
                                                       su2_count=args.su2_count)
        if args.workload == "batched":
            nsample = len(out)
            flops_sample = sum(c.flops for *_, c in wl.pairs[:nsample])
            got = wl.results()[:nsample]
            num = sum(float(np.linalg.norm(g - o.data) ** 2) for g, o in zip(got, out))
            den = sum(float(np.linalg.norm(o.data) ** 2) for o in out)
            parity = (num / den) ** 0.5
            cpu_val = flops_sample / t_cpu / 1e12
        elif args.workload == "su2_fmove":
            flops_sample = None
            from tenet_amd.tensor import Runtime as _R  # noqa: F401
            from tenet_amd.workloads import SU2FMoveContract as _S, SU2FMoveSwapContract as _S3
            small = (_S3 if args.fmove == "swap" else _S)(rt, 8, 8, "f64")
            flops_sample = small.flops
            got = small.run().to_numpy()
            parity = float(np.linalg.norm(got - out.data) / np.linalg.norm(out.data))
            cpu_val = flops_sample / t_cpu / 1e12
        else:
            got = wl.run().to_numpy()
            parity = float(np.linalg.norm(got - out.data) / np.linalg.norm(out.data))
            cpu_val = wl.flops / t_cpu / 1e12
        cpu = {"value": pflops / pt / 1e12, "unit": "TFLOP/s", "cores": int(pthreads), "kind": "port",
               "sample": f"{psample}; C++/OpenMP port of the host route (oracle/cpu_port.cpp), FP64, "
                         f"median of {pruns} warm run(s) of {pt:.2f}s",
               "cpu_model": P.cpu_model(), "gemm": "cache-blocked DGEMM of the port (AVX2+FMA), no system BLAS",
               "numpy_oracle": {"value": cpu_val, "unit": "TFLOP/s", "blas_threads": int(cores),
                                "sample": f"{sample}; one run of {t_cpu:.2f}s"}}
        log(f"[bench] numpy oracle {cpu_val * 1e3:.2f} GFLOP/s on {cores} BLAS threads; GPU vs oracle rel err {parity:.2e}")

    rec = {
        "metric": "contract TFLOP/s (FP64 block-sparse contraction incl. recoupling)",
        "value": value,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f64": "f64", "f32": "f32 (f64 in/out)", "split_bf16": "split-bf16 (f64 in/out)",
                  "split_i8": "split-int8 (f64 in/out)"}[args.precision],
        "data": "synthetic (seeded normal(0,1) on device)",
        "config": cfg,
        "roofline": roofline(args.precision, dom, cfg["workload"] + (
            f" D={args.D}" if "bond_dim" in cfg else f" mult={args.mult}" if "multiplicity" in cfg else "")),
        "gemm_all_launches": {"tflops": gemm_tflops, "launches": gemm_launches,
                              "ms_per_step": gemm_ms / args.steps},
        "ms_per_step_with_kernel_events": ms_events,
        "clock_settle": {"ms": args.settle_ms, "steps": settle_steps,
                         "note": "steps run back to back before the timed region (not timed, not counted in warmup)"},
        "launch": "hipGraph replay of the step" if use_graph else "eager launches",
        "recoupling": {"bound": "hbm", "achieved": xf_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": xf_gbs / HBM_PEAK_GBS, "launches": xf_launches,
                       "avg_launch_ms": xf_ms / max(1, xf_launches),
                       "bytes_per_step": xf_bytes / max(1, args.steps)},
        "cpu_baseline": cpu,
        "r_only_swap": r_swap,
        "parity_rel_err_vs_oracle": parity,
        "parity_rel_err_vs_cpu_port_full_size": parity_port,
        "cold_compile_s": cold,
        "tile_configs": tuned,
        "rel_err_vs_f64": batched_err if args.workload == "batched" else None,
    }
    if args.secondary == "on":
        rt.close()
        rec["secondary"] = run_secondary()
        print(json.dumps(rec), flush=True)
        return
    print(json.dumps(rec), flush=True)
    rt.close()


if __name__ == "__main__":
    main()
