#!/usr/bin/env python3
"""Headline benchmark: batched in-place FP32 LU inversion on one MI355X.

Workload (the metric's configuration): 1,000,000 matrices of n=32, entries
i.i.d. U[-1,1) from mt19937(42), partial pivoting by the reference's rule
(LUB_PIVOT_PARALLEL), inverted in place with the int32 pivot vectors written
out.  One step = one lu_batched_inverse launch over the whole batch.

Timing: every timed step gets its own pristine copy of the input, already
resident in HBM (288 GB leaves room for K copies), so the K steps run back to
back inside one region bracketed by stream synchronisation, with a HIP event
pair around each launch for the per-launch duration used by the roofline.

Prints one JSON line on stdout; diagnostics go to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0  # MI355X published HBM3E bandwidth
METRIC = "matrices/s & %HBM peak (n=32 fp32, 1M, pivoted); p50/p99 ms mixed stream"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _repeat(run, reset, target_s: float):
    """Runs `run` (after `reset`) until about target_s seconds of it were
    timed; returns (repetitions, seconds timed)."""
    reset()
    t0 = time.perf_counter()
    run()
    total = time.perf_counter() - t0
    reps = max(1, int(target_s / max(total, 1e-6)))
    for _ in range(reps - 1):
        reset()
        t0 = time.perf_counter()
        run()
        total += time.perf_counter() - t0
    return reps, total


def cpu_baseline(n: int, target_s: float = 10.0, gepp: bool = False) -> dict:
    """The reference's host path, timed on this host's cores on bounded samples
    of the same workload:

    * inversion: the oracle's FP32 restatement of the reference algorithm
      (first-max pivots on the un-updated column, Doolittle LU, column
      solves) -- or, for the partial-pivoting mode, its FP32 Gauss-Jordan with
      getrf's rule -- with OpenMP over matrices on all cores (`value`) and on
      one core (`value_1core`);
    * the check: the oracle's restatement of verifyInv (parallel_pivot/
      verify.hpp:50-103) single-threaded, as the reference builds it without
      -fopenmp (README.md:51), and on all cores.
    """
    import oracle

    def invert(work, piv, threads):
        if gepp:
            oracle.batch_gepp_inplace(work, piv, threads)
        else:
            oracle.batch_inverse_inplace(work, piv, oracle.PIV_FIRSTMAX, threads)

    threads = oracle.openmp_threads()
    sample, sample1 = 65536, 8192
    pristine = oracle.uniform_matrices(42, sample, n)
    piv = np.empty((sample, n), dtype=np.int32)
    work = pristine.copy()
    reps, total = _repeat(lambda: invert(work, piv, threads), lambda: np.copyto(work, pristine), target_s)
    w1 = pristine[:sample1].copy()
    reps1, total1 = _repeat(lambda: invert(w1, piv[:sample1], 1), lambda: np.copyto(w1, pristine[:sample1]),
                            target_s / 2)
    x = pristine[:sample1].copy()
    invert(x, piv[:sample1], threads)
    a1 = pristine[:sample1]
    vreps1, vtotal1 = _repeat(lambda: oracle.verify_inv(a1, x, 1e-3, threads=1), lambda: None, target_s / 3)
    vreps, vtotal = _repeat(lambda: oracle.verify_inv(a1, x, 1e-3, threads=0), lambda: None, target_s / 5)
    rule = ("FP32 Gauss-Jordan with getrf's partial pivoting" if gepp else
            "FP32 Doolittle LU + column solves with the reference's first-max pivot rule")
    return {"value": reps * sample / total, "unit": "matrices/s", "cores": threads, "kind": "port",
            "sample": f"inversion: first {sample} matrices of the workload (n={n}, mt19937(42) U[-1,1)) "
                      f"inverted {reps} times, {rule} (oracle/lu_oracle.c), OpenMP over matrices on {threads} "
                      f"threads, {total:.1f} s; 1 core: first {sample1}, {reps1} times, {total1:.1f} s; "
                      f"verifyInv restatement on {sample1} inverses: 1 thread {vreps1} times ({vtotal1:.1f} s), "
                      f"{threads} threads {vreps} times ({vtotal:.1f} s)",
            "value_1core": reps1 * sample1 / total1,
            "verify_inv_1thread_matrices_per_s": vreps1 * sample1 / vtotal1,
            "verify_inv_all_cores_matrices_per_s": vreps * sample1 / vtotal,
            "cpu_model": cpu_model(), "host_logical_cpus": os.cpu_count()}


def cpu_baseline_ragged(ns_all: np.ndarray, flat: np.ndarray, offsets: np.ndarray, target_s: float = 12.0) -> dict:
    """The oracle over a bounded sample of the ragged batch (its first 65,536
    matrices, grouped by n), OpenMP over matrices, repeated to ~target_s."""
    import oracle

    threads = oracle.openmp_threads()
    sample = min(65536, ns_all.size)
    groups = []
    for n in np.unique(ns_all[:sample]):
        sel = np.nonzero(ns_all[:sample] == n)[0]
        a = np.stack([flat[offsets[i]:offsets[i] + n * n].reshape(n, n) for i in sel])
        groups.append((int(n), a, np.empty((sel.size, n), np.int32)))

    def once():
        t = 0.0
        for n, a, p in groups:
            w = a.copy()
            t0 = time.perf_counter()
            oracle.batch_inverse_inplace(w, p, oracle.PIV_FIRSTMAX, threads)
            t += time.perf_counter() - t0
        return t

    first = once()
    reps = max(1, int(target_s / max(first, 1e-6)))
    total = first + sum(once() for _ in range(reps - 1))
    return {"value": reps * sample / total, "unit": "matrices/s", "cores": threads, "kind": "port",
            "sample": f"first {sample} matrices of the ragged batch (n~U{{2..32}}), grouped by n, inverted {reps} "
                      f"times by the oracle's FP32 restatement (first-max pivots), OpenMP over matrices, "
                      f"{total:.1f} s"}


def ragged_measure(batch: int, pivot: str, steps: int, warmup: int, n_lo: int = 2, n_hi: int = 32,
                   graph: bool = True, order: str = "random") -> dict:
    """Mixed-n batch (`batch` matrices, n ~ U{n_lo..n_hi}, packed with int64
    offsets): the ragged entry (plan + inversion kernel) eager and replayed
    from a hipGraph, against one plain launch per n (a) over the same matrices
    pre-grouped into per-n contiguous buffers (the plain path's best case: no
    gather cost counted) and (b) on the packed batch itself (gather each class,
    invert, scatter back).  graph=False launches everything directly (no
    hipGraph), for profilers that do not follow graph launches; the step time
    is then the eager one.  order="sorted" packs the same matrices grouped by
    n (largest first) instead of in generation order.  Returns the
    measurements; `ok` is False when any result differs bitwise."""
    import matrixinversion_amd as mia
    from matrixinversion_amd import workloads

    flags = mia.PIVOT_NAMES[pivot]
    t0 = time.perf_counter()
    flat, offsets, ns, poffs = workloads.ragged_batch(42, batch, n_lo, n_hi)
    if order == "sorted":  # the same matrices, packed class by class
        perm = np.argsort(-ns, kind="stable")
        flat = np.concatenate([flat[offsets[i]:offsets[i] + int(ns[i]) ** 2] for i in perm])
        ns = ns[perm]
        sq = ns.astype(np.int64) ** 2
        offsets = np.zeros(batch, np.int64)
        np.cumsum(sq[:-1], out=offsets[1:])
        poffs = np.zeros(batch, np.int64)
        np.cumsum(ns[:-1].astype(np.int64), out=poffs[1:])
    log(f"[ragged] generated {batch} matrices ({flat.nbytes / 1e6:.0f} MB) in {time.perf_counter() - t0:.1f}s; "
        f"E[n]={ns.mean():.2f} E[n^2]={(ns.astype(np.int64) ** 2).mean():.1f}")
    total_piv = int(ns.sum())
    stream = mia.Stream()
    pristine = mia.DeviceBuffer.from_array(flat)
    O = mia.DeviceBuffer.from_array(offsets)
    NS = mia.DeviceBuffer.from_array(ns)
    PO = mia.DeviceBuffer.from_array(poffs)
    P = mia.DeviceBuffer(total_piv * 4)
    bufs = [mia.DeviceBuffer(flat.nbytes) for _ in range(steps)]
    plan = mia.RaggedPlan(batch)

    nfloats = flat.size

    def restore():
        for b in bufs:
            mia.device.copy_kernel(b, pristine, nfloats, stream)
        stream.synchronize()

    # A copy of the batch of its own for the untimed load run between each
    # restore and the timed region that follows it: launch times settle only
    # under sustained load, and a restore's long copy unsettles them again.
    wbuf = mia.DeviceBuffer(flat.nbytes)
    mia.device.copy_kernel(wbuf, pristine, nfloats, stream)

    def settle():
        for _ in range(max(1, warmup)):
            mia.lu_batched_inverse_ragged(wbuf, O, NS, batch, flags, P, PO, plan=plan, stream=stream)
        stream.synchronize()

    def timed(run_one):
        """K steps back to back between one event pair (events between graph
        launches would add their own cost): the mean step time."""
        e0, e1 = mia.Event(), mia.Event()
        stream.synchronize()
        e0.record(stream)
        for i in range(steps):
            run_one(i)
        e1.record(stream)
        stream.synchronize()
        return e0.elapsed_ms(e1) / steps

    def ragged(i):
        mia.lu_batched_inverse_ragged(bufs[i], O, NS, batch, flags, P, PO, plan=plan, stream=stream)

    def graphs_of(fn):
        out = []
        for i in range(steps):
            g = mia.Graph(stream)
            g.begin()
            fn(i)
            g.end()
            out.append(g)
        for _ in range(max(1, warmup)):  # every graph's first launch is slower: keep it out of the timing
            for g in out:
                g.launch()
        stream.synchronize()
        return out

    # eager ragged
    for i in range(warmup):
        restore()
        ragged(i % steps)
        stream.synchronize()
    restore()
    settle()
    t_eager = timed(ragged)
    x_ragged = bufs[-1].download(flat.shape, np.float32)
    p_ragged = P.download((total_piv,), np.int32)
    bad = int(plan.invalid_count())

    # graph-replayed ragged: one graph per resident copy
    graphs, t_graph, graph_same = [], None, True
    if graph:
        graphs = graphs_of(ragged)
        restore()
        settle()
        t_graph = timed(lambda i: graphs[i].launch())
        x_graph = bufs[-1].download(flat.shape, np.float32)
        graph_same = bool(np.array_equal(x_graph.view(np.uint32), x_ragged.view(np.uint32)))

    # baseline: one plain launch per n on per-n contiguous copies of the same matrices
    classes = []
    for n in range(1, 33):
        sel = np.nonzero(ns == n)[0]
        if sel.size == 0:
            continue
        idx = (offsets[sel][:, None] + np.arange(n * n)[None, :]).ravel()
        a = flat[idx]
        src = mia.DeviceBuffer.from_array(a)
        work = [mia.DeviceBuffer(a.nbytes) for _ in range(steps)]
        classes.append((n, sel, src, work, mia.DeviceBuffer(sel.size * n * 4)))
    del flat

    def restore_plain():
        for n, sel, src, work, pv in classes:
            for w in work:
                mia.device.copy_kernel(w, src, sel.size * n * n, stream)
        stream.synchronize()

    def plain(i):
        for n, sel, src, work, pv in classes:
            mia.lu_batched_inverse(work[i], n, sel.size, flags, pv, stream)

    # the same per-n launches on the packed batch itself: gather each class
    # to a contiguous buffer, invert, scatter back (what a caller of the one-n
    # entry point does with packed mixed-size data)
    from matrixinversion_amd._native import check, lib as _lib
    class_offs = [mia.DeviceBuffer.from_array(offsets[sel]) for n, sel, *_ in classes]

    def plain_packed(i):
        for (n, sel, src, work, pv), co in zip(classes, class_offs):
            check(_lib().lubx_gather_f32(work[i].ptr, bufs[i].ptr, co.ptr, n * n, sel.size, stream.handle), "gather")
            mia.lu_batched_inverse(work[i], n, sel.size, flags, pv, stream)
            check(_lib().lubx_scatter_f32(bufs[i].ptr, work[i].ptr, co.ptr, n * n, sel.size, stream.handle),
                  "scatter")

    for i in range(warmup):
        plain(i % steps)
    stream.synchronize()
    restore_plain()
    settle()
    t_plain = timed(plain)
    plain_graphs, packed_graphs, t_plain_graph = [], [], None
    if graph:
        plain_graphs = graphs_of(plain)
        restore_plain()
        settle()
        t_plain_graph = timed(lambda i: plain_graphs[i].launch())
        packed_graphs = graphs_of(plain_packed)
        restore()
        settle()
        t_packed_graph = timed(lambda i: packed_graphs[i].launch())
    else:
        restore()
        settle()
        t_packed_graph = timed(plain_packed)
    x_packed = bufs[-1].download(x_ragged.shape, np.float32)
    packed_same = bool(np.array_equal(x_packed.view(np.uint32), x_ragged.view(np.uint32)))

    # the ragged result must equal the per-n plain results bit for bit
    same = True
    for n, sel, src, work, pv in classes:
        xp = work[-1].download((sel.size, n * n), np.float32)
        idx = (offsets[sel][:, None] + np.arange(n * n)[None, :])
        same &= bool(np.array_equal(x_ragged[idx].view(np.uint32), xp.view(np.uint32)))
        pp = pv.download((sel.size, n), np.int32)
        same &= bool(np.array_equal(p_ragged[(poffs[sel][:, None] + np.arange(n)[None, :])], pp))
    def ms(t):
        return "-" if t is None else f"{t:.3f}"

    log(f"[ragged] eager {ms(t_eager)} ms, graph {ms(t_graph)} ms, per-n plain {ms(t_plain)} ms "
        f"(graph {ms(t_plain_graph)} ms; on the packed batch with gather/scatter {ms(t_packed_graph)} ms); "
        f"ragged == plain bitwise: {same}; graph == eager: {graph_same}; invalid n: {bad}")

    for g in graphs + plain_graphs + packed_graphs:
        g.close()
    plan.close()
    for b in bufs + [pristine, O, NS, PO, P, wbuf] + class_offs:
        b.close()
    for n, sel, src, work, pv in classes:
        for b in [src, pv] + work:
            b.close()
    stream.close()
    algo = workloads.ragged_bytes(ns, pivots=flags != 0)
    t_step = t_graph if graph else t_eager
    return {"batch": batch, "timing": "graph" if graph else "eager", "step_ms_mean": t_step,
            "graph_ms_mean": t_graph, "eager_ms_mean": t_eager,
            "plain_per_n_ms_mean": t_plain, "plain_per_n_graph_ms_mean": t_plain_graph,
            "speedup_vs_plain_per_n": t_plain / t_step,
            "speedup_vs_plain_per_n_graph": None if t_plain_graph is None else t_plain_graph / t_step,
            "plain_per_n_packed_graph_ms_mean": t_packed_graph,
            "speedup_vs_plain_per_n_on_packed_data": t_packed_graph / t_step,
            "plain_on_packed_equals_ragged_bitwise": packed_same,
            "ragged_equals_plain_bitwise": same, "graph_equals_eager_bitwise": graph_same,
            "invalid_n": bad, "algorithmic_bytes_per_launch": algo,
            "pct_hbm": 100 * algo / (t_step * 1e-3) / 1e9 / HBM_PEAK_GBPS,
            "plain_launches_per_step": len(classes),
            "ok": same and graph_same and packed_same and bad == 0,
            "baselines": "plain_per_n: one lu_batched_inverse per n on matrices pre-grouped by n outside the timed "
                         "region; plain_per_n_packed: per n, gather the class from the packed batch, invert, "
                         "scatter back (graph-captured unless timing is eager)"}


def bench_ragged(args) -> int:
    """The ragged workload as its own bench line (value = matrices/s of the
    graph-replayed ragged launch; of the direct launch under --no-graph)."""
    import matrixinversion_amd as mia

    batch = args.batch or 500_000
    info = mia.device_info(0)
    log(f"[ragged] device {info.arch} CUs={info.cu_count}")
    d = ragged_measure(batch, args.pivot, args.steps, args.warmup, args.n_lo, args.n_hi, graph=not args.no_graph,
                       order=args.ragged_order)
    ms = d["step_ms_mean"]
    algo = d["algorithmic_bytes_per_launch"]
    achieved = algo / (ms * 1e-3) / 1e9
    traffic = (load_traffic(f"traffic_lu_ragged_n{args.n_lo}-{args.n_hi}_{args.pivot}_{batch}.json")
               if args.ragged_order == "random" else {"bytes": None, "build_id": None, "stale": None})
    out = {
        "metric": "matrices/s (ragged mixed-n batch, pivoted)",
        "value": batch / (ms * 1e-3),
        "unit": "matrices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": f"synthetic: sizes mt19937(43) n={args.n_lo}+x%{args.n_hi - args.n_lo + 1}, entries mt19937(42) "
                f"U[-1,1), packed {'in matrix order' if args.ragged_order == 'random' else 'grouped by n'}",
        "config": {"workload": f"ragged in-place LU inversion, {batch} matrices, n~U{{{args.n_lo}..{args.n_hi}}}, "
                               f"pivot={args.pivot}, "
                               f"int64 offsets, plan + inversion launch "
                               f"({'launched directly' if args.no_graph else 'one hipGraph replay'})",
                   "batch": batch, "pivot": args.pivot, "global_batch": batch, "order": args.ragged_order},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": traffic["bytes"],
                     "traffic_build_id": traffic["build_id"], "traffic_stale": traffic["stale"]},
        "detail": d,
    }
    if not args.no_cpu_baseline:
        from matrixinversion_amd import workloads
        log("[ragged] timing the CPU baseline ...")
        flat2, offsets2, ns2, _ = workloads.ragged_batch(42, min(batch, 65536), args.n_lo, args.n_hi)
        out["cpu_baseline"] = cpu_baseline_ragged(ns2, flat2, offsets2)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    return 0 if d["ok"] else 3


def load_traffic(name: str) -> dict:
    """HBM bytes per launch from a committed counter profile (profiles/<name>),
    kept only when it was measured on the sources of the loaded library
    (its build_id equals lu_build_id()); otherwise reported as stale."""
    import matrixinversion_amd as mia

    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return {"bytes": None, "build_id": None, "stale": None}
    with open(path) as f:
        t = json.load(f)
    stale = t.get("build_id") != mia.build_id()
    return {"bytes": None if stale else t.get("hbm_bytes_per_launch"), "build_id": t.get("build_id"), "stale": stale}


def serve_sweep(rates, duration: float, budget_gb: float = 64.0, pool: int = 200_000, driver: str = "native") -> list:
    """Poisson serving runs (matrixinversion_amd.serve) at each rate, driven
    by the library's host loop (native) or by the Python loop."""
    from matrixinversion_amd import serve

    pools = (serve.Pool(42, pool), serve.Pool(43, pool, dominant=True))
    run = serve.poisson_stream_native if driver == "native" else serve.poisson_stream
    out = []
    with serve.Server(int(budget_gb * 2**30)) as srv:
        # untimed: the first kernel launches load code objects, and the stream
        # settles only after some time at the timed rate
        serve.warm(srv, pools, rate=float(rates[0]), duration=0.5, driver=driver)
        for r in rates:
            res = run(srv, pools, float(r), duration)
            log(f"[serve] {r:>8.0f} req/s: p50 {res['p50_ms']:.3f} ms, p99 {res['p99_ms']:.3f} ms, "
                f"{res['matrices_per_s'] / 1e6:.2f} M matrices/s, {res['matrices_per_batch']:.0f} matrices/batch")
            out.append(res)
    for p in pools:
        p.close()
    return out


def bench_serve(args) -> int:
    """Mixed serving stream: Poisson arrivals, 1..4096 matrices per request
    (log-uniform), n ~ U{4..32}, 70 % pivoted U[-1,1) / 30 % unpivoted
    diagonally dominant, 64 GB HBM admission budget.  One JSON line; value =
    p99 latency at the reference rate (the last of --rates)."""
    rates = [float(x) for x in args.rates.split(",")]
    sweep = serve_sweep(rates, args.duration, driver=args.serve_driver)
    top = sweep[-1]
    out = {
        "metric": "p50/p99 ms, mixed serving stream",
        "value": top["p99_ms"],
        "unit": "ms (p99)",
        "n_gpus": 1,
        "steps": len(sweep),
        "warmup": 0,
        "ms_per_step": top["mean_ms"],
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: request pools from mt19937(42/43), Poisson arrivals from RandomState(7)",
        "config": {"workload": "serving: Poisson arrivals, log-uniform 1..4096 matrices per request, n~U{4..32}, "
                               "70% pivoted / 30% unpivoted dominant, 64 GB admission budget",
                   "rates_req_s": rates, "duration_s": args.duration, "driver": args.serve_driver},
        "detail": {"p50_ms": top["p50_ms"], "p99_ms": top["p99_ms"], "matrices_per_s": top["matrices_per_s"],
                   "sweep": sweep},
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    return 0


def bench_refine(args) -> int:
    """Precision workload: 100,000 matrices of n=32, A = Q1 diag(sigma) Q2^T,
    sigma log-spaced 1..10^k for k in {2, 4, 6}; plain FP32 LU inverse vs one
    refinement step (FP64 residual, FP32 MFMA correction), relative error
    against the FP64 inverse of the same FP32 A.  Timed: the plain inverse
    (in place), the fused refined inverse (lu_batched_inverse_refined: A read
    once, X written once) and the separate refinement pass over A and X.
    One JSON line; value = the median relative error after refinement at
    k = 4."""
    import matrixinversion_amd as mia
    from matrixinversion_amd import workloads
    from matrixinversion_amd._native import check, lib

    batch, n = args.batch or 100_000, 32
    stream = mia.Stream()
    steps = max(2, args.steps)
    per_k = {}

    def per_launch(fn, before=None):
        """Mean ms of `steps` launches of fn, each bracketed by an event pair
        on the stream (`before`, untimed, runs ahead of each)."""
        ev = [mia.Event() for _ in range(2 * steps)]
        for _ in range(max(1, args.warmup)):
            if before:
                before()
            fn()
        for i in range(steps):
            if before:
                before()
            ev[2 * i].record(stream)
            fn()
            ev[2 * i + 1].record(stream)
        stream.synchronize()
        return float(np.mean([ev[2 * i].elapsed_ms(ev[2 * i + 1]) for i in range(steps)]))

    for k in (2, 4, 6):
        t0 = time.perf_counter()
        a = workloads.conditioned_matrices(100 + k, batch, n, k)
        log(f"[refine] k={k}: generated {batch} matrices in {time.perf_counter() - t0:.1f}s")
        A = mia.DeviceBuffer.from_array(a)
        X = mia.DeviceBuffer(a.nbytes)
        X2 = mia.DeviceBuffer(a.nbytes)
        P = mia.DeviceBuffer(batch * n * 4)

        def restore():
            mia.device.copy_kernel(X, A, batch * n * n, stream)

        t_inv = per_launch(lambda: mia.lu_batched_inverse(X, n, batch, mia.PIVOT_PARALLEL, P, stream), restore)
        stream.synchronize()
        x0 = X.download(a.shape, np.float32)
        t_ref = per_launch(lambda: mia.refine(A, X, n, batch, stream))   # X: the plain inverse, refined again each call
        restore()
        mia.lu_batched_inverse(X, n, batch, mia.PIVOT_PARALLEL, P, stream)
        mia.refine(A, X, n, batch, stream)
        stream.synchronize()
        x1 = X.download(a.shape, np.float32)
        t_fused = per_launch(lambda: check(lib().lu_batched_inverse_refined(
            A.ptr, X2.ptr, P.ptr, n, batch, mia.PIVOT_PARALLEL, stream.handle), "inverse_refined"))
        x2 = X2.download(a.shape, np.float32)
        for b in (A, X, X2, P):
            b.close()
        inv = np.linalg.inv(a.astype(np.float64))
        nrm = np.linalg.norm(inv, axis=(1, 2))
        e0 = np.linalg.norm(x0 - inv, axis=(1, 2)) / nrm
        e1 = np.linalg.norm(x1 - inv, axis=(1, 2)) / nrm
        e2 = np.linalg.norm(x2 - inv, axis=(1, 2)) / nrm
        per_k[k] = {"median_rel_err_plain": float(np.median(e0)), "median_rel_err_refined": float(np.median(e1)),
                    "median_rel_err_fused": float(np.median(e2)),
                    "p99_rel_err_plain": float(np.percentile(e0, 99)),
                    "p99_rel_err_refined": float(np.percentile(e1, 99)),
                    "p99_rel_err_fused": float(np.percentile(e2, 99)),
                    "inverse_ms": t_inv, "refine_pass_ms": t_ref, "fused_refined_inverse_ms": t_fused,
                    "fused_over_plain": t_fused / t_inv,
                    "refine_pass_GBps": batch * 12 * n * n / (t_ref * 1e-3) / 1e9,
                    "fused_GBps": batch * (8 * n * n + 4 * n) / (t_fused * 1e-3) / 1e9}
        log(f"[refine] k={k}: median rel err {np.median(e0):.2e} -> {np.median(e1):.2e} (fused {np.median(e2):.2e}), "
            f"p99 {np.percentile(e0, 99):.2e} -> {np.percentile(e1, 99):.2e}; inverse {t_inv:.3f} ms, "
            f"fused inverse+refine {t_fused:.3f} ms ({t_fused / t_inv:.2f}x), refinement pass {t_ref:.3f} ms")
    fb = per_k[4]["fused_GBps"]
    out = {
        "metric": "relative error vs FP64 inverse after one refinement step (median, k=4)",
        "value": per_k[4]["median_rel_err_fused"],
        "unit": "relative Frobenius error",
        "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": per_k[4]["fused_refined_inverse_ms"],
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32 (FP64 residual)",
        "data": "synthetic: A = Q1 diag(logspace(0,k,32)) Q2^T, Q from QR of RandomState(100+k) Gaussians",
        "config": {"workload": f"refinement, {batch} x n=32, k in (2, 4, 6), pivot=parallel", "batch": batch},
        "detail": per_k,
        # the fused kernel: 8 n^2 + 4 n algorithmic bytes per matrix (A read, X and pivots written)
        "roofline": {"bound": "hbm", "achieved": fb, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": fb / HBM_PEAK_GBPS, "traffic": None},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    return 0


def bench_rocsolver(args) -> int:
    """This library against rocSOLVER's strided-batched getrf + getri (LAPACK
    partial pivoting) on the same device: 1M matrices per n, U[-1,1) from
    mt19937(42), both in place; kernel time from HIP events on the launch
    stream, inputs restored (untimed) before every launch.  value = speedup
    at n = 32."""
    import ctypes

    import matrixinversion_amd as mia
    from matrixinversion_amd import workloads
    from matrixinversion_amd._native import check, cmp_lib

    batch = args.batch or 1_000_000
    stream = mia.Stream()
    lib = cmp_lib()
    h = ctypes.c_void_p()
    check(lib.lubref_rocsolver_create(ctypes.byref(h), stream.handle), "rocsolver create")
    per_n = {}
    for n in (8, 16, 32):
        a = workloads.uniform_matrices(42, batch, n)
        src = mia.DeviceBuffer.from_array(a)
        work = mia.DeviceBuffer(a.nbytes)
        P = mia.DeviceBuffer(batch * n * 4)
        I = mia.DeviceBuffer(batch * 4)
        e0, e1 = mia.Event(), mia.Event()

        def timed(fn):
            out = []
            for i in range(args.warmup + args.steps):
                mia.device.copy_kernel(work, src, batch * n * n, stream)
                e0.record(stream)
                fn()
                e1.record(stream)
                stream.synchronize()
                if i >= args.warmup:
                    out.append(e0.elapsed_ms(e1))
            return float(np.mean(out))

        t_ours = timed(lambda: mia.lu_batched_inverse(work, n, batch, mia.PIVOT_PARALLEL, P, stream))
        t_roc = timed(lambda: check(lib.lubref_rocsolver_inverse(h, work.ptr, P.ptr, I.ptr, n, batch), "rocsolver"))
        info = I.download((batch,), np.int32)
        per_n[n] = {"ours_ms": t_ours, "rocsolver_ms": t_roc, "speedup": t_roc / t_ours,
                    "ours_Mmat_s": batch / t_ours / 1e3, "rocsolver_Mmat_s": batch / t_roc / 1e3,
                    "rocsolver_info_nonzero": int((info != 0).sum())}
        log(f"[rocsolver] n={n}: ours {t_ours:.3f} ms, rocSOLVER getrf+getri {t_roc:.3f} ms "
            f"({t_roc / t_ours:.1f}x)")
        for b in (src, work, P, I):
            b.close()
    lib.lubref_rocsolver_destroy(h)
    out = {
        "metric": "speedup over rocSOLVER strided-batched getrf+getri (n=32, 1M, fp32)",
        "value": per_n[32]["speedup"], "unit": "x", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": per_n[32]["ours_ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic: mt19937(42) U[-1,1) entries",
        "config": {"workload": f"in-place batched inverse, {batch} matrices per n, pivoted", "batch": batch},
        "detail": per_n, "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    return 0


def bench_ragged_classes(args) -> int:
    """Single-class batches through the ragged entry against the plain launch
    of the same matrices: per n, `batch` (default 200k) matrices packed
    contiguously, both entry points timed with HIP events over K back-to-back
    launches after W warmup launches.  Shows which size classes the ragged
    kernel loses on (its planning kernels are included)."""
    import matrixinversion_amd as mia

    batch = args.batch or 200_000
    flags = mia.PIVOT_NAMES[args.pivot]
    steps, warmup = args.steps, max(1, args.warmup)
    stream = mia.Stream()
    rng = np.random.RandomState(1)

    def timeit(fn):
        for _ in range(warmup):
            fn()
        stream.synchronize()
        e0, e1 = mia.Event(), mia.Event()
        e0.record(stream)
        for _ in range(steps):
            fn()
        e1.record(stream)
        stream.synchronize()
        return e0.elapsed_ms(e1) / steps

    rows = {}
    for n in (32, 31, 28, 24, 20, 17, 16, 13, 12, 9, 8, 5, 3):
        a = rng.uniform(-1, 1, size=batch * n * n).astype(np.float32)
        A = mia.DeviceBuffer.from_array(a)
        O = mia.DeviceBuffer.from_array(np.arange(batch, dtype=np.int64) * n * n)
        NS = mia.DeviceBuffer.from_array(np.full(batch, n, np.int32))
        PO = mia.DeviceBuffer.from_array(np.arange(batch, dtype=np.int64) * n)
        P = mia.DeviceBuffer(batch * n * 4)
        with mia.RaggedPlan(batch) as plan:
            t_plain = timeit(lambda: mia.lu_batched_inverse(A, n, batch, flags, P, stream))
            t_rag = timeit(lambda: mia.lu_batched_inverse_ragged(A, O, NS, batch, flags, P, PO, plan=plan,
                                                                  stream=stream))
        for b in (A, O, NS, PO, P):
            b.close()
        rows[n] = {"plain_us": t_plain * 1e3, "ragged_us": t_rag * 1e3, "ratio": t_rag / t_plain}
        log(f"[classes] n={n:2d} plain {t_plain * 1e3:7.1f} us ragged {t_rag * 1e3:7.1f} us ratio {t_rag / t_plain:.2f}")
    out = {
        "metric": "ragged / plain launch time, single-class batches", "value": rows[32]["ratio"],
        "unit": "x (at n=32)", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": rows[32]["ragged_us"] / 1e3, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic: RandomState(1) U[-1,1), one n per batch",
        "config": {"workload": f"single-class ragged vs plain, {batch} matrices per n, pivot={args.pivot}",
                   "batch": batch, "pivot": args.pivot, "global_batch": batch},
        "detail": rows, "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=32)
    ap.add_argument("--batch", type=int, default=None)
    ap.add_argument("--pivot", default="parallel", choices=["none", "parallel", "serial", "tree", "gepp"],
                    help="pivot rule; gepp is LAPACK's partial pivoting (not a reference rule)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--n-lo", type=int, default=2, help="ragged: smallest n")
    ap.add_argument("--n-hi", type=int, default=32, help="ragged: largest n")
    ap.add_argument("--ragged-order", choices=["random", "sorted"], default="random",
                    help="ragged: matrices packed in generation order (random n) or grouped by n")
    ap.add_argument("--workload", default="headline", choices=["headline", "ragged", "ragged_classes", "serve", "refine", "rocsolver"],
                    help="headline: the metric's configuration; ragged: mixed-n batch (--batch defaults to 500k); "
                         "serve: Poisson serving sweep; refine: precision workload (100k n=32); "
                         "rocsolver: against rocSOLVER getrf+getri")
    ap.add_argument("--rates", default="1000,4000,16000", help="serve: request rates (req/s)")
    ap.add_argument("--duration", type=float, default=2.0, help="serve: seconds per rate")
    ap.add_argument("--no-serve", action="store_true", help="headline: skip the short serving run")
    ap.add_argument("--serve-driver", choices=["native", "python"], default="native",
                    help="serving runs: the library's host loop over lu_server_* (native) or the Python loop")
    ap.add_argument("--no-graph", action="store_true",
                    help="ragged: launch directly instead of replaying hipGraphs (profiled runs: rocprofv3's "
                         "kernel trace stops the process inside hipGraphLaunch on this image)")
    ap.add_argument("--no-extras", action="store_true",
                    help="headline: skip the dominant-input gate and the mixed-n leg (profiling runs)")
    ap.add_argument("--traffic-from", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"), default=None,
                    help="write the traffic JSON of --traffic-kernel from two rocprofv3 --pmc output folders "
                         "(FETCH_SIZE, WRITE_SIZE) of a bench run, calibrated by its copy kernel, and exit")
    ap.add_argument("--traffic-kernel", default="lu_inverse_kernel", help="kernel name prefix for --traffic-from")
    ap.add_argument("--traffic-out", default=None, help="--traffic-from: output file under profiles/")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU: the batch inversion has no cross-GPU exchange and this "
              f"build runs on a single MI355X (asked for --gpus {args.gpus}).", file=sys.stderr)
        return 2

    if args.traffic_from:
        return traffic_from(args)
    if args.workload == "ragged":
        return bench_ragged(args)
    if args.workload == "ragged_classes":
        return bench_ragged_classes(args)
    if args.workload == "serve":
        return bench_serve(args)
    if args.workload == "refine":
        return bench_refine(args)
    if args.workload == "rocsolver":
        return bench_rocsolver(args)

    import matrixinversion_amd as mia
    from matrixinversion_amd import workloads

    n, batch = args.n, args.batch or 1_000_000
    flags = mia.PIVOT_NAMES[args.pivot]
    info = mia.device_info(0)
    log(f"[bench] device {info.arch} CUs={info.cu_count} HBM={info.total_mem / 2**30:.0f} GiB")

    serving = None
    if not args.no_serve:
        # the metric's second half, measured first: a serving process does not
        # follow a multi-GB inversion burst, and the server frees its arena
        # before the headline's buffers are allocated
        top = serve_sweep([16000.0], 2.0, driver=args.serve_driver)[0]
        serving = {"rate_req_s": top["rate_req_s"], "p50_ms": top["p50_ms"], "p99_ms": top["p99_ms"],
                   "matrices_per_s": top["matrices_per_s"],
                   "workload": "Poisson arrivals, log-uniform 1..4096 matrices/request, n~U{4..32}, "
                               "70% pivoted / 30% unpivoted dominant, 64 GB admission budget",
                   "driver": top["driver"], "busy_retries": top["busy_retries"], "measured": "before the headline"}

    t0 = time.perf_counter()
    a = workloads.uniform_matrices(42, batch, n)
    log(f"[bench] generated {batch} x {n}x{n} inputs in {time.perf_counter() - t0:.1f}s")
    nbytes = a.nbytes
    pristine = mia.DeviceBuffer.from_array(a)
    del a
    stream = mia.Stream()
    steps, warmup = args.steps, args.warmup
    bufs = [mia.DeviceBuffer(nbytes) for _ in range(steps)]
    piv = mia.DeviceBuffer(batch * n * 4)

    def restore(bs):
        for b in bs:  # plain dwordx4 copy kernel: also the counter calibration of known size
            mia.device.copy_kernel(b, pristine, batch * n * n, stream)
        stream.synchronize()

    # Restore the timed buffers first, then warm up on a copy of their own:
    # back-to-back launches, as in the timed region (the device needs a few
    # tens of milliseconds of sustained load before its launch times settle,
    # and a long copy just before the timed region unsettles them again).
    restore(bufs)
    if warmup > 0:
        warm = mia.DeviceBuffer(nbytes)
        restore([warm])
        for i in range(warmup):
            mia.lu_batched_inverse(warm, n, batch, flags, piv, stream)
        stream.synchronize()
        warm.close()

    ev = [mia.Event() for _ in range(steps + 1)]
    stream.synchronize()
    t_start = time.perf_counter()
    ev[0].record(stream)
    for i in range(steps):
        mia.lu_batched_inverse(bufs[i], n, batch, flags, piv, stream)
        ev[i + 1].record(stream)
    stream.synchronize()
    wall_ms = (time.perf_counter() - t_start) * 1e3
    per_launch = np.array([ev[i].elapsed_ms(ev[i + 1]) for i in range(steps)])
    total_ms = ev[0].elapsed_ms(ev[steps])

    # check of the last step on the device: verifyInv over the whole batch
    correct, worst = mia.residual(pristine, bufs[-1], n, batch, 1e-3, stream)
    p = piv.download((min(batch, 1024), n), np.int32)
    if not (np.sort(p, axis=1) == np.arange(n)).all():
        log("[bench] pivot vectors are not permutations")
        return 3
    log(f"[bench] check: verifyInv correct {correct}/{batch} at 1e-3, max|AX-I|={worst:.2e}")
    for b in bufs:
        b.close()
    pristine.close()

    ms_per_step = total_ms / steps
    mats_per_s = batch / (ms_per_step * 1e-3)
    algo_bytes = batch * (8 * n * n + (4 * n if flags != 0 else 0))
    achieved = algo_bytes / (per_launch.mean() * 1e-3) / 1e9

    traffic = load_traffic(f"traffic_lu_inverse_n{n}_{args.pivot}_{batch}.json")

    out = {
        "metric": METRIC,
        "value": mats_per_s,
        "unit": "matrices/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: mt19937(42) U[-1,1) entries, pinned float mapping",
        "config": {"workload": f"batched in-place LU inversion, {batch} x n={n}, pivot={args.pivot}",
                   "n": n, "batch": batch, "pivot": args.pivot, "layout": "row-major contiguous, in place",
                   "global_batch": batch},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": traffic["bytes"],
                     "traffic_build_id": traffic["build_id"], "traffic_stale": traffic["stale"],
                     "build_id": mia.build_id()},
        "detail": {"kernel_ms_mean": float(per_launch.mean()), "kernel_ms_min": float(per_launch.min()),
                   "kernel_ms_max": float(per_launch.max()), "timed_wall_ms": wall_ms,
                   "algorithmic_bytes_per_launch": algo_bytes, "pct_hbm": 100 * achieved / HBM_PEAK_GBPS,
                   "verifyInv_correct": correct, "max_abs_AXmI": worst,
                   "accuracy_note": "U[-1,1) inputs are not uniformly well-conditioned; the 1e-4 accuracy gate "
                                    "is checked on diagonally dominant inputs (well_conditioned_1M)"},
    }
    piv.close()
    ok = True
    if not args.no_extras:
        # the accuracy target on well-conditioned inputs, at full size: A = U[-1,1) + n*I
        wc = dominant_gate(n, batch, flags, stream)
        out["detail"]["well_conditioned_1M"] = wc
        ok &= wc["gate_passed"]
        if n == 32 and args.pivot in ("parallel", "serial", "tree"):
            # the mixed-size target: the ragged path against per-n plain launches
            log("[bench] mixed-n leg (500k matrices, n~U{2..32}) ...")
            rg = ragged_measure(500_000, args.pivot, 5, 20)  # 20 settle launches, as --workload ragged
            out["detail"]["mixed_n_500k"] = rg
            ok &= rg["ok"]
    if serving is not None:
        out["serving"] = serving
    if not args.no_cpu_baseline:
        log("[bench] timing the CPU baseline ...")
        out["cpu_baseline"] = cpu_baseline(n, gepp=args.pivot == "gepp")
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    return 0 if ok else 3


def dominant_gate(n: int, batch: int, flags: int, stream) -> dict:
    """The accuracy target at full size: `batch` matrices A = U[-1,1) + n*I
    (mt19937(43)), inverted with `flags` on the device; max|A X - I| over the
    whole batch must be <= 1e-4; the reference's verifyInv count (1e-3) and
    the count within 1e-4 come from the device check."""
    import matrixinversion_amd as mia
    from matrixinversion_amd import workloads

    t0 = time.perf_counter()
    a = workloads.dominant_matrices(43, batch, n)
    A = mia.DeviceBuffer.from_array(a)
    X = mia.DeviceBuffer(a.nbytes)
    P = mia.DeviceBuffer(batch * n * 4)
    del a
    try:
        X.copy_from(A, stream)
        mia.lu_batched_inverse(X, n, batch, flags, P, stream)
        correct3, worst = mia.residual(A, X, n, batch, 1e-3, stream)
        correct4, _ = mia.residual(A, X, n, batch, 1e-4, stream)
    finally:
        for b in (A, X, P):
            b.close()
    passed = bool(worst <= 1e-4 and correct4 == batch)
    log(f"[bench] well-conditioned {batch} x n={n}: max|AX-I|={worst:.2e}, within 1e-4: {correct4}/{batch} "
        f"({'pass' if passed else 'FAIL'}) [{time.perf_counter() - t0:.1f}s]")
    return {"inputs": f"U[-1,1) + {n}*I, mt19937(43)", "max_abs_AXmI": worst, "verifyInv_correct_1e-3": correct3,
            "within_1e-4": correct4, "gate": "max|AX-I| <= 1e-4 for every matrix", "gate_passed": passed}


def traffic_from(args) -> int:
    """HBM bytes per launch of one kernel from two rocprofv3 counter runs of a
    bench.py command (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch), each
    scaled by what the counter reads for the copy kernel of the same run,
    whose bytes are known (it restores every timed input: n*n*batch floats in
    and out).  Writes profiles/<out> with the loaded library's build id."""
    import csv
    import glob

    import matrixinversion_amd as mia

    def per_dispatch(folder, counter):
        rows = {}
        for path in glob.glob(os.path.join(folder, "**", "*counter_collection.csv"), recursive=True):
            with open(path) as f:
                for r in csv.DictReader(f):
                    if r.get("Counter_Name") != counter:
                        continue
                    key = (r["Dispatch_Id"], r["Kernel_Name"], int(r.get("Grid_Size") or 0))
                    rows[key] = rows.get(key, 0.0) + float(r["Counter_Value"])
        return rows

    if args.workload == "ragged":
        from matrixinversion_amd import workloads
        batch = args.batch or 500_000
        nn = int((workloads.ragged_sizes(43, batch, args.n_lo, args.n_hi).astype(np.int64) ** 2).sum())
    else:
        batch = args.batch or 1_000_000
        nn = batch * args.n * args.n
    copy_bytes = (nn // 4) * 16  # the dwordx4 part of one full restore (lubx_copy_f32)
    out = {}
    for counter, folder in zip(("FETCH_SIZE", "WRITE_SIZE"), args.traffic_from):
        rows = per_dispatch(folder, counter)
        # the full-batch restores are the copy_kernel dispatches with the largest grid
        grid = max([g for (d, k, g) in rows if "copy_kernel" in k], default=0)
        copies = [v for (d, k, g), v in rows.items() if "copy_kernel" in k and g == grid]
        kern = [v for (d, k, g), v in rows.items() if args.traffic_kernel in k]
        if not copies or not kern:
            log(f"[traffic] {counter}: no copy_kernel or {args.traffic_kernel} dispatches under {folder}")
            return 4
        factor = copy_bytes / (float(np.median(copies)) * 1024)
        out[counter] = {"kernel_kib": float(np.mean(kern)), "copy_kib": float(np.median(copies)),
                        "factor": factor, "bytes": float(np.mean(kern)) * 1024 * factor,
                        "dispatches": len(kern)}
    total = out["FETCH_SIZE"]["bytes"] + out["WRITE_SIZE"]["bytes"]
    res = {"hbm_bytes_per_launch": total, "fetch_bytes": out["FETCH_SIZE"]["bytes"],
           "write_bytes": out["WRITE_SIZE"]["bytes"], "fetch_factor": out["FETCH_SIZE"]["factor"],
           "write_factor": out["WRITE_SIZE"]["factor"], "counters": out, "kernel": args.traffic_kernel,
           "build_id": mia.build_id(), "copy_bytes_each_way": copy_bytes}
    path = os.path.join(ROOT, "profiles", args.traffic_out)
    with open(path, "w") as f:
        json.dump(res, f, indent=1)
    log(f"[traffic] {args.traffic_kernel}: {total / 1e9:.4f} GB per launch -> {path}")
    print(json.dumps(res))
    return 0


if __name__ == "__main__":
    sys.exit(main())
