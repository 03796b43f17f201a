"""Sweep harness: the replacement of the reference's run.py.

The reference's sweep (parallel_pivot/run.py:169-261) recompiles the CUDA
driver for every (n, batch, threads) with -D macros, profiles it under ncu,
runs the binary 10 times as separate processes (one cold launch each), parses
"Kernel execution time" from stdout and writes one JSON per batch size.  Its
correctness field is never filled (the parse sits after a `break`,
run.py:136-151).

Here n, batch and the pivot rule are runtime arguments of one library: the
sweep runs in one process against device-resident inputs, times every launch
with HIP events on the launch stream (warm launches; the first, cold one is
recorded separately), checks every result on the device (lu_batched_residual,
the reference's verifyInv count) and writes JSON with the reference's keys --
matrix_size, num_matrices, num_threads, runtimes, runtime_avg, variance,
std_dev, incorrect_inversions (now populated) -- plus throughput and HBM
fractions.  Output layout: <out>/runtime_results/<100k|500k|1M|...>/
benchmark_results_<label>.json, keys "m{n}_n{batch}_t{threads}".

Beyond the reference's keys each record carries max_abs_AXmI (device),
max_rel_err_vs_fp64 (max |X - A^-1| / max |A^-1| against numpy's FP64 inverse,
over the first CHECK_SAMPLE matrices), pivot_mismatches (matrices of that
sample whose pivot vector differs from the reference driver's own host
check, pivotedA: first maximum |a| of the un-updated column,
parallel_pivot/luBatchedInplace.cu:148-149 -> verify.hpp:106-155; null for the
tree and partial-pivoting rules, which pivotedA does not describe) and
cpu_cores (host cores available to the process).

Inputs: "random" (default) fills every matrix from one mt19937 stream (seeded
per n) with a = 2*((x>>8)*2^-24) - 1; "dominant" adds n*I; "template:<file>"
reproduces the reference's input construction -- the first n*n numbers of a
text fixture replicated batch times (parallel_pivot/luBatchedInplace.cu:44-53,
88-101).

    python -m matrixinversion_amd.harness --sizes 1-32 --batches 100000,500000,1000000 --pivot parallel
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import api
from .device import DeviceBuffer, Event, Stream, copy_kernel, device_info
from .workloads import make_inputs

HBM_PEAK_GBPS = 8000.0
THREADS_PER_WORKGROUP = 64  # one wave owns a tile of 4 matrices
CHECK_SAMPLE = 1024         # matrices checked on the host per configuration


def cpu_cores() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def pivoted_a(a: np.ndarray) -> np.ndarray:
    """The reference driver's host pivot check (pivotedA, parallel_pivot/
    verify.hpp:106-155) over a batch: at step k the row at positions k..n-1
    with the largest |a[.][k]| of the row-permuted input (no elimination) is
    swapped to position k; strict '>' so the first maximum wins, a NaN at
    position k is never beaten and a NaN elsewhere never wins.  Returns
    perm (batch, n): perm[p] = original row at position p."""
    b, n, _ = a.shape
    perm = np.tile(np.arange(n, dtype=np.int64), (b, 1))
    mag = np.abs(a.astype(np.float32))
    idx = np.arange(b)
    for k in range(n):
        vals = np.take_along_axis(mag[:, :, k], perm[:, k:], axis=1)
        keep = np.isnan(vals[:, 0])
        best = np.argmax(np.where(np.isnan(vals), -np.inf, vals), axis=1) + k
        best = np.where(keep, k, best)
        pk = perm[idx, k].copy()
        perm[idx, k] = perm[idx, best]
        perm[idx, best] = pk
    return perm.astype(np.int32)


def host_checks(a: np.ndarray, x: np.ndarray, piv: np.ndarray, flags: int):
    """(max_rel_err_vs_fp64, pivot_mismatches) over a sample: a, x (count, n, n),
    piv (count, n)."""
    a64 = a.astype(np.float64)
    try:
        inv = np.linalg.inv(a64)
    except np.linalg.LinAlgError:
        inv = np.stack([np.linalg.pinv(m) for m in a64])
    num = np.abs(x.astype(np.float64) - inv)
    num = np.where(np.isnan(num), np.inf, num).max(axis=(1, 2))
    rel = float((num / np.abs(inv).max(axis=(1, 2))).max()) if len(a) else 0.0
    if flags in (api.PIVOT_PARALLEL, api.PIVOT_SERIAL):
        mism = int((pivoted_a(a) != piv).any(axis=1).sum())
    elif flags == api.PIVOT_NONE:
        mism = int((piv != np.arange(a.shape[-1])).any(axis=1).sum())
    else:
        mism = None
    return rel, mism


def _batch_label(batch: int) -> str:
    if batch % 1_000_000 == 0:
        return f"{batch // 1_000_000}M"
    if batch % 1000 == 0:
        return f"{batch // 1000}k"
    return str(batch)


def run_config(n: int, batch: int, flags: int, runs: int, inputs: str, seed: int, stream: Stream) -> dict:
    a = make_inputs(inputs, n, batch, seed)
    pristine = DeviceBuffer.from_array(a)
    work = DeviceBuffer(a.nbytes)
    piv = DeviceBuffer(max(batch * n * 4, 4))
    sample = min(batch, CHECK_SAMPLE)
    a_sample = a[:sample].copy()
    del a
    e0, e1 = Event(), Event()
    times = []
    bad_counts = []
    max_res = 0.0
    try:
        for r in range(runs + 1):
            copy_kernel(work, pristine, batch * n * n, stream)
            stream.synchronize()
            e0.record(stream)
            api.lu_batched_inverse(work, n, batch, flags, piv, stream)
            e1.record(stream)
            stream.synchronize()
            times.append(e0.elapsed_ms(e1))
            correct, worst = api.residual(pristine, work, n, batch, 1e-3, stream)
            bad_counts.append(batch - correct)
            max_res = max(max_res, worst) if worst == worst else float("nan")
        x_sample = work.download((sample, n, n), np.float32)
        p_sample = piv.download((sample, n), np.int32)
    finally:
        pristine.close()
        work.close()
        piv.close()
    rel, mism = host_checks(a_sample, x_sample, p_sample, flags)
    cold, warm = times[0], np.array(times[1:])
    bytes_per = 8 * n * n + (4 * n if flags != api.PIVOT_NONE else 0)
    avg = float(warm.mean())
    return {
        "matrix_size": n,
        "num_matrices": batch,
        "num_threads": THREADS_PER_WORKGROUP,
        "runtimes": [float(t) for t in warm],
        "runtime_avg": avg,
        "variance": float(np.var(warm)),
        "std_dev": float(np.std(warm)),
        "incorrect_inversions": bad_counts[1:],
        "cold_runtime": float(cold),
        "p50_ms": float(np.percentile(warm, 50)),
        "p99_ms": float(np.percentile(warm, 99)),
        "matrices_per_s": batch / (avg * 1e-3),
        "gbps": batch * bytes_per / (avg * 1e-3) / 1e9,
        "pct_hbm": 100.0 * batch * bytes_per / (avg * 1e-3) / 1e9 / HBM_PEAK_GBPS,
        "max_abs_AXmI": max_res,
        "max_rel_err_vs_fp64": rel,
        "pivot_mismatches": mism,
        "check_sample": sample,
        "cpu_cores": cpu_cores(),
        "pivot": [k for k, v in api.PIVOT_NAMES.items() if v == flags][0],
        "inputs": inputs,
    }


def parse_sizes(spec: str) -> list[int]:
    out = []
    for part in spec.split(","):
        if "-" in part:
            lo, hi = part.split("-")
            out.extend(range(int(lo), int(hi) + 1))
        else:
            out.append(int(part))
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--sizes", default="1-32")
    ap.add_argument("--batches", default="100000,500000,1000000")
    ap.add_argument("--pivot", default="parallel", choices=sorted(api.PIVOT_NAMES))
    ap.add_argument("--runs", type=int, default=10)
    ap.add_argument("--inputs", default="random", help="random | dominant | template:<file>")
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--out", default="benchmark_results")
    ap.add_argument("--label", default=None)
    args = ap.parse_args(argv)

    flags = api.PIVOT_NAMES[args.pivot]
    label = args.label or args.pivot
    info = device_info(0)
    stream = Stream()
    for batch in [int(b) for b in args.batches.split(",")]:
        outdir = os.path.join(args.out, "runtime_results", _batch_label(batch))
        os.makedirs(outdir, exist_ok=True)
        path = os.path.join(outdir, f"benchmark_results_{label}.json")
        results = {}
        if os.path.exists(path):  # resume: keep configurations already measured
            with open(path) as f:
                results = json.load(f)
        for n in parse_sizes(args.sizes):
            key = f"m{n}_n{batch}_t{THREADS_PER_WORKGROUP}"
            if key in results:
                continue
            t0 = time.time()
            res = run_config(n, batch, flags, args.runs, args.inputs, args.seed + n, stream)
            res["device"] = {"arch": info.arch, "cu_count": info.cu_count}
            results[key] = res
            tmp = path + ".tmp"
            with open(tmp, "w") as f:
                json.dump(results, f, indent=2)
            os.replace(tmp, path)
            print(f"{key}: {res['runtime_avg']:.4f} ms  {res['matrices_per_s'] / 1e6:9.1f} M/s  "
                  f"{res['gbps']:7.0f} GB/s ({res['pct_hbm']:.1f}%)  incorrect={res['incorrect_inversions'][-1]}  "
                  f"max|AX-I|={res['max_abs_AXmI']:.2e}  [{time.time() - t0:.1f}s]", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
