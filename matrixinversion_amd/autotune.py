"""Per-n launch tuning for the plain and ragged entry points.

The reference fixes its launch shape by hand per build (NUMTHREADS and the
derived matrices per block, parallel_pivot/run.py:39-41 and
luBatchedInplace.cu:8-11).  Here the shape is a runtime table, one entry per
n, read by liblubatch whenever flags carry LUB_LAYOUT_AUTO:

  layout[n]  0 = 16 lanes per matrix (4 matrices per wave), 1 = one 4-lane
             quad per matrix (16 per wave; plain path n <= 16, ragged n <= 12)
  grab[n]    ragged path: wave tiles per work item (up to 64 matrices)
  pad[n]     plain path, n % 4 == 0: LDS rows padded by 4 floats (1) or not (0)

This module measures both on the current device and writes the table, which
load_table() installs in a process (the library's built-in defaults are a
table measured this way on MI355X):

    python -m matrixinversion_amd.autotune --out tuning.json

The layout is timed with the plain kernel (1 launch over `batch` matrices of
one n, pivoted); grab is timed with the ragged kernel on a single-class
batch of the same n, so each choice is made where it matters.  With
--objective mix the ragged columns are instead tuned on a mixed batch
(n ~ U{2..32}), where the classes share the chip in one launch.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import api
from .device import DeviceBuffer, Event, Stream, copy_kernel, device_info, synchronize
from .workloads import uniform_matrices

PLAIN_QUAD_MAX_N = 16
RAGGED_QUAD_MAX_N = 12
MAX_PER = 64


def _time(fn, stream: Stream, reps: int, trials: int = 3) -> float:
    """Median over `trials` of the mean ms per call of `fn` over `reps`
    back-to-back calls on `stream`."""
    fn()
    stream.synchronize()
    e0, e1 = Event(), Event()
    out = []
    for _ in range(trials):
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        stream.synchronize()
        out.append(e0.elapsed_ms(e1) / reps)
    return float(np.median(out))


def tune_layout(n: int, batch: int, stream: Stream, reps: int = 20) -> tuple[int, dict]:
    """Plain path: 16-lane rows vs quads for one n (n <= 16)."""
    a = uniform_matrices(1000 + n, batch, n)
    src = DeviceBuffer.from_array(a)
    work = DeviceBuffer(a.nbytes)
    piv = DeviceBuffer(batch * n * 4)
    times = {}
    try:
        for name, flag in (("row16", api.LAYOUT_ROW16), ("quad4", api.LAYOUT_QUAD4)):
            if flag == api.LAYOUT_QUAD4 and n > PLAIN_QUAD_MAX_N:
                continue

            def run():
                # inverting in place alternates A and A^-1; restore every call so
                # each launch sees the original inputs (the copy is timed too, but
                # equally for both layouts)
                copy_kernel(work, src, batch * n * n, stream)
                api.lu_batched_inverse(work, n, batch, api.PIVOT_PARALLEL | flag, piv, stream)

            def copy_only():
                copy_kernel(work, src, batch * n * n, stream)

            times[name] = _time(run, stream, reps) - _time(copy_only, stream, reps)
    finally:
        for b in (src, work, piv):
            b.close()
    best = min(times, key=times.get)
    return (1 if best == "quad4" else 0), times


def tune_pad(n: int, layout: int, batch: int, stream: Stream, reps: int = 20) -> tuple[int, dict]:
    """Plain path, n % 4 == 0: LDS rows padded by 4 floats or not, in the
    chosen layout."""
    a = uniform_matrices(1500 + n, batch, n)
    src = DeviceBuffer.from_array(a)
    work = DeviceBuffer(a.nbytes)
    piv = DeviceBuffer(batch * n * 4)
    flag = api.LAYOUT_QUAD4 if layout == 1 else api.LAYOUT_ROW16
    base = api.lds_pad_get()
    times = {}
    try:
        for pad in (0, 1):
            p = base.copy()
            p[n] = pad
            api.lds_pad_set(p)

            def run():
                copy_kernel(work, src, batch * n * n, stream)
                api.lu_batched_inverse(work, n, batch, api.PIVOT_PARALLEL | flag, piv, stream)

            def copy_only():
                copy_kernel(work, src, batch * n * n, stream)

            times[pad] = _time(run, stream, reps) - _time(copy_only, stream, reps)
    finally:
        api.lds_pad_set(base)
        for b in (src, work, piv):
            b.close()
    best = min(times, key=times.get)
    return best, times


def tune_grab(n: int, layout: int, batch: int, stream: Stream, reps: int = 20) -> tuple[int, dict]:
    """Ragged path: wave tiles per work item for one n under `layout`."""
    mpw = 16 if (layout == 1 and n <= RAGGED_QUAD_MAX_N) else 4
    a = uniform_matrices(2000 + n, batch, n).reshape(-1)
    ns = np.full(batch, n, np.int32)
    offs = np.arange(batch, dtype=np.int64) * (n * n)
    poffs = np.arange(batch, dtype=np.int64) * n
    src = DeviceBuffer.from_array(a)
    work = DeviceBuffer(a.nbytes)
    O, NS, PO = (DeviceBuffer.from_array(x) for x in (offs, ns, poffs))
    P = DeviceBuffer(batch * n * 4)
    plan = api.RaggedPlan(batch)
    base_l, base_g = api.tuning_get()
    times = {}
    try:
        for grab in (1, 2, 4, 8, 16):
            if mpw * grab > MAX_PER:
                continue
            g = base_g.copy()
            g[n] = grab
            api.tuning_set(base_l, g)

            def run():
                copy_kernel(work, src, a.size, stream)
                api.lu_batched_inverse_ragged(work, O, NS, batch, api.PIVOT_PARALLEL, P, PO, plan=plan,
                                              stream=stream)

            def copy_only():
                copy_kernel(work, src, a.size, stream)

            times[grab] = _time(run, stream, reps) - _time(copy_only, stream, reps)
    finally:
        api.tuning_set(base_l, base_g)
        plan.close()
        for b in (src, work, O, NS, PO, P):
            b.close()
    best = min(times, key=times.get)
    return best, times


def autotune(sizes=range(1, 33), plain_batch: int = 262144, ragged_batch: int = 131072, log=None) -> dict:
    stream = Stream()
    layout, grab = api.tuning_get()
    layout = layout.copy()
    grab = grab.copy()
    pad = api.lds_pad_get()
    detail = {}
    for n in sizes:
        t0 = time.perf_counter()
        if n <= PLAIN_QUAD_MAX_N:
            layout[n], lt = tune_layout(n, plain_batch, stream)
        else:
            layout[n], lt = 0, {}
        api.tuning_set(layout, grab)
        if n % 4 == 0:
            pad[n], pt = tune_pad(n, int(layout[n]), plain_batch, stream)
            api.lds_pad_set(pad)
        else:
            pt = {}
        grab[n], gt = tune_grab(n, int(layout[n]), ragged_batch, stream)
        api.tuning_set(layout, grab)
        detail[n] = {"layout_ms": lt, "pad_ms": {str(k): v for k, v in pt.items()},
                     "grab_ms": {str(k): v for k, v in gt.items()}}
        if log:
            log(f"[tune] n={n:2d} layout={'quad4' if layout[n] else 'row16'} pad={pad[n]} grab={grab[n]} "
                f"({time.perf_counter() - t0:.1f}s) {lt} {pt} {gt}")
    stream.close()
    info = device_info(0)
    return {"device": info.name, "arch": info.arch, "layout": [int(x) for x in layout],
            "grab": [int(x) for x in grab], "pad": [int(x) for x in pad], "plain_batch": plain_batch,
            "ragged_batch": ragged_batch, "detail": detail}


def tune_mix(batch: int = 500_000, n_lo: int = 2, n_hi: int = 32, rounds: int = 1, reps: int = 10,
             log=None) -> dict:
    """Ragged path, optimised on a mixed batch instead of single classes: in one
    launch the classes share the chip, so a per-class choice that wins alone
    can unbalance the mix.  Coordinate descent over the classes, largest n
    first: for each n every admissible (layout, grab) is tried with the rest
    of the table fixed, timing the whole mixed batch (n ~ U{n_lo..n_hi},
    pivoted), and the fastest is kept.  Starts from the installed table;
    returns the table, the mix time before and after, and the trials."""
    from .workloads import ragged_batch

    stream = Stream()
    flat, offs, ns, poffs = ragged_batch(42, batch, n_lo, n_hi)
    src = DeviceBuffer.from_array(flat)
    work = DeviceBuffer(flat.nbytes)
    O, NS, PO = (DeviceBuffer.from_array(x) for x in (offs, ns, poffs))
    P = DeviceBuffer(int(ns.sum()) * 4)
    plan = api.RaggedPlan(batch)
    layout, grab = (x.copy() for x in api.tuning_get())
    pad = api.lds_pad_get()

    def mix_ms(lay, grb) -> float:
        api.tuning_set(lay, grb)

        def run():
            api.lu_batched_inverse_ragged(work, O, NS, batch, api.PIVOT_PARALLEL, P, PO, plan=plan, stream=stream)

        # inverting in place alternates A and A^-1: both are timed alike
        copy_kernel(work, src, flat.size, stream)
        return _time(run, stream, reps)

    trials = []
    try:
        start = best = mix_ms(layout, grab)
        for _ in range(rounds):
            for n in range(n_hi, n_lo - 1, -1):
                options = []
                for lay in ((0, 1) if n <= RAGGED_QUAD_MAX_N else (0,)):
                    mpw = 16 if lay == 1 else 4
                    options += [(lay, g) for g in (1, 2, 4, 8, 16) if mpw * g <= MAX_PER]
                for lay, g in options:
                    if (lay, g) == (layout[n], grab[n]):
                        continue
                    l2, g2 = layout.copy(), grab.copy()
                    l2[n], g2[n] = lay, g
                    t = mix_ms(l2, g2)
                    trials.append({"n": n, "layout": lay, "grab": g, "ms": t})
                    if t < best * 0.995:  # keep only clear wins (timing noise is about 0.5 %)
                        best, layout, grab = t, l2, g2
                if log:
                    log(f"[tune-mix] n={n:2d} layout={'quad4' if layout[n] else 'row16'} grab={grab[n]} "
                        f"mix {best:.4f} ms")
        final = mix_ms(layout, grab)
    finally:
        plan.close()
        for b in (src, work, O, NS, PO, P):
            b.close()
        stream.close()
    return {"layout": [int(x) for x in layout], "grab": [int(x) for x in grab], "pad": [int(x) for x in pad],
            "mix_batch": batch,
            "mix_ms_start": start, "mix_ms_best": best, "mix_ms_final": final, "trials": trials}


def apply_table(table: dict) -> None:
    """Install a table written by autotune() (33 entries per column)."""
    api.tuning_set(np.array(table["layout"], np.int8), np.array(table["grab"], np.int8))
    if "pad" in table:
        api.lds_pad_set(np.array(table["pad"], np.int8))


def load_table(path: str) -> None:
    with open(path) as f:
        apply_table(json.load(f))


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default="tuning.json")
    ap.add_argument("--sizes", default="1-32")
    ap.add_argument("--plain-batch", type=int, default=262144)
    ap.add_argument("--ragged-batch", type=int, default=131072)
    ap.add_argument("--objective", choices=["class", "mix"], default="class",
                    help="class: each n timed alone (plain layout, ragged grab); mix: the ragged table "
                         "tuned on the mixed batch (--mix-batch matrices, n ~ U{2..32})")
    ap.add_argument("--mix-batch", type=int, default=500_000)
    ap.add_argument("--rounds", type=int, default=1, help="mix: coordinate-descent sweeps over the classes")
    args = ap.parse_args(argv)
    lo, _, hi = args.sizes.partition("-")
    sizes = range(int(lo), int(hi or lo) + 1)
    log = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731
    if args.objective == "mix":
        table = tune_mix(args.mix_batch, rounds=args.rounds, log=log)
    else:
        table = autotune(sizes, args.plain_batch, args.ragged_batch, log=log)
    synchronize()
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1)
    print(json.dumps({"layout": table["layout"], "grab": table["grab"]}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
