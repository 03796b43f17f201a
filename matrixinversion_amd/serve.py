"""Serving: continuous batching of inversion requests (liblubatch lu_server_*).

``Server`` wraps the native engine: requests are admitted into an HBM arena
under a byte budget, coalesced per pivot rule into ragged batches whenever
that rule's stream is idle, and stay resident until released.

``poisson_stream`` drives a server with the mixed serving workload and
reports latency percentiles and throughput:

  * arrivals: Poisson at `rate` requests/s for `duration` seconds;
  * a request holds a log-uniform count of 1..4096 matrices, n ~ U{4..32}
    per matrix, drawn as a contiguous run of a pre-generated device pool;
  * 70 % ask for pivoting (U[-1,1) entries, LUB_PIVOT_PARALLEL), 30 % are
    diagonally dominant (U[-1,1) + n*I) and ask for none;
  * latency = completion seen by the driver - scheduled arrival (it includes
    any wait for admission under the budget).

``warm`` runs a short untimed stream first, so that the first launches of
each kernel (code object loading) stay out of the measured one.

    python -m matrixinversion_amd.serve --rates 2000,8000,20000 --duration 2
"""
from __future__ import annotations

import argparse
import ctypes
import json
import sys
import time

import numpy as np

from . import api
from ._native import NativeError, check, lib
from .device import DeviceBuffer, device_info
from .workloads import ragged_batch

LUB_EBUSY = 9
QUEUED, RUNNING, DONE = 0, 1, 2


class Server:
    def __init__(self, budget_bytes: int, max_batch: int = 1 << 20, ring_slots: int = 1 << 21):
        h = ctypes.c_void_p()
        check(lib().lu_server_create(ctypes.byref(h), int(budget_bytes), int(max_batch), int(ring_slots)),
              "lu_server_create")
        self.handle = h.value
        self._lib = lib()
        self._poll_buf = (ctypes.c_uint64 * 4096)()

    def submit(self, src_ptr: int, ns: np.ndarray, flags: int):
        """Ticket, or None when the request does not fit right now (LUB_EBUSY)."""
        ns = np.ascontiguousarray(ns, dtype=np.int32)
        t = ctypes.c_uint64()
        rc = self._lib.lu_server_submit(self.handle, src_ptr, ns.ctypes.data, int(ns.size), flags, ctypes.byref(t))
        if rc == LUB_EBUSY:
            return None
        check(rc, "lu_server_submit")
        return t.value

    def progress(self):
        la, co = ctypes.c_int(), ctypes.c_int()
        check(self._lib.lu_server_progress(self.handle, ctypes.byref(la), ctypes.byref(co)), "lu_server_progress")
        return la.value, co.value

    def query(self, ticket: int):
        """(state, result device pointer, pivot device pointer or None, latency ms)."""
        st = ctypes.c_int()
        res, piv = ctypes.c_void_p(), ctypes.c_void_p()
        lat = ctypes.c_double()
        check(self._lib.lu_server_query(self.handle, ticket, ctypes.byref(st), ctypes.byref(res), ctypes.byref(piv),
                                        ctypes.byref(lat)), "lu_server_query")
        return st.value, res.value, piv.value, lat.value

    def release(self, ticket: int) -> None:
        check(self._lib.lu_server_release(self.handle, ticket), "lu_server_release")

    def poll(self) -> list:
        """Tickets finished since the last poll, in completion order."""
        out = []
        cnt = ctypes.c_int()
        while True:
            check(self._lib.lu_server_poll(self.handle, self._poll_buf, len(self._poll_buf), ctypes.byref(cnt)),
                  "lu_server_poll")
            out.extend(self._poll_buf[:cnt.value])
            if cnt.value < len(self._poll_buf):
                return out

    def counters(self):
        """(batches launched, most batches of one pivot rule running at once)."""
        b, p = ctypes.c_int64(), ctypes.c_int()
        check(self._lib.lu_server_counters(self.handle, ctypes.byref(b), ctypes.byref(p)), "lu_server_counters")
        return b.value, p.value

    def stats(self):
        q, r, b = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        check(self._lib.lu_server_stats(self.handle, ctypes.byref(q), ctypes.byref(r), ctypes.byref(b)),
              "lu_server_stats")
        return q.value, r.value, b.value

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().lu_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def download(ptr: int, nbytes: int, dtype) -> np.ndarray:
    """Copy `nbytes` of device memory at `ptr` to a new host array."""
    out = np.empty(nbytes // np.dtype(dtype).itemsize, dtype=dtype)
    if nbytes:
        check(lib().lubx_memcpy_dtoh(out.ctypes.data, ptr, nbytes), "memcpy D2H")
    return out


class Pool:
    """Pre-generated request payloads in HBM: a packed ragged batch of
    `size` matrices, n ~ U{n_lo..n_hi}; a request is a contiguous run."""

    def __init__(self, seed: int, size: int, n_lo: int = 4, n_hi: int = 32, dominant: bool = False):
        self.flat, self.offsets, self.ns, _ = ragged_batch(seed, size, n_lo, n_hi, dominant=dominant)
        self.buf = DeviceBuffer.from_array(self.flat)

    def request(self, start: int, count: int):
        """(device pointer, ns) of matrices start .. start+count-1."""
        return self.buf.ptr + int(self.offsets[start]) * 4, self.ns[start:start + count]

    def host(self, start: int, count: int) -> np.ndarray:
        a = int(self.offsets[start])
        b = int(self.offsets[start + count - 1]) + int(self.ns[start + count - 1]) ** 2
        return self.flat[a:b]

    def close(self):
        self.buf.close()


def request_mix(seed: int, count: int, pool_size: int, max_count: int = 4096, pivot_share: float = 0.7):
    """(matrix counts, pool starts, pivot?) of `count` requests: log-uniform
    counts in 1..max_count, uniform starts, Bernoulli(pivot_share) rules."""
    rng = np.random.RandomState(seed)
    counts = np.minimum(np.floor(np.exp(rng.uniform(0.0, np.log(max_count + 1), count))), max_count).astype(np.int64)
    counts = np.maximum(counts, 1)
    starts = (rng.uniform(0, 1, count) * (pool_size - counts + 1)).astype(np.int64)
    pivot = rng.uniform(0, 1, count) < pivot_share
    return counts, starts, pivot


def schedule(pools, rate: float, duration: float, seed: int = 7, max_count: int = 4096):
    """The request stream both drivers replay: (arrival times in s, matrix
    counts, pool starts, pivoted?, flags)."""
    rng = np.random.RandomState(seed)
    n_req = max(1, int(rng.poisson(rate * duration)))
    arrivals = np.sort(rng.uniform(0.0, duration, n_req))
    counts, starts, pivot = request_mix(seed + 1, n_req, min(p.ns.size for p in pools), max_count)
    flags = np.where(pivot, api.PIVOT_PARALLEL, api.PIVOT_NONE)
    return arrivals, counts, starts, pivot, flags


def _summary(rate, lat, counts, pivot, wall, batches, busy_retries, peak, driver) -> dict:
    mats = int(counts.sum())
    return {"rate_req_s": rate, "requests": int(lat.size), "matrices": mats, "wall_s": wall,
            "matrices_per_s": mats / wall, "requests_per_s": lat.size / wall,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "max_ms": float(lat.max()), "batches": batches,
            "matrices_per_batch": mats / max(batches, 1), "busy_retries": busy_retries,
            "pivot_share": float(pivot.mean()), "peak_rule_inflight_since_create": peak, "driver": driver}


def poisson_stream_native(server: Server, pools, rate: float, duration: float, seed: int = 7,
                          max_count: int = 4096) -> dict:
    """The same stream as poisson_stream, replayed by the library's host loop
    (lubx_serve_replay): submit, progress, poll and release with no
    interpreter between them, so the latencies are the engine's."""
    arrivals, counts, starts, pivot, flags = schedule(pools, rate, duration, seed, max_count)
    n_req = arrivals.size
    src = np.empty(n_req, np.uint64)
    ns_off = np.empty(n_req, np.int64)
    ns_parts = []
    off = 0
    for i in range(n_req):
        pool = pools[0] if pivot[i] else pools[1]
        ptr, ns = pool.request(int(starts[i]), int(counts[i]))
        src[i] = ptr
        ns_off[i] = off
        ns_parts.append(ns)
        off += ns.size
    ns_all = np.ascontiguousarray(np.concatenate(ns_parts), dtype=np.int32)
    cnt = np.ascontiguousarray(counts, dtype=np.int32)
    flg = np.ascontiguousarray(flags, dtype=np.int32)
    arr = np.ascontiguousarray(arrivals, dtype=np.float64)
    lat = np.full(n_req, np.nan)
    stats = np.zeros(2, np.int64)
    _, peak0 = server.counters()
    t0 = time.perf_counter()
    check(server._lib.lubx_serve_replay(server.handle, int(n_req), arr.ctypes.data, src.ctypes.data,
                                        ns_all.ctypes.data, ns_off.ctypes.data, cnt.ctypes.data, flg.ctypes.data,
                                        lat.ctypes.data, stats.ctypes.data), "lubx_serve_replay")
    wall = time.perf_counter() - t0
    _, peak = server.counters()
    return _summary(rate, lat, counts, pivot, wall, int(stats[0]), int(stats[1]), peak, "native")


def poisson_stream(server: Server, pools, rate: float, duration: float, seed: int = 7,
                   max_count: int = 4096) -> dict:
    """Drive `server` with Poisson arrivals from a Python loop; returns
    latency percentiles and throughput.  pools = (pivoted pool, no-pivot
    pool)."""
    arrivals, counts, starts, pivot, flags = schedule(pools, rate, duration, seed, max_count)
    n_req = arrivals.size
    lat = np.full(n_req, np.nan)
    busy_retries = 0
    inflight = {}          # ticket -> request index
    nxt = 0
    done = 0
    batches = 0
    b0, _ = server.counters()
    t0 = time.perf_counter()
    while done < n_req:
        now = time.perf_counter() - t0
        while nxt < n_req and arrivals[nxt] <= now:
            pool = pools[0] if pivot[nxt] else pools[1]
            ptr, ns = pool.request(int(starts[nxt]), int(counts[nxt]))
            t = server.submit(ptr, ns, int(flags[nxt]))
            if t is None:        # over budget: retry after progress
                busy_retries += 1
                break
            inflight[t] = nxt
            nxt += 1
        launched, completed = server.progress()
        batches += launched
        if completed:
            now = time.perf_counter() - t0
            for t in server.poll():
                i = inflight.pop(t)
                lat[i] = (now - arrivals[i]) * 1e3
                server.release(t)
                done += 1
    wall = time.perf_counter() - t0
    b1, peak = server.counters()
    return _summary(rate, lat, counts, pivot, wall, batches, busy_retries, peak, "python")


def warm(server: Server, pools, rate: float = 4000.0, duration: float = 0.15, driver: str = "python") -> None:
    """An untimed stream through `server` (both pivot rules): the first launch
    of each kernel loads its code object, which a measured stream should not
    see."""
    (poisson_stream_native if driver == "native" else poisson_stream)(server, pools, rate, duration, seed=1001)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="Poisson serving sweep over lu_server_*")
    ap.add_argument("--rates", default="1000,4000,16000")
    ap.add_argument("--duration", type=float, default=2.0)
    ap.add_argument("--budget-gb", type=float, default=64.0)
    ap.add_argument("--pool", type=int, default=200_000)
    ap.add_argument("--driver", choices=["native", "python"], default="native",
                    help="native: the library's host loop (lubx_serve_replay); python: this module's loop")
    args = ap.parse_args(argv)
    info = device_info(0)
    print(f"[serve] {info.name} {info.arch}", file=sys.stderr, flush=True)
    pools = (Pool(42, args.pool), Pool(43, args.pool, dominant=True))
    out = []
    run = poisson_stream_native if args.driver == "native" else poisson_stream
    with Server(int(args.budget_gb * 2**30)) as srv:
        warm(srv, pools, driver=args.driver)
        for r in (float(x) for x in args.rates.split(",")):
            res = run(srv, pools, r, args.duration)
            print(json.dumps(res), flush=True)
            out.append(res)
    for p in pools:
        p.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
