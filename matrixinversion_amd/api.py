"""Python face of the C ABI (include/lubatch.h).

``lu_batched_inverse`` mirrors the reference's kernel surface
(``batched_lu_subwarp<T, matrixSize, ...>(T* A)``, parallel_pivot/
luBatchedInplace.cuh:127-129, and the runtime form of no_templ/
luBatchedInplace.cuh:82-85): a device batch of contiguous row-major FP32 n x n
matrices is inverted in place, with the pivot rule chosen by ``flags``.
Unlike the reference it also returns the pivot vectors and reports errors as
exceptions instead of exiting.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._native import check, lib
from .device import DeviceBuffer, Stream

PIVOT_NONE = 0x0
PIVOT_PARALLEL = 0x1
PIVOT_SERIAL = 0x2
PIVOT_TREE = 0x3
PIVOT_GEPP = 0x4  # LAPACK getrf's partial pivoting (updated column); not a reference rule

LAYOUT_AUTO = 0x00
LAYOUT_ROW16 = 0x10
LAYOUT_QUAD4 = 0x20

PIVOT_NAMES = {"none": PIVOT_NONE, "parallel": PIVOT_PARALLEL, "serial": PIVOT_SERIAL, "tree": PIVOT_TREE, "gepp": PIVOT_GEPP}

MAX_N = 32


def lu_batched_inverse(A: DeviceBuffer, n: int, batch: int, flags: int = PIVOT_PARALLEL,
                       piv: DeviceBuffer | None = None, stream: Stream | None = None) -> None:
    """Invert `batch` n x n matrices held in device buffer `A` in place.

    With a stream the call is asynchronous (lu_batched_inverse_async64);
    without one it waits for completion (lu_batched_inverse)."""
    if batch * n * n * 4 > A.nbytes:
        raise ValueError("A is smaller than batch*n*n floats")
    if piv is not None and batch * n * 4 > piv.nbytes:
        raise ValueError("piv is smaller than batch*n ints")
    pp = piv.ptr if piv is not None else None
    if stream is None:
        if batch > 2**31 - 1:
            raise ValueError("synchronous entry takes a 32-bit batch; pass a stream")
        check(lib().lu_batched_inverse(A.ptr, pp, n, batch, flags), "lu_batched_inverse")
    else:
        check(lib().lu_batched_inverse_async64(A.ptr, pp, n, batch, flags, stream.handle),
              "lu_batched_inverse_async64")


def residual(A: DeviceBuffer, X: DeviceBuffer, n: int, batch: int, thr: float = 1e-3,
             stream: Stream | None = None):
    """verifyInv on the device: (number of matrices whose A X is within thr of I
    entry-wise, max |A X - I|)."""
    out = DeviceBuffer(16)
    try:
        out.fill_bytes(0)
        check(lib().lu_batched_residual(A.ptr, X.ptr, n, batch, thr, out.ptr, out.ptr + 8,
                                        stream.handle if stream else None), "lu_batched_residual")
        if stream is not None:
            stream.synchronize()
        else:
            from .device import synchronize
            synchronize()
        raw = out.download((4,), np.uint32)
    finally:
        out.close()
    correct = int(raw[0]) | (int(raw[1]) << 32)
    worst = float(np.array([raw[2]], dtype=np.uint32).view(np.float32)[0])
    if raw[2] == 0x7f800001:
        worst = float("nan")
    return correct, worst


def invert(a: np.ndarray, flags: int = PIVOT_PARALLEL, return_pivots: bool = True):
    """Host convenience: upload a (batch, n, n) FP32 array, invert on the GPU,
    download.  Returns (inverses, pivots) or inverses."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.ndim == 2:
        a = a[None]
    batch, n, n2 = a.shape
    if n != n2:
        raise ValueError("matrices must be square")
    A = DeviceBuffer.from_array(a)
    P = DeviceBuffer(max(batch * n * 4, 4)) if return_pivots else None
    try:
        lu_batched_inverse(A, n, batch, flags, P)
        x = A.download(a.shape, np.float32)
        if return_pivots:
            return x, P.download((batch, n), np.int32)
        return x
    finally:
        A.close()
        if P is not None:
            P.close()


def invert_lapack(a: np.ndarray, flags: int = PIVOT_PARALLEL):
    """Host convenience with getrf/getri-shaped outputs: (inverses, ipiv
    (1-based interchanges, batch x n), info (batch))."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.ndim == 2:
        a = a[None]
    batch, n, _ = a.shape
    A = DeviceBuffer.from_array(a)
    P = DeviceBuffer(max(batch * n * 4, 4))
    I = DeviceBuffer(max(batch * 4, 4))
    try:
        check(lib().lu_batched_inverse_lapack(A.ptr, P.ptr, I.ptr, n, batch, flags, None), "lu_batched_inverse_lapack")
        from .device import synchronize
        synchronize()
        return (A.download(a.shape, np.float32), P.download((batch, n), np.int32), I.download((batch,), np.int32))
    finally:
        for b in (A, P, I):
            b.close()


def refine(A: DeviceBuffer, X: DeviceBuffer, n: int, batch: int, stream: Stream | None = None) -> None:
    """X <- X + X (I - A X): FP64 residual, FP32 MFMA correction
    (lu_batched_refine).  Waits for completion when no stream is given."""
    if batch * n * n * 4 > min(A.nbytes, X.nbytes):
        raise ValueError("A or X is smaller than batch*n*n floats")
    check(lib().lu_batched_refine(A.ptr, X.ptr, n, batch, stream.handle if stream else None), "lu_batched_refine")
    if stream is None:
        from .device import synchronize
        synchronize()


def invert_refined(a: np.ndarray, flags: int = PIVOT_PARALLEL):
    """Host convenience: (plain FP32 inverse, inverse after one refinement
    step) of a (batch, n, n) FP32 array."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.ndim == 2:
        a = a[None]
    batch, n, _ = a.shape
    A = DeviceBuffer.from_array(a)
    X = DeviceBuffer(a.nbytes)
    P = DeviceBuffer(max(batch * n * 4, 4))
    try:
        X.copy_from(A)
        lu_batched_inverse(X, n, batch, flags, P)
        x0 = X.download(a.shape, np.float32)
        refine(A, X, n, batch)
        x1 = X.download(a.shape, np.float32)
    finally:
        for b in (A, X, P):
            b.close()
    return x0, x1


class RaggedPlan:
    """Device workspace for ragged batches of up to `max_batch` matrices
    (lu_ragged_plan_create).  Enqueueing on a plan allocates nothing and never
    synchronises, so it can be captured in a Graph."""

    def __init__(self, max_batch: int):
        h = ctypes.c_void_p()
        check(lib().lu_ragged_plan_create(ctypes.byref(h), int(max_batch)), "lu_ragged_plan_create")
        self.handle = h.value
        self.max_batch = int(max_batch)

    def invalid_count(self) -> int:
        out = ctypes.c_int()
        check(lib().lu_ragged_invalid_count(self.handle, ctypes.byref(out)), "lu_ragged_invalid_count")
        return out.value

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().lu_ragged_plan_destroy(self.handle)
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


def lu_batched_inverse_ragged(A: DeviceBuffer, offsets: DeviceBuffer, ns: DeviceBuffer, batch: int,
                              flags: int = PIVOT_PARALLEL, piv: DeviceBuffer | None = None,
                              piv_offsets: DeviceBuffer | None = None, plan: RaggedPlan | None = None,
                              stream: Stream | None = None) -> None:
    """Invert a packed mixed-n batch in place: matrix i has size ns[i] and
    starts at float offset offsets[i] (int64) of A; pivots go to
    piv + piv_offsets[i].  Asynchronous on `stream` (default stream if None).
    With a plan the call is lu_batched_inverse_ragged_plan, else the
    per-thread-plan form."""
    if batch > offsets.nbytes // 8 or batch > ns.nbytes // 4:
        raise ValueError("offsets / ns hold fewer than batch entries")
    if (piv is None) != (piv_offsets is None):
        raise ValueError("piv and piv_offsets go together")
    pp = piv.ptr if piv is not None else None
    po = piv_offsets.ptr if piv_offsets is not None else None
    sh = stream.handle if stream is not None else None
    if plan is not None:
        check(lib().lu_batched_inverse_ragged_plan(plan.handle, A.ptr, offsets.ptr, ns.ptr, pp, po, batch, flags, sh),
              "lu_batched_inverse_ragged_plan")
    else:
        check(lib().lu_batched_inverse_ragged(A.ptr, offsets.ptr, ns.ptr, pp, po, batch, flags, sh),
              "lu_batched_inverse_ragged")


def pack_ragged(mats):
    """Host packing of a list of square FP32 matrices: (flat A, offsets int64,
    ns int32, pivot offsets int64)."""
    ns = np.array([m.shape[0] for m in mats], dtype=np.int32)
    sizes = ns.astype(np.int64) ** 2
    offsets = np.zeros(len(mats), dtype=np.int64)
    if len(mats) > 1:
        offsets[1:] = np.cumsum(sizes)[:-1]
    poffs = np.zeros(len(mats), dtype=np.int64)
    if len(mats) > 1:
        poffs[1:] = np.cumsum(ns.astype(np.int64))[:-1]
    flat = (np.concatenate([np.ascontiguousarray(m, dtype=np.float32).ravel() for m in mats])
            if mats else np.zeros(0, np.float32))
    return flat, offsets, ns, poffs


def invert_ragged(mats, flags: int = PIVOT_PARALLEL):
    """Host convenience for a list of square matrices of mixed n: returns
    (inverses, pivots) as lists."""
    flat, offsets, ns, poffs = pack_ragged(mats)
    bufs = [DeviceBuffer.from_array(x) for x in (flat, offsets, ns)]
    P = DeviceBuffer(max(int(ns.sum()) * 4, 4))
    PO = DeviceBuffer.from_array(poffs)
    try:
        A, O, NS = bufs
        lu_batched_inverse_ragged(A, O, NS, len(mats), flags, P, PO)
        from .device import synchronize
        synchronize()
        x = A.download(flat.shape, np.float32)
        p = P.download((int(ns.sum()),), np.int32)
    finally:
        for b in bufs + [P, PO]:
            b.close()
    inv = [x[o:o + n * n].reshape(n, n) for o, n in zip(offsets, ns)]
    pv = [p[o:o + n] for o, n in zip(poffs, ns)]
    return inv, pv


def tuning_get():
    """The per-n (layout, grab) table used with LAYOUT_AUTO (33 entries each)."""
    layout = ctypes.create_string_buffer(33)
    grab = ctypes.create_string_buffer(33)
    check(lib().lu_tuning_get(layout, grab), "lu_tuning_get")
    return (np.frombuffer(layout.raw, dtype=np.int8).copy(), np.frombuffer(grab.raw, dtype=np.int8).copy())


def tuning_set(layout=None, grab=None) -> None:
    """Replace the per-n table (None keeps the current column)."""
    cur_l, cur_g = tuning_get()
    lay = np.asarray(cur_l if layout is None else layout, dtype=np.int8)
    grb = np.asarray(cur_g if grab is None else grab, dtype=np.int8)
    if lay.shape != (33,) or grb.shape != (33,):
        raise ValueError("tables hold 33 entries (index n, entry 0 unused)")
    check(lib().lu_tuning_set(lay.tobytes(), grb.tobytes()), "lu_tuning_set")


def lds_pad_get():
    """Per-n LDS row padding of the plain kernels (33 entries, 1 = padded)."""
    pad = ctypes.create_string_buffer(33)
    check(lib().lu_tuning_get_lds_pad(pad), "lu_tuning_get_lds_pad")
    return np.frombuffer(pad.raw, dtype=np.int8).copy()


def lds_pad_set(pad) -> None:
    """Replace the per-n LDS row padding (only n % 4 == 0 may be 1)."""
    p = np.asarray(pad, dtype=np.int8)
    if p.shape != (33,):
        raise ValueError("the table holds 33 entries (index n, entry 0 unused)")
    check(lib().lu_tuning_set_lds_pad(p.tobytes()), "lu_tuning_set_lds_pad")


def version() -> str:
    return lib().lu_version().decode()


def build_id() -> str:
    """Hash of the sources the loaded library was built from."""
    return lib().lu_build_id().decode()
