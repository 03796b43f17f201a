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

import numpy as np

from ._native import check, lib
from .device import DeviceBuffer, Stream

PIVOT_NONE = 0x0
PIVOT_PARALLEL = 0x1
PIVOT_SERIAL = 0x2
PIVOT_TREE = 0x3

LAYOUT_AUTO = 0x00
LAYOUT_ROW16 = 0x10
LAYOUT_QUAD4 = 0x20

PIVOT_NAMES = {"none": PIVOT_NONE, "parallel": PIVOT_PARALLEL, "serial": PIVOT_SERIAL, "tree": PIVOT_TREE}

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


def version() -> str:
    return lib().lu_version().decode()
