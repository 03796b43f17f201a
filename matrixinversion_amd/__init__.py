"""matrixinversion_amd -- batched small-matrix (n <= 32) FP32 LU inversion for
AMD Instinct MI355X (gfx950).

The product is the C-ABI library ``_lib/liblubatch.so`` (include/lubatch.h)
with hand-written HIP kernels; this package binds it with ctypes and carries
the sweep harness that replaces the reference's run.py.
"""
from .api import (MAX_N, PIVOT_NAMES, PIVOT_NONE, PIVOT_PARALLEL, PIVOT_SERIAL, PIVOT_TREE, invert,
                  lu_batched_inverse, residual, version)
from .device import DeviceBuffer, Event, Stream, device_count, device_info, synchronize

__all__ = [
    "MAX_N", "PIVOT_NAMES", "PIVOT_NONE", "PIVOT_PARALLEL", "PIVOT_SERIAL", "PIVOT_TREE", "invert",
    "lu_batched_inverse", "residual", "version", "DeviceBuffer", "Event", "Stream", "device_count", "device_info",
    "synchronize",
]
