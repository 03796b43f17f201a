"""matrixinversion_amd -- batched small-matrix (n <= 32) FP32 LU inversion for
AMD Instinct MI355X (gfx950).

The product is the C-ABI library ``_lib/liblubatch.so`` (include/lubatch.h)
with hand-written HIP kernels; this package binds it with ctypes and carries
the sweep harness that replaces the reference's run.py, the autotuner and the
serving driver.
"""
from .api import (
    LAYOUT_AUTO, LAYOUT_QUAD4, LAYOUT_ROW16, MAX_N, PIVOT_NAMES, PIVOT_NONE, PIVOT_PARALLEL, PIVOT_SERIAL,
    PIVOT_TREE, PIVOT_GEPP, RaggedPlan, invert, invert_lapack, invert_ragged, invert_refined, lu_batched_inverse,
    lu_batched_inverse_ragged, pack_ragged, refine, residual, tuning_get, tuning_set, version, build_id, lds_pad_get, lds_pad_set,
)
from .device import DeviceBuffer, Event, Graph, Stream, device_count, device_info, synchronize

__all__ = [
    "LAYOUT_AUTO", "LAYOUT_QUAD4", "LAYOUT_ROW16", "MAX_N", "PIVOT_NAMES", "PIVOT_NONE", "PIVOT_PARALLEL",
    "PIVOT_SERIAL", "PIVOT_TREE", "PIVOT_GEPP", "RaggedPlan", "invert", "invert_lapack", "invert_ragged", "invert_refined",
    "lu_batched_inverse", "lu_batched_inverse_ragged", "pack_ragged", "refine", "residual", "tuning_get",
    "tuning_set", "version", "build_id", "lds_pad_get", "lds_pad_set", "DeviceBuffer", "Event", "Graph", "Stream", "device_count", "device_info",
    "synchronize",
]
