"""Device memory, streams and events over liblubatch's utility exports.

The harness needs HBM buffers, host<->device copies and event timing but no
framework: these classes wrap the corresponding HIP runtime calls exported by
liblubatch.so (include/lubatch_util.h).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from ._native import NativeError, check, lib


@dataclass
class DeviceInfo:
    name: str
    arch: str
    cu_count: int
    clock_khz: int
    total_mem: int
    mem_clock_khz: int
    mem_bus_width: int
    lds_per_block: int


def device_count() -> int:
    return int(lib().lubx_device_count())


def device_info(dev: int = 0) -> DeviceInfo:
    name = ctypes.create_string_buffer(256)
    arch = ctypes.create_string_buffer(256)
    cu, clk, mclk, bw = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    mem, lds = ctypes.c_size_t(), ctypes.c_size_t()
    check(lib().lubx_device_info(dev, name, 256, arch, 256, ctypes.byref(cu), ctypes.byref(clk), ctypes.byref(mem),
                                 ctypes.byref(mclk), ctypes.byref(bw), ctypes.byref(lds)), "device_info")
    return DeviceInfo(name.value.decode(), arch.value.decode(), cu.value, clk.value, mem.value, mclk.value,
                      bw.value, lds.value)


def synchronize() -> None:
    check(lib().lubx_device_sync(), "device sync")


class DeviceBuffer:
    """A device allocation of `nbytes` bytes (freed on close / garbage collection)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = lib().lubx_malloc(max(self.nbytes, 1))
        if not self.ptr:
            raise NativeError(f"device allocation of {self.nbytes} bytes failed")

    @classmethod
    def from_array(cls, arr: np.ndarray) -> "DeviceBuffer":
        arr = np.ascontiguousarray(arr)
        buf = cls(arr.nbytes)
        buf.upload(arr)
        return buf

    def upload(self, arr: np.ndarray, offset: int = 0) -> None:
        arr = np.ascontiguousarray(arr)
        if offset + arr.nbytes > self.nbytes:
            raise ValueError("upload exceeds buffer")
        if arr.nbytes:
            check(lib().lubx_memcpy_htod(self.ptr + offset, arr.ctypes.data, arr.nbytes), "memcpy H2D")

    def download(self, shape, dtype, offset: int = 0) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        if offset + out.nbytes > self.nbytes:
            raise ValueError("download exceeds buffer")
        if out.nbytes:
            check(lib().lubx_memcpy_dtoh(out.ctypes.data, self.ptr + offset, out.nbytes), "memcpy D2H")
        return out

    def copy_from(self, other: "DeviceBuffer", stream=None, nbytes: int | None = None) -> None:
        n = self.nbytes if nbytes is None else nbytes
        check(lib().lubx_memcpy_dtod_async(self.ptr, other.ptr, n, stream.handle if stream else None), "memcpy D2D")

    def fill_bytes(self, value: int) -> None:
        """Set every byte; returns once the device memory holds the value, so
        work enqueued next on any stream sees it."""
        check(lib().lubx_memset(self.ptr, value, self.nbytes), "memset")

    def close(self) -> None:
        if getattr(self, "ptr", None):
            lib().lubx_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Stream:
    """A non-blocking, default-priority HIP stream."""

    def __init__(self):
        self.handle = lib().lubx_stream_create()
        if not self.handle:
            raise NativeError("stream creation failed")

    def synchronize(self) -> None:
        check(lib().lubx_stream_sync(self.handle), "stream sync")

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().lubx_stream_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Event:
    def __init__(self):
        self.handle = lib().lubx_event_create()
        if not self.handle:
            raise NativeError("event creation failed")

    def record(self, stream: Stream | None = None) -> None:
        check(lib().lubx_event_record(self.handle, stream.handle if stream else None), "event record")

    def synchronize(self) -> None:
        check(lib().lubx_event_sync(self.handle), "event sync")

    def elapsed_ms(self, end: "Event") -> float:
        ms = float(lib().lubx_event_elapsed_ms(self.handle, end.handle))
        if ms < 0:
            raise NativeError("event elapsed time failed")
        return ms

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().lubx_event_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def copy_kernel(dst: DeviceBuffer, src: DeviceBuffer, count: int, stream: Stream | None = None) -> None:
    """Run the plain dwordx4 copy kernel over `count` floats."""
    check(lib().lubx_copy_f32(dst.ptr, src.ptr, count, stream.handle if stream else None), "copy kernel")


class Graph:
    """Work enqueued on `stream` between begin() and end(), replayed by launch()
    (hipStreamBeginCapture / hipGraphInstantiate / hipGraphLaunch)."""

    def __init__(self, stream: Stream):
        self.stream = stream
        self.handle = None

    def begin(self) -> None:
        check(lib().lubx_graph_capture_begin(self.stream.handle), "graph capture begin")

    def end(self) -> None:
        h = ctypes.c_void_p()
        check(lib().lubx_graph_capture_end(self.stream.handle, ctypes.byref(h)), "graph capture end")
        self.handle = h.value

    def launch(self) -> None:
        check(lib().lubx_graph_launch(self.handle, self.stream.handle), "graph launch")

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().lubx_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
