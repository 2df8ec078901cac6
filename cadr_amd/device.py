"""Device memory, streams and events over the C ABI (thin Python wrappers used by
tests and the benchmark; the product API is the C ABI and the C++ CadR facade)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as N


def set_device(index: int = 0) -> None:
    N.check(N.load().cadrSetDevice(index))


def device_count() -> int:
    c = ctypes.c_int(0)
    N.check(N.load().cadrGetDeviceCount(ctypes.byref(c)))
    return c.value


def compute_units() -> int:
    c = ctypes.c_int(0)
    N.check(N.load().cadrGetComputeUnitCount(ctypes.byref(c)))
    return c.value


def synchronize() -> None:
    N.check(N.load().cadrDeviceSynchronize())


class DeviceBuffer:
    """A hipMalloc allocation; freed on close() or garbage collection."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        a = ctypes.c_uint64(0)
        N.check(N.load().cadrDeviceMalloc(self.nbytes, ctypes.byref(a)))
        self.address = a.value

    @classmethod
    def from_array(cls, arr: np.ndarray, stream=None) -> "DeviceBuffer":
        arr = np.ascontiguousarray(arr)
        b = cls(max(arr.nbytes, 16))
        b.upload(arr, stream)
        return b

    def upload(self, arr: np.ndarray, stream=None, offset: int = 0) -> None:
        arr = np.ascontiguousarray(arr)
        assert offset + arr.nbytes <= self.nbytes
        N.check(N.load().cadrMemcpyHtoD(self.address + offset, N.ptr(arr), arr.nbytes, stream))
        if stream is None:
            N.check(N.load().cadrStreamSynchronize(None))

    def download(self, dtype, count: int | None = None, stream=None, offset: int = 0) -> np.ndarray:
        dtype = np.dtype(dtype)
        if count is None:
            count = (self.nbytes - offset) // dtype.itemsize
        out = np.empty(count, dtype=dtype)
        assert offset + out.nbytes <= self.nbytes
        N.check(N.load().cadrMemcpyDtoH(N.ptr(out), self.address + offset, out.nbytes, stream))
        N.check(N.load().cadrStreamSynchronize(stream))
        return out

    def fill_u32(self, value: int, stream=None, nbytes: int | None = None) -> None:
        n = self.nbytes if nbytes is None else nbytes
        N.check(N.load().cadrFillU32(self.address, value & 0xFFFFFFFF, n - n % 4, stream))

    def close(self) -> None:
        if self.address:
            N.check(N.load().cadrDeviceFree(self.address))
            self.address = 0

    def __del__(self):
        try:
            if self.address:
                N.load().cadrDeviceFree(self.address)
                self.address = 0
        except Exception:
            pass


class Stream:
    def __init__(self):
        s = ctypes.c_void_p(0)
        N.check(N.load().cadrStreamCreate(ctypes.byref(s)))
        self.handle = s.value

    def synchronize(self) -> None:
        N.check(N.load().cadrStreamSynchronize(self.handle))

    def close(self) -> None:
        if self.handle:
            N.check(N.load().cadrStreamDestroy(self.handle))
            self.handle = None


class Event:
    def __init__(self):
        e = ctypes.c_void_p(0)
        N.check(N.load().cadrEventCreate(ctypes.byref(e)))
        self.handle = e.value

    def record(self, stream=None) -> None:
        N.check(N.load().cadrEventRecord(self.handle, stream))

    def synchronize(self) -> None:
        N.check(N.load().cadrEventSynchronize(self.handle))

    def elapsed_ms(self, end: "Event") -> float:
        ms = ctypes.c_float(0)
        N.check(N.load().cadrEventElapsedTime(ctypes.byref(ms), self.handle, end.handle))
        return float(ms.value)

    def close(self) -> None:
        if self.handle:
            N.check(N.load().cadrEventDestroy(self.handle))
            self.handle = None


def time_next_launch(start: "Event | None", stop: "Event | None") -> None:
    """The calling thread's next pass launch records start / stop on its own
    kernel dispatch (cadrTimeNextLaunch): the kernel's duration with no event
    commands queued between kernels."""
    N.check(N.load().cadrTimeNextLaunch(start.handle if start else None, stop.handle if stop else None))
