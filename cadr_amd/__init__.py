"""cadr_amd — MI355X-native drawable-processing pass of Rendering-FIT/CADR.

The product is libcadr_hip.so: hand-written HIP kernels for gfx950 behind the
C ABI of include/cadr_hip.h, plus the C++ CadR facade (include/CadR/).  This
Python package is a thin ctypes driver used by the tests and the benchmark.
"""
from . import _native
from ._native import CadrError, load
from .device import DeviceBuffer, Event, Stream, compute_units, device_count, set_device, synchronize

__all__ = ["CadrError", "DeviceBuffer", "Event", "Stream", "compute_units", "device_count", "load",
           "set_device", "synchronize", "_native"]
