"""Minimal HIP runtime plumbing over ctypes: streams and events for timing.

Only what the host mirror and the benchmark need; device memory goes through
the C-ABI (rtx_malloc / rtx_memcpy_*).
"""
import ctypes
from ctypes import POINTER, byref, c_float, c_int, c_uint, c_void_p

_hip = None


def hiplib():
    global _hip
    if _hip is None:
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so", "libamdhip64.so.7"):
            try:
                _hip = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise RuntimeError("libamdhip64.so not found")
        _hip.hipStreamCreate.argtypes = [POINTER(c_void_p)]
        _hip.hipStreamDestroy.argtypes = [c_void_p]
        _hip.hipStreamSynchronize.argtypes = [c_void_p]
        _hip.hipEventCreate.argtypes = [POINTER(c_void_p)]
        _hip.hipEventDestroy.argtypes = [c_void_p]
        _hip.hipEventRecord.argtypes = [c_void_p, c_void_p]
        _hip.hipEventSynchronize.argtypes = [c_void_p]
        _hip.hipEventElapsedTime.argtypes = [POINTER(c_float), c_void_p, c_void_p]
        _hip.hipGetErrorString.restype = ctypes.c_char_p
        _hip.hipGetErrorString.argtypes = [c_int]
        _hip.hipSetDevice.argtypes = [c_int]
        _hip.hipDeviceSynchronize.argtypes = []
        _hip.hipGetDeviceCount.argtypes = [POINTER(c_int)]
    return _hip


def _ok(e, what):
    if e != 0:
        raise RuntimeError("%s: %s" % (what, hiplib().hipGetErrorString(e).decode()))


def device_count():
    n = c_int(0)
    try:
        e = hiplib().hipGetDeviceCount(byref(n))
    except RuntimeError:
        return 0
    return n.value if e == 0 else 0


def synchronize():
    _ok(hiplib().hipDeviceSynchronize(), "hipDeviceSynchronize")


class Stream:
    """A HIP stream at the default priority."""

    def __init__(self):
        self.handle = c_void_p()
        _ok(hiplib().hipStreamCreate(byref(self.handle)), "hipStreamCreate")

    @property
    def ptr(self):
        return self.handle.value

    def synchronize(self):
        _ok(hiplib().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def destroy(self):
        if self.handle.value:
            hiplib().hipStreamDestroy(self.handle)
            self.handle = c_void_p()


class Event:
    def __init__(self):
        self.handle = c_void_p()
        _ok(hiplib().hipEventCreate(byref(self.handle)), "hipEventCreate")

    def record(self, stream=None):
        s = stream.handle if isinstance(stream, Stream) else c_void_p(stream)
        _ok(hiplib().hipEventRecord(self.handle, s), "hipEventRecord")

    def synchronize(self):
        _ok(hiplib().hipEventSynchronize(self.handle), "hipEventSynchronize")

    def elapsed_ms(self, end):
        ms = c_float(0.0)
        _ok(hiplib().hipEventElapsedTime(byref(ms), self.handle, end.handle), "hipEventElapsedTime")
        return ms.value

    def destroy(self):
        if self.handle.value:
            hiplib().hipEventDestroy(self.handle)
            self.handle = c_void_p()
