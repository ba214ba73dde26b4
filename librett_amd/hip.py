"""Minimal ctypes binding of the HIP runtime (libamdhip64) for device buffers,
streams and events.

The library's own ``librett.so`` links the same ``libamdhip64.so``, so
pointers and streams made here are valid for it. PyTorch is not needed: its
wheel bundles a different HIP runtime build, and mixing two runtimes in one
process would make streams from one invalid in the other.
"""
import ctypes
import os

import numpy as np

_HIP = None

HIP_MEMCPY_H2D = 1
HIP_MEMCPY_D2H = 2
HIP_MEMCPY_D2D = 3


class HipError(RuntimeError):
    pass


def _candidates():
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    return [os.path.join(rocm, "lib", "libamdhip64.so"), "libamdhip64.so"]


def runtime():
    global _HIP
    if _HIP is None:
        err = None
        for path in _candidates():
            try:
                _HIP = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
                break
            except OSError as e:  # pragma: no cover - depends on the image
                err = e
        if _HIP is None:
            raise HipError("cannot load libamdhip64.so: %s" % err)
        H = _HIP
        vp = ctypes.c_void_p
        H.hipGetErrorString.restype = ctypes.c_char_p
        H.hipGetErrorString.argtypes = [ctypes.c_int]
        H.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
        H.hipFree.argtypes = [vp]
        H.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
        H.hipMemcpyAsync.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int, vp]
        H.hipMemset.argtypes = [vp, ctypes.c_int, ctypes.c_size_t]
        H.hipMemsetAsync.argtypes = [vp, ctypes.c_int, ctypes.c_size_t, vp]
        H.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
        H.hipStreamDestroy.argtypes = [vp]
        H.hipStreamSynchronize.argtypes = [vp]
        H.hipDeviceSynchronize.argtypes = []
        H.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
        H.hipEventDestroy.argtypes = [vp]
        H.hipEventRecord.argtypes = [vp, vp]
        H.hipEventSynchronize.argtypes = [vp]
        H.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
        H.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
        H.hipSetDevice.argtypes = [ctypes.c_int]
        H.hipGetLastError.argtypes = []
        H.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
        H.hipDeviceGetAttribute.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int]
    return _HIP


def check(rc, what="hip call"):
    if rc != 0:
        msg = runtime().hipGetErrorString(rc)
        raise HipError("%s failed: %d (%s)" % (what, rc, msg.decode() if msg else "?"))


def device_count():
    n = ctypes.c_int(0)
    rc = runtime().hipGetDeviceCount(ctypes.byref(n))
    if rc != 0:
        runtime().hipGetLastError()
        return 0
    return n.value


def malloc(nbytes):
    p = ctypes.c_void_p()
    check(runtime().hipMalloc(ctypes.byref(p), max(1, int(nbytes))), "hipMalloc(%d)" % nbytes)
    return p.value


def free(ptr):
    if ptr:
        check(runtime().hipFree(ctypes.c_void_p(ptr)), "hipFree")


def htod(dst, arr):
    arr = np.ascontiguousarray(arr)
    check(runtime().hipMemcpy(ctypes.c_void_p(dst), ctypes.c_void_p(arr.ctypes.data), arr.nbytes,
                              HIP_MEMCPY_H2D), "hipMemcpy H2D")


def dtoh(arr, src):
    assert arr.flags["C_CONTIGUOUS"]
    check(runtime().hipMemcpy(ctypes.c_void_p(arr.ctypes.data), ctypes.c_void_p(src), arr.nbytes,
                              HIP_MEMCPY_D2H), "hipMemcpy D2H")
    return arr


def memset(ptr, value, nbytes, stream=None):
    if stream is None:
        check(runtime().hipMemset(ctypes.c_void_p(ptr), value, nbytes), "hipMemset")
    else:
        check(runtime().hipMemsetAsync(ctypes.c_void_p(ptr), value, nbytes, ctypes.c_void_p(stream)),
              "hipMemsetAsync")


def memcpy_d2d_async(dst, src, nbytes, stream):
    check(runtime().hipMemcpyAsync(ctypes.c_void_p(dst), ctypes.c_void_p(src), nbytes, HIP_MEMCPY_D2D,
                                   ctypes.c_void_p(stream)), "hipMemcpyAsync D2D")


def stream_create():
    s = ctypes.c_void_p()
    check(runtime().hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
    return s.value


def stream_destroy(s):
    if s:
        check(runtime().hipStreamDestroy(ctypes.c_void_p(s)), "hipStreamDestroy")


def stream_sync(s):
    check(runtime().hipStreamSynchronize(ctypes.c_void_p(s)), "hipStreamSynchronize")


def device_sync():
    check(runtime().hipDeviceSynchronize(), "hipDeviceSynchronize")


class Event:
    def __init__(self):
        e = ctypes.c_void_p()
        check(runtime().hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        self.ptr = e.value

    def record(self, stream):
        check(runtime().hipEventRecord(ctypes.c_void_p(self.ptr), ctypes.c_void_p(stream)), "hipEventRecord")

    def synchronize(self):
        check(runtime().hipEventSynchronize(ctypes.c_void_p(self.ptr)), "hipEventSynchronize")

    def elapsed_ms(self, start):
        ms = ctypes.c_float()
        check(runtime().hipEventElapsedTime(ctypes.byref(ms), ctypes.c_void_p(start.ptr),
                                            ctypes.c_void_p(self.ptr)), "hipEventElapsedTime")
        return ms.value

    def destroy(self):
        if self.ptr:
            runtime().hipEventDestroy(ctypes.c_void_p(self.ptr))
            self.ptr = None


class DeviceBuffer:
    """Owned device allocation."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = malloc(self.nbytes)

    def upload(self, arr):
        assert np.ascontiguousarray(arr).nbytes <= self.nbytes
        htod(self.ptr, arr)

    def download(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        return dtoh(out, self.ptr)

    def free(self):
        if self.ptr:
            free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
