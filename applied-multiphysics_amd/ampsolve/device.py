"""Device (HBM) arrays for the host layer, allocated through the same HIP runtime that libampsolve.so links.

The library's *_device entry points take plain device pointers. Callers that keep inputs resident in HBM (the bench,
the batched and sweep drivers, the tests) hold them in DeviceArray: a typed, shaped buffer from hipMalloc with
explicit host copies. Going through the library's own runtime keeps one HIP / HSA runtime per process (a second copy,
such as the one a framework bundles, would be initialised and torn down beside it).
"""
import ctypes as C

import numpy as np

from . import _capi as A

_H2D, _D2H, _D2D = 1, 2, 3
_hip = None


def hip():
    """The HIP runtime of libampsolve.so (resolved by soname after the library is loaded)."""
    global _hip
    if _hip is None:
        A.lib()
        h = C.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        h.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
        h.hipDeviceSynchronize.argtypes = []
        h.hipGetDeviceCount.argtypes = [C.POINTER(C.c_int)]
        h.hipGetErrorString.argtypes = [C.c_int]
        h.hipGetErrorString.restype = C.c_char_p
        for f in ("hipMalloc", "hipFree", "hipMemcpy", "hipMemcpyAsync", "hipMemset", "hipDeviceSynchronize", "hipGetDeviceCount"):
            getattr(h, f).restype = C.c_int
        _hip = h
    return _hip


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what}: {hip().hipGetErrorString(rc).decode()}")


def device_count():
    n = C.c_int(0)
    rc = hip().hipGetDeviceCount(C.byref(n))
    return n.value if rc == 0 else 0


def available():
    try:
        return device_count() > 0
    except (OSError, RuntimeError):
        return False


def synchronize():
    _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceArray:
    """A C-contiguous array in HBM. Indexing the first axis with an integer or a slice gives a view (no copy)."""

    is_cuda = True  # duck type of a device tensor for the host layer

    def __init__(self, shape, dtype=np.float64, _base=None, _ptr=None):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.size = int(np.prod(self.shape)) if self.shape else 1
        self.nbytes = self.size * self.dtype.itemsize
        self._base = _base
        if _ptr is not None:
            self._p = _ptr
        else:
            p = C.c_void_p()
            _ok(hip().hipMalloc(C.byref(p), max(self.nbytes, 256)), "hipMalloc")
            self._p = p.value

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    @classmethod
    def zeros(cls, shape, dtype=np.float64):
        d = cls(shape, dtype)
        _ok(hip().hipMemset(C.c_void_p(d._p), 0, d.nbytes), "hipMemset")
        return d

    def data_ptr(self):
        return self._p

    def numel(self):
        return self.size

    def __len__(self):
        return self.shape[0] if self.shape else 1

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.size != self.size:
            raise ValueError(f"upload: {a.size} entries into a device array of {self.size}")
        if self.nbytes:
            _ok(hip().hipMemcpy(C.c_void_p(self._p), a.ctypes.data_as(C.c_void_p), self.nbytes, _H2D), "upload")
        return self

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            _ok(hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self._p), self.nbytes, _D2H), "download")
        return out

    cpu = numpy

    def copy_(self, other):
        if other.nbytes != self.nbytes:
            raise ValueError("copy_: sizes differ")
        if self.nbytes:
            _ok(hip().hipMemcpy(C.c_void_p(self._p), C.c_void_p(other._p), self.nbytes, _D2D), "copy")
        return self

    def copy_async_(self, other):
        """Device-to-device copy queued on the null stream (returns before it completes)."""
        if other.nbytes != self.nbytes:
            raise ValueError("copy_async_: sizes differ")
        if self.nbytes:
            _ok(hip().hipMemcpyAsync(C.c_void_p(self._p), C.c_void_p(other._p), self.nbytes, _D2D, None), "copy")
        return self

    def new_empty(self, shape=None):
        return DeviceArray(self.shape if shape is None else shape, self.dtype)

    def clone(self):
        return DeviceArray(self.shape, self.dtype).copy_(self)

    def zero_(self):
        _ok(hip().hipMemset(C.c_void_p(self._p), 0, self.nbytes), "hipMemset")
        return self

    def __getitem__(self, k):
        if not self.shape:
            raise IndexError("scalar device array")
        row = self.size // self.shape[0] if self.shape[0] else 0
        es = self.dtype.itemsize
        if isinstance(k, (int, np.integer)):
            k = int(k) + (self.shape[0] if k < 0 else 0)
            if not 0 <= k < self.shape[0]:
                raise IndexError(k)
            return DeviceArray(self.shape[1:], self.dtype, _base=self, _ptr=self._p + k * row * es)
        if isinstance(k, slice):
            a, b, st = k.indices(self.shape[0])
            if st != 1:
                raise IndexError("device arrays take contiguous slices only")
            return DeviceArray((max(b - a, 0),) + self.shape[1:], self.dtype, _base=self, _ptr=self._p + a * row * es)
        raise IndexError("device arrays take an integer or a contiguous slice")

    def view(self, shape):
        if int(np.prod(shape)) != self.size:
            raise ValueError("view: sizes differ")
        return DeviceArray(shape, self.dtype, _base=self, _ptr=self._p)

    def __del__(self):
        if getattr(self, "_base", None) is None and getattr(self, "_p", None) and _hip is not None:
            try:
                _hip.hipFree(C.c_void_p(self._p))
            except Exception:
                pass
            self._p = None
