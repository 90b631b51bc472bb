"""Device buffers for the device-pointer entry points (pikgpu_fwd_adj with
PIKGPU_ADJ_DEVICE_POINTERS), allocated through the HIP runtime that
libpikgpu.so itself links, so a process holds one HIP runtime.

A DeviceArray owns one hipMalloc allocation; indexing its first axis gives
a non-owning view of a plane. `data_ptr()` is the device address.
"""
import ctypes

import numpy as np

from . import abi

_H2D, _D2H = 1, 2
_hip = None


def hip():
    """The HIP runtime libpikgpu.so is linked against (already loaded by
    it; dlopen by soname returns that copy)."""
    global _hip
    if _hip is None:
        abi.load()
        L = ctypes.CDLL("libamdhip64.so.7")
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        L.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
        L.hipFree.argtypes = [vp]
        L.hipMemcpy.argtypes = [vp, vp, sz, ctypes.c_int]
        L.hipMemset.argtypes = [vp, ctypes.c_int, sz]
        L.hipSetDevice.argtypes = [ctypes.c_int]
        L.hipHostMalloc.argtypes = [ctypes.POINTER(vp), sz, ctypes.c_uint]
        L.hipHostFree.argtypes = [vp]
        L.hipGetErrorString.restype = ctypes.c_char_p
        L.hipGetErrorString.argtypes = [ctypes.c_int]
        _hip = L
    return _hip


def _check(rc, what):
    if rc != 0:
        raise RuntimeError("%s: %s" % (what, hip().hipGetErrorString(rc)
                                       .decode()))


class DeviceArray:
    def __init__(self, shape, dtype, device=0, _base=None, _ptr=None):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self._base = _base
        if _ptr is not None:
            self._ptr = _ptr
            self._owned = False
            return
        L = hip()
        _check(L.hipSetDevice(device), "hipSetDevice")
        p = ctypes.c_void_p()
        _check(L.hipMalloc(ctypes.byref(p), max(self.nbytes, 4)), "hipMalloc")
        self._ptr = p.value
        self._owned = True

    @classmethod
    def from_numpy(cls, a, device=0):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype, device)
        _check(hip().hipMemcpy(d._ptr, a.ctypes.data, a.nbytes, _H2D),
               "hipMemcpy")
        return d

    @classmethod
    def zeros(cls, shape, dtype, device=0):
        d = cls(shape, dtype, device)
        _check(hip().hipMemset(d._ptr, 0, d.nbytes), "hipMemset")
        return d

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        _check(hip().hipMemcpy(out.ctypes.data, self._ptr, self.nbytes, _D2H),
               "hipMemcpy")
        return out

    def data_ptr(self):
        return self._ptr

    def __getitem__(self, i):
        if not 0 <= i < self.shape[0]:
            raise IndexError(i)
        sub = self.shape[1:]
        step = int(np.prod(sub)) * self.dtype.itemsize
        return DeviceArray(sub, self.dtype, _base=self,
                           _ptr=self._ptr + i * step)

    def free(self):
        if getattr(self, "_owned", False) and self._ptr:
            hip().hipFree(self._ptr)
        self._ptr = None
        self._owned = False

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class PinnedArray:
    """Page-locked host memory (hipHostMalloc) viewed as a numpy array
    (`.array`): host <-> device copies from it run at the link's rate
    instead of through the runtime's staging of pageable memory."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        nbytes = max(int(np.prod(self.shape)) * self.dtype.itemsize, 4)
        p = ctypes.c_void_p()
        _check(hip().hipHostMalloc(ctypes.byref(p), nbytes, 0),
               "hipHostMalloc")
        self._ptr = p.value
        buf = (ctypes.c_char * nbytes).from_address(self._ptr)
        self.array = np.frombuffer(buf, self.dtype,
                                   int(np.prod(self.shape))).reshape(
                                       self.shape)

    @classmethod
    def copy_of(cls, a):
        a = np.asarray(a)
        p = cls(a.shape, a.dtype)
        p.array[...] = a
        return p

    def free(self):
        if self._ptr:
            self.array = None
            hip().hipHostFree(self._ptr)
        self._ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
