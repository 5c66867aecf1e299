"""Device vectors owned through the libvtsa C ABI (HBM allocation, copies, default stream).

The package drives the GPU only through libvtsa: it needs no other HIP runtime in the process
(PyTorch-ROCm ships its own, which must not be mixed with the one libvtsa links).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi

_checked = False


def _ensure_device():
    global _checked
    if not _checked:
        cnt = C.c_int32(0)
        rc = _abi.lib().vtsa_device_count(C.byref(cnt))
        if rc != 0 or cnt.value < 1:
            raise RuntimeError("no HIP device is visible: the dac_vtsa_amd compute path needs an MI355X GPU")
        _checked = True


class DeviceArray:
    """A contiguous vector in device memory (FP64 unless another dtype is given)."""

    __slots__ = ("n", "dtype", "_p", "__weakref__")

    def __init__(self, n: int, dtype=np.float64):
        _ensure_device()
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        p = C.c_void_p()
        _abi.call("vtsa_device_alloc", C.c_size_t(self.dtype.itemsize * max(self.n, 1)), C.byref(p))
        self._p = p

    @property
    def nbytes(self):
        return self.dtype.itemsize * self.n

    @property
    def ptr(self):
        return self._p

    def __len__(self):
        return self.n

    def __del__(self):
        p = getattr(self, "_p", None)
        if p is not None and p.value:
            try:
                _abi.lib().vtsa_device_free(p)
            except Exception:
                pass
            self._p = None

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype).ravel()
        if a.size != self.n:
            raise ValueError(f"size mismatch {a.size} != {self.n}")
        _abi.call("vtsa_copy_h2d", self._p, a.ctypes.data_as(C.c_void_p), C.c_size_t(self.nbytes), None)
        return self

    def numpy(self):
        out = np.empty(self.n, dtype=self.dtype)
        _abi.call("vtsa_copy_d2h", out.ctypes.data_as(C.c_void_p), self._p, C.c_size_t(self.nbytes), None)
        return out

    def copy_from(self, other: "DeviceArray"):
        if other.n != self.n or other.dtype != self.dtype:
            raise ValueError("size mismatch")
        _abi.call("vtsa_copy_d2d", self._p, other._p, C.c_size_t(self.nbytes), None)
        return self

    def clone(self) -> "DeviceArray":
        return DeviceArray(self.n, self.dtype).copy_from(self)

    def view(self, offset: int, n: int) -> "DeviceView":
        """A window [offset, offset + n) of this array (shares its memory)."""
        if offset < 0 or n < 0 or offset + n > self.n:
            raise ValueError("view out of range")
        return DeviceView(self, offset, n)

    def fill(self, v: float):
        if self.dtype != np.float64:
            raise TypeError("fill is FP64 only")
        _abi.call("vtsa_fill", self._p, float(v), C.c_size_t(self.n), None)
        return self


class DeviceView:
    """A window of a DeviceArray; keeps its parent alive."""

    __slots__ = ("base", "n", "dtype", "_p")

    def __init__(self, base: DeviceArray, offset: int, n: int):
        self.base = base
        self.n = n
        self.dtype = base.dtype
        self._p = C.c_void_p(int(base.ptr.value) + int(offset) * base.dtype.itemsize)

    @property
    def ptr(self):
        return self._p

    def numpy(self):
        out = np.empty(self.n, dtype=self.dtype)
        _abi.call("vtsa_copy_d2h", out.ctypes.data_as(C.c_void_p), self._p,
                  C.c_size_t(self.n * self.dtype.itemsize), None)
        return out


def empty(n, dtype=np.float64):
    return DeviceArray(n, dtype)


def zeros(n):
    return DeviceArray(n).fill(0.0)


def to_dev(a, dtype=np.float64):
    a = np.asarray(a, dtype=dtype)
    return DeviceArray(a.size, dtype).upload(a)


def to_host(x):
    return x.numpy()


def ptr(x):
    """Device address of an array (or None)."""
    return None if x is None else x.ptr


def stream():
    """All work runs on the default stream (NULL)."""
    return None


def synchronize():
    _abi.call("vtsa_sync", None)
