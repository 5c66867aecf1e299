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
    """A contiguous FP64 vector in device memory."""

    __slots__ = ("n", "_p", "__weakref__")

    def __init__(self, n: int):
        _ensure_device()
        self.n = int(n)
        p = C.c_void_p()
        _abi.call("vtsa_device_alloc", C.c_size_t(8 * max(self.n, 1)), C.byref(p))
        self._p = p

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
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        if a.size != self.n:
            raise ValueError(f"size mismatch {a.size} != {self.n}")
        _abi.call("vtsa_copy_h2d", self._p, a.ctypes.data_as(C.c_void_p), C.c_size_t(8 * self.n), None)
        return self

    def numpy(self):
        out = np.empty(self.n)
        _abi.call("vtsa_copy_d2h", out.ctypes.data_as(C.c_void_p), self._p, C.c_size_t(8 * self.n), None)
        return out

    def copy_from(self, other: "DeviceArray"):
        if other.n != self.n:
            raise ValueError("size mismatch")
        _abi.call("vtsa_copy_d2d", self._p, other._p, C.c_size_t(8 * self.n), None)
        return self

    def clone(self) -> "DeviceArray":
        return DeviceArray(self.n).copy_from(self)

    def fill(self, v: float):
        _abi.call("vtsa_fill", self._p, float(v), C.c_size_t(self.n), None)
        return self


def empty(n):
    return DeviceArray(n)


def zeros(n):
    return DeviceArray(n).fill(0.0)


def to_dev(a):
    a = np.asarray(a, dtype=np.float64)
    return DeviceArray(a.size).upload(a)


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
