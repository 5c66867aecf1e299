"""Device vectors owned through the libvtsa C ABI (HBM allocation, copies, streams).

The package drives the GPU only through libvtsa: it needs no other HIP runtime in the process
(PyTorch-ROCm ships its own, which must not be mixed with the one libvtsa links).

Streams: every call takes the calling thread's current stream (``stream()``), the null stream
unless ``use_stream`` set another one, so that independent cases driven from different threads
run concurrently.  Allocation: freed arrays go to a cache keyed by (stream, bytes) and are
reused by later allocations on the same stream, which keeps hipMalloc/hipFree (and the device
synchronisation of hipFree) out of the integrator's loops; reuse on the same stream is ordered
after every earlier use by the stream itself.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import threading

import numpy as np

from . import _abi

_checked = False
_tls = threading.local()
_cache = {}                      # (stream value, nbytes) -> [pointer values]
# Reentrant: DeviceArray.__del__ takes it, and a garbage collection triggered inside a critical
# section can run another DeviceArray's __del__ on the same thread.
_cache_lock = threading.RLock()


def _ensure_device():
    global _checked
    if not _checked:
        cnt = C.c_int32(0)
        rc = _abi.lib().vtsa_device_count(C.byref(cnt))
        if rc != 0 or cnt.value < 1:
            raise RuntimeError("no HIP device is visible: the dac_vtsa_amd compute path needs an MI355X GPU")
        _checked = True


def stream():
    """The calling thread's current stream (None = the null stream)."""
    return getattr(_tls, "stream", None)


@contextlib.contextmanager
def use_stream(s):
    """Run the enclosed calls of this thread on stream s."""
    prev = stream()
    _tls.stream = s
    try:
        yield s
    finally:
        _tls.stream = prev


def new_stream():
    _ensure_device()
    s = C.c_void_p()
    _abi.call("vtsa_stream_create", C.byref(s))
    return s


def free_stream(s):
    """Synchronise and destroy a stream; its cached buffers are freed."""
    if s is None or not s.value:
        return
    _abi.call("vtsa_stream_destroy", s)
    release_cache(s)


def release_cache(s=None):
    """Free the cached buffers of stream s (all streams when s is None)."""
    key = None if s is None else (s.value if isinstance(s, C.c_void_p) else s)
    with _cache_lock:
        keys = [k for k in _cache if s is None or k[0] == key]
        blocks = [p for k in keys for p in _cache.pop(k)]
    if blocks:
        _abi.call("vtsa_sync", None)
        for p in blocks:
            _abi.lib().vtsa_device_free(C.c_void_p(p))


def _stream_key():
    s = stream()
    return s.value if isinstance(s, C.c_void_p) else s


class DeviceArray:
    """A contiguous vector in device memory (FP64 unless another dtype is given)."""

    __slots__ = ("n", "dtype", "_p", "_key", "__weakref__")

    def __init__(self, n: int, dtype=np.float64):
        _ensure_device()
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        nb = self.dtype.itemsize * max(self.n, 1)
        key = (_stream_key(), nb)
        with _cache_lock:
            lst = _cache.get(key)
            pv = lst.pop() if lst else None
        if pv is None:
            p = C.c_void_p()
            _abi.call("vtsa_device_alloc", C.c_size_t(nb), C.byref(p))
        else:
            p = C.c_void_p(pv)
        self._p = p
        self._key = key

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
                pv, fresh = p.value, []          # allocate before taking the lock
                with _cache_lock:
                    _cache.setdefault(self._key, fresh).append(pv)
            except Exception:
                pass
            self._p = None

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype).ravel()
        if a.size != self.n:
            raise ValueError(f"size mismatch {a.size} != {self.n}")
        _abi.call("vtsa_copy_h2d", self._p, a.ctypes.data_as(C.c_void_p), C.c_size_t(self.nbytes), stream())
        return self

    def numpy(self):
        out = np.empty(self.n, dtype=self.dtype)
        _abi.call("vtsa_copy_d2h", out.ctypes.data_as(C.c_void_p), self._p, C.c_size_t(self.nbytes), stream())
        return out

    def copy_from(self, other: "DeviceArray"):
        if other.n != self.n or other.dtype != self.dtype:
            raise ValueError("size mismatch")
        _abi.call("vtsa_copy_d2d", self._p, other._p, C.c_size_t(self.nbytes), stream())
        return self

    def clone(self) -> "DeviceArray":
        return DeviceArray(self.n, self.dtype).copy_from(self)

    # a copy owns memory of its own (two owners of one pointer would both return it to the cache)
    def __copy__(self):
        return self.clone()

    def __deepcopy__(self, memo):
        return self.clone()

    def view(self, offset: int, n: int) -> "DeviceView":
        """A window [offset, offset + n) of this array (shares its memory)."""
        if offset < 0 or n < 0 or offset + n > self.n:
            raise ValueError("view out of range")
        return DeviceView(self, offset, n)

    def fill(self, v: float):
        if self.dtype != np.float64:
            raise TypeError("fill is FP64 only")
        _abi.call("vtsa_fill", self._p, float(v), C.c_size_t(self.n), stream())
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
                  C.c_size_t(self.n * self.dtype.itemsize), stream())
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


def synchronize():
    """Wait for the calling thread's current stream."""
    _abi.call("vtsa_sync", stream())
