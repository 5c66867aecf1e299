"""Device buffers and streams (PyTorch-ROCm is used as the allocator and stream source only)."""
from __future__ import annotations

import ctypes as C

import numpy as np

_torch = None


def torch():
    global _torch
    if _torch is None:
        import torch as t
        if not t.cuda.is_available():
            raise RuntimeError("no HIP device is visible: the dac_vtsa_amd compute path needs an MI355X GPU")
        _torch = t
    return _torch


def empty(n, dtype=np.float64):
    t = torch()
    return t.empty(int(n), dtype=t.float64 if dtype == np.float64 else t.float32, device="cuda")


def zeros(n):
    t = torch()
    return t.zeros(int(n), dtype=t.float64, device="cuda")


def to_dev(a):
    t = torch()
    return t.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to("cuda")


def to_host(x):
    return x.detach().to("cpu").numpy()


def ptr(x):
    """Device address of a tensor (or None)."""
    return None if x is None else C.c_void_p(x.data_ptr())


def stream():
    """The current HIP stream as a raw handle for the C ABI."""
    return C.c_void_p(torch().cuda.current_stream().cuda_stream)
