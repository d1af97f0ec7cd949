"""Device-memory plumbing for the drop-in functions (PyTorch-ROCm tensors).

Drop-ins accept NumPy arrays (the reference's types) or torch tensors already
on the GPU.  NumPy inputs are copied to HBM, results copied back; tensor
inputs stay on the device and results are returned as tensors.
"""
import numpy as np
import torch

from . import _lib

DEVICE = torch.device("cuda", 0)

_TORCH = {np.dtype(np.float64): torch.float64, np.dtype(np.int32): torch.int32,
          np.dtype(np.uint8): torch.uint8, np.dtype(np.bool_): torch.uint8,
          np.dtype(np.float32): torch.float32}


def require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("meeuuw_amd needs a ROCm GPU (no CPU fallback)")


def stream():
    return ctypes_stream(torch.cuda.current_stream())


def ctypes_stream(s):
    return s.cuda_stream


def to_dev(a, dtype):
    """NumPy array or tensor -> contiguous device tensor of the given numpy dtype."""
    dt = np.dtype(dtype)
    if isinstance(a, torch.Tensor):
        t = a.to(device=DEVICE)
        want = _TORCH[dt]
        if t.dtype == torch.bool:
            t = t.to(torch.uint8)
        if t.dtype != want:
            t = t.to(want)
        return t.contiguous()
    arr = np.asarray(a)
    if dt == np.bool_ or dt == np.uint8:
        arr = np.ascontiguousarray(arr.astype(bool, copy=False)).view(np.uint8)
    else:
        arr = np.ascontiguousarray(arr, dtype=dt)
    return torch.from_numpy(arr).to(DEVICE, non_blocking=False)


def empty(n, dtype=np.float64):
    return torch.empty(int(n), dtype=_TORCH[np.dtype(dtype)], device=DEVICE)


def zeros(n, dtype=np.float64):
    return torch.zeros(int(n), dtype=_TORCH[np.dtype(dtype)], device=DEVICE)


def ptr(t):
    return None if t is None else t.data_ptr()


def host(t, like_numpy=True):
    torch.cuda.current_stream().synchronize()
    return t.cpu().numpy() if like_numpy else t


def is_dev(*arrays):
    return any(isinstance(a, torch.Tensor) for a in arrays)


def out(t, numpy_out, shape=None, dtype=None):
    """Return t as numpy (copy back) or as the device tensor."""
    if shape is not None:
        t = t.reshape(shape)
    if not numpy_out:
        return t
    a = host(t)
    if dtype is not None:
        a = a.astype(dtype, copy=False)
    return a


class Incidence:
    """node -> sorted (element, local index) incidence of a connectivity."""

    def __init__(self, icon_dev, nn):
        m, nel = icon_dev.shape
        self.nn, self.m, self.nel = int(nn), int(m), int(nel)
        self.ptr = empty(nn + 1, np.int32)
        self.list = empty(m * nel, np.int32)
        work = empty((_lib.lib.mw_incidence_workspace_bytes(int(nn)) + 7) // 8, np.float64)
        _lib.call("mw_incidence_build", self.nel, self.m, ptr(icon_dev), self.nn, ptr(self.ptr),
                  ptr(self.list), ptr(work), stream())


_INC_CACHE = {}


def incidence(icon_dev, nn, key=None):
    """Incidence for a device connectivity; cached per (key, shape, nn) when key given."""
    if key is None:
        return Incidence(icon_dev, nn)
    k = (key, tuple(icon_dev.shape), int(nn))
    inc = _INC_CACHE.get(k)
    if inc is None:
        if len(_INC_CACHE) > 16:
            _INC_CACHE.clear()
        inc = Incidence(icon_dev, nn)
        _INC_CACHE[k] = inc
    return inc


def array_key(a):
    """Content key of a NumPy connectivity (shape + hash of the bytes); None for
    device tensors, whose memory may be reused, so their incidence is rebuilt."""
    if isinstance(a, torch.Tensor):
        return None
    arr = np.ascontiguousarray(a)
    return ("n", arr.shape, arr.dtype.str, hash(arr.tobytes()))
