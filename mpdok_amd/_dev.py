"""Device-buffer plumbing: torch tensors on the current ROCm device hold every
array; the library receives raw pointers and the current torch stream."""
import numpy as np
import torch

from . import _lib

_DT = {torch.float64: _lib.F64, torch.float32: _lib.F32}


def device():
    if not torch.cuda.is_available():
        raise RuntimeError("mpdok_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def stream():
    """Raw hipStream_t of torch's current stream."""
    return ctypes_ptr(torch.cuda.current_stream().cuda_stream)


def ctypes_ptr(v):
    return None if v is None or v == 0 else int(v)


def ptr(t):
    return None if t is None else t.data_ptr()


def dtcode(t):
    return _DT[t.dtype]


def _from_any(a, dtype):
    """Anything array-like -> torch tensor (device chosen by the caller)."""
    if isinstance(a, torch.Tensor):
        return a.to(dtype=dtype)
    if isinstance(a, np.ndarray):
        return torch.from_numpy(np.ascontiguousarray(a) if not (a.flags.c_contiguous or a.flags.f_contiguous) else a).to(dtype=dtype)
    if hasattr(a, "__dlpack__"):
        return torch.from_dlpack(a).to(dtype=dtype)
    if hasattr(a, "array"):  # _ManagedArray-like wrapper
        return _from_any(a.array, dtype)
    return torch.as_tensor(np.asarray(a), dtype=dtype)


def as_device(a, dtype=torch.float64, order="C"):
    """Return a device tensor with the values of `a`, in `dtype`, whose storage
    is C-contiguous (order='C') or Fortran-contiguous (order='F', i.e. a 2-D
    tensor with strides (1, rows)). No copy when `a` already qualifies."""
    dev = device()
    t = _from_any(a, dtype)
    if t.device != dev:
        t = t.to(dev, non_blocking=False)
    if order == "C" or t.dim() != 2:
        return t.contiguous()
    if is_fortran(t):
        return t
    return t.t().contiguous().t()


def is_fortran(t):
    return t.dim() == 2 and t.stride(0) == 1 and t.stride(1) >= max(1, t.shape[0])


def layout(t):
    """(order code, leading dimension) of a 2-D device tensor."""
    if t.dim() != 2:
        raise ValueError("expected a 2-D tensor")
    if t.stride(1) == 1 and t.stride(0) >= t.shape[1]:
        return _lib.ROW_MAJOR, t.stride(0)
    if is_fortran(t):
        return _lib.COL_MAJOR, t.stride(1)
    raise ValueError(f"unsupported strides {t.stride()} for shape {tuple(t.shape)}")


def empty(shape, dtype, order="C"):
    dev = device()
    if order == "F" and len(shape) == 2:
        return torch.empty((shape[1], shape[0]), dtype=dtype, device=dev).t()
    return torch.empty(shape, dtype=dtype, device=dev)


def zeros(shape, dtype):
    return torch.zeros(shape, dtype=dtype, device=device())


_WS = {}


def workspace(nbytes, key="default"):
    """Cached uint8 device buffer of at least `nbytes` (grown on demand).
    Calls that share a key must not run concurrently on different streams."""
    nbytes = max(int(nbytes), 256)
    k = (torch.cuda.current_device(), key)
    buf = _WS.get(k)
    if buf is None or buf.numel() < nbytes:
        buf = torch.empty(nbytes, dtype=torch.uint8, device=device())
        _WS[k] = buf
    return buf


def release_workspaces():
    _WS.clear()


def to_host_float(t):
    """Synchronising read of a one-element device tensor."""
    return float(t.item())
