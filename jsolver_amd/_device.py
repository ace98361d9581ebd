"""Device plumbing: torch tensors as the array type, the current HIP stream."""
import numpy as np
import torch

from . import _lib


def device():
    if not torch.cuda.is_available():
        raise _lib.JsvError("jsolver_amd needs a ROCm GPU (gfx950); none is visible")
    return torch.device("cuda", torch.cuda.current_device())


def stream_ptr():
    return ctypes_void(torch.cuda.current_stream().cuda_stream)


def ctypes_void(v):
    import ctypes
    return ctypes.c_void_p(int(v))


def ptr(t):
    """Device pointer of a contiguous tensor (None -> NULL)."""
    import ctypes
    if t is None:
        return ctypes.c_void_p(0)
    return ctypes.c_void_p(t.data_ptr())


def is_scalar(v):
    """The reference's jnp.isscalar: Python numbers and 0-d arrays."""
    if isinstance(v, (bool, int, float, np.number)):
        return True
    if isinstance(v, np.ndarray) or torch.is_tensor(v):
        return v.ndim == 0
    return False


def dtype_name(v):
    if torch.is_tensor(v):
        return {torch.float64: "float64", torch.float32: "float32"}.get(v.dtype, str(v.dtype))
    if isinstance(v, np.ndarray) or isinstance(v, np.number):
        return str(v.dtype)
    if isinstance(v, bool):
        return "bool"
    if isinstance(v, int):
        return "int64"
    if isinstance(v, float):
        return "float64"
    if v is None:
        return "float64"
    return type(v).__name__


def shape_of(v):
    if torch.is_tensor(v) or isinstance(v, np.ndarray):
        return tuple(v.shape)
    return ()


def to_device(v, dev):
    """float64 contiguous device tensor (copy when needed; input never mutated)."""
    if torch.is_tensor(v):
        t = v.detach()
        if t.device != dev or t.dtype != torch.float64:
            t = t.to(device=dev, dtype=torch.float64)
        return t.contiguous()
    a = np.ascontiguousarray(np.asarray(v, dtype=np.float64))
    return torch.from_numpy(a).to(dev)


def scalar_value(v):
    if torch.is_tensor(v):
        return float(v.item())
    return float(v)
