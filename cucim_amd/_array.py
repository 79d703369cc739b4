"""Array plumbing: DLPack intake, dtype codes, output allocation.

Device arrays are ``torch.Tensor`` objects on the current ROCm device; any
object exporting ``__dlpack__`` (CuPy arrays, JAX arrays, ...) is taken
through ``torch.from_dlpack`` without a copy.  Host numpy arrays are copied to
the device.  The kernels run on torch's current stream.
"""

import numpy as np
import torch

from . import _lib

_TORCH_TO_HC = {
    torch.bool: _lib.BOOL,
    torch.int8: _lib.INT8,
    torch.uint8: _lib.UINT8,
    torch.int16: _lib.INT16,
    torch.uint16: _lib.UINT16,
    torch.int32: _lib.INT32,
    torch.uint32: _lib.UINT32,
    torch.int64: _lib.INT64,
    torch.uint64: _lib.UINT64,
    torch.float16: _lib.FLOAT16,
    torch.float32: _lib.FLOAT32,
    torch.float64: _lib.FLOAT64,
    torch.bfloat16: _lib.BFLOAT16,
}

_NP_TO_TORCH = {
    np.dtype(np.bool_): torch.bool,
    np.dtype(np.int8): torch.int8,
    np.dtype(np.uint8): torch.uint8,
    np.dtype(np.int16): torch.int16,
    np.dtype(np.uint16): torch.uint16,
    np.dtype(np.int32): torch.int32,
    np.dtype(np.uint32): torch.uint32,
    np.dtype(np.int64): torch.int64,
    np.dtype(np.uint64): torch.uint64,
    np.dtype(np.float16): torch.float16,
    np.dtype(np.float32): torch.float32,
    np.dtype(np.float64): torch.float64,
}
_TORCH_TO_NP = {v: k for k, v in _NP_TO_TORCH.items()}


def np_dtype(dt):
    """numpy dtype of a torch dtype / numpy dtype / string / type."""
    if isinstance(dt, torch.dtype):
        if dt in (torch.complex64, torch.complex128):
            return np.dtype(np.complex64 if dt == torch.complex64 else np.complex128)
        return _TORCH_TO_NP[dt]
    return np.dtype(dt)


def torch_dtype(dt):
    if isinstance(dt, torch.dtype):
        return dt
    d = np.dtype(dt)
    if d.kind == "c":
        raise NotImplementedError("complex dtypes are not supported by cucim_amd")
    try:
        return _NP_TO_TORCH[d]
    except KeyError:
        raise TypeError(f"unsupported dtype {d}")


def device():
    return torch.device("cuda", torch.cuda.current_device())


def as_tensor(x):
    """Device tensor view of ``x`` (no copy for device arrays)."""
    if isinstance(x, torch.Tensor):
        t = x
    elif hasattr(x, "__dlpack__"):
        t = torch.from_dlpack(x)
    else:
        t = torch.from_numpy(np.ascontiguousarray(np.asarray(x)))
    if t.is_complex():
        raise NotImplementedError("complex inputs are not supported by cucim_amd")
    if t.device.type != "cuda":
        t = t.to(device())
    return t


def hc_array(t):
    """``hc_array`` descriptor of a tensor (element strides)."""
    if t.dtype not in _TORCH_TO_HC:
        raise TypeError(f"unsupported dtype {t.dtype}")
    if t.dim() > _lib.MAX_NDIM:
        raise ValueError(f"at most {_lib.MAX_NDIM} dimensions are supported")
    a = _lib.HCArray()
    a.data = t.data_ptr()
    a.dtype = _TORCH_TO_HC[t.dtype]
    a.ndim = t.dim()
    for d in range(t.dim()):
        a.shape[d] = t.shape[d]
        a.strides[d] = t.stride(d)
    return a


def hc_arrays(ts):
    arr = (_lib.HCArray * len(ts))()
    for i, t in enumerate(ts):
        arr[i] = hc_array(t)
    return arr


def stream():
    return torch.cuda.current_stream().cuda_stream


def empty(shape, dtype):
    return torch.empty(tuple(shape), dtype=torch_dtype(dtype), device=device())


def shares_memory(a, b):
    """Conservative overlap test of two tensors' storage ranges."""
    if a.numel() == 0 or b.numel() == 0:
        return False
    sa, sb = a.untyped_storage(), b.untyped_storage()
    if sa.data_ptr() != sb.data_ptr():
        # distinct storages may still overlap through views of raw pointers
        pass
    def extent(t):
        lo = t.data_ptr()
        span = sum((s - 1) * abs(st) for s, st in zip(t.shape, t.stride())) + 1
        return lo, lo + span * t.element_size()
    a0, a1 = extent(a)
    b0, b1 = extent(b)
    return a0 < b1 and b0 < a1


def copy_cast(src, dst):
    """dst[...] = cast<dst dtype>(src) with the reference's cast rules."""
    lib = _lib.lib()
    a, b = hc_array(src), hc_array(dst)
    _lib.check(lib.hc_copy_cast(a, b, stream()), "copy")
    return dst


def get_output(output, inp, shape=None):
    """``_vendored/_ndimage_util.py:49-71``: None -> input dtype; a dtype or
    dtype string -> new array; an array -> shape check."""
    shape = tuple(inp.shape) if shape is None else tuple(shape)
    if output is None:
        return empty(shape, inp.dtype)
    if isinstance(output, (type, np.dtype, str, torch.dtype)):
        dt = torch_dtype(output) if not isinstance(output, torch.dtype) else output
        if np_dtype(dt).kind == "c":
            raise NotImplementedError("complex outputs are not supported by cucim_amd")
        return empty(shape, dt)
    out = as_tensor(output)
    if tuple(out.shape) != shape:
        raise RuntimeError("output shape not correct")
    return out


def binary(a, b, out, op):
    """out = a - b (``_lib.BIN_SUB``), a * b (``_lib.BIN_MUL``) or a + b
    (``_lib.BIN_ADD``): one rounding per element in the arrays' float dtype,
    wrapping for integer dtypes; out may be a or b."""
    _lib.check(_lib.lib().hc_binary(hc_array(a), hc_array(b), hc_array(out), op, stream()),
               "elementwise op")
    return out
