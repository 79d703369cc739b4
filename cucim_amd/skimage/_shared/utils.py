"""dtype / shape policy of the path (``_shared/utils.py`` of the reference,
paths relative to ``python/cucim/src/cucim/skimage/``)."""

from collections.abc import Iterable

import numpy as np

# ``_shared/utils.py:1131-1149``: cuCIM keeps float32 for 8/16-bit ints and bool
new_float_type = {
    "f": np.float32,
    "d": np.float64,
    "F": np.complex64,
    "D": np.complex128,
    "e": np.float32,
    "g": np.float64,
    "G": np.complex128,
    "b": np.float32,
    "B": np.float32,
    "h": np.float32,
    "H": np.float32,
    "?": np.float32,
}


def _supported_float_type(input_dtype, allow_complex=False):
    """``_shared/utils.py:1151-1179``."""
    if isinstance(input_dtype, Iterable) and not isinstance(input_dtype, str):
        return np.result_type(*(_supported_float_type(d) for d in input_dtype))
    from ..._array import np_dtype

    input_dtype = np_dtype(input_dtype)
    if not allow_complex and input_dtype.kind == "c":
        raise ValueError("complex valued input is not supported")
    return np.dtype(new_float_type.get(input_dtype.char, np.float64))


def check_nD(array, ndim, arg_name="image"):
    """``_shared/utils.py:965-988``."""
    array_ndim = array.dim() if hasattr(array, "dim") else np.asarray(array).ndim
    msg_incorrect_dim = "The parameter `%s` must be a %s-dimensional array"
    msg_empty_array = "The parameter `%s` cannot be an empty array"
    if isinstance(ndim, int):
        ndim = [ndim]
    numel = array.numel() if hasattr(array, "numel") else np.asarray(array).size
    if numel == 0:
        raise ValueError(msg_empty_array % (arg_name))
    if array_ndim not in ndim:
        raise ValueError(
            msg_incorrect_dim % (arg_name, "-or-".join([str(n) for n in ndim])))


def convert_to_float(image, preserve_range):
    """``_shared/utils.py:1019-1052``."""
    from ..._array import np_dtype
    from ..util.dtype import img_as_float

    dt = np_dtype(image.dtype)
    if dt == np.float16:
        return astype(image, np.float32)
    if preserve_range:
        if dt.char not in "df":
            image = astype(image, _supported_float_type(dt))
    else:
        image = img_as_float(image)
    return image


def astype(t, dtype):
    """Plain dtype cast on the device (C conversion), no copy if equal."""
    from ... import _array

    dt = _array.torch_dtype(dtype)
    if t.dtype == dt:
        return t
    out = _array.empty(t.shape, dt)
    _array.copy_cast(t, out)
    return out
