"""Integer -> float image conversion (``util/dtype.py:219-517`` of the
reference, paths relative to ``python/cucim/src/cucim/skimage/``)."""

import numpy as np

from ... import _array, _lib
from .._shared.utils import _supported_float_type, astype


def _to_float(image, dtype):
    """``_convert`` int/bool/float -> float: uint scales by 1/imax, int also
    clips at -1, computed in the first of (dtype, float32, float64) wide
    enough for the input."""
    x = _array.as_tensor(image)
    dt = np.dtype(dtype)
    src = _array.np_dtype(x.dtype)
    if src == dt:
        return x
    if src.kind in "fb":
        return astype(x, dt)
    xc = x.contiguous()
    out = _array.empty(x.shape, dt)
    _lib.check(_lib.lib().hc_to_float(_array.hc_array(xc), _array.hc_array(out),
                                      _array.stream()), "img_as_float")
    return out


def img_as_float32(image, force_copy=False):
    out = _to_float(image, np.float32)
    return out.clone() if (force_copy and out is image) else out


def img_as_float64(image, force_copy=False):
    out = _to_float(image, np.float64)
    return out.clone() if (force_copy and out is image) else out


def img_as_float(image, force_copy=False):
    """float16/float32 and 8/16-bit ints -> float32, others -> float64."""
    x = _array.as_tensor(image)
    out = _to_float(x, _supported_float_type(x.dtype))
    return out.clone() if (force_copy and out is x) else out
