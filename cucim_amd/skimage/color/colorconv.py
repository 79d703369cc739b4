"""``rgb2gray`` (``color/colorconv.py:1037-1095`` of the reference, paths
relative to ``python/cucim/src/cucim/skimage/``)."""

import numpy as np

from ... import _array, _lib
from .._shared.utils import _supported_float_type


def rgb2gray(rgb, *, channel_axis=-1):
    """Luminance Y = 0.2125 R + 0.7154 G + 0.0721 B.  8/16-bit integer input
    gives float32 (cuCIM's dtype policy), float64 input float64."""
    x = _array.as_tensor(rgb)
    if x.dim() < 2:
        raise ValueError("Input image expected to be RGB, RGBA or gray.")
    if not isinstance(channel_axis, int):
        raise TypeError("channel_axis must be an integer")
    ca = channel_axis % x.dim()
    if x.shape[ca] != 3:
        raise ValueError(
            f"the input array must have size 3 along `channel_axis`, got {tuple(x.shape)}")
    if ca != x.dim() - 1:
        x = x.movedim(ca, -1)
    x = x.contiguous()
    fdt = _supported_float_type(x.dtype)
    out = _array.empty(tuple(x.shape[:-1]), fdt)
    _lib.check(_lib.lib().hc_rgb2gray(_array.hc_array(x), _array.hc_array(out),
                                      _array.stream()), "rgb2gray")
    return out
