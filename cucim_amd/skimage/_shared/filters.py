"""``gaussian`` (``_shared/filters.py:21-166`` of the reference, paths
relative to ``python/cucim/src/cucim/skimage/``)."""

from collections.abc import Iterable

import numpy as np

from ... import _array
from .._vendored import _ndimage_filters as ndi_filters
from .._vendored import ndimage as ndi
from .utils import _supported_float_type, astype, convert_to_float


def gaussian(image, sigma=1, mode="nearest", cval=0, preserve_range=False,
             truncate=4.0, *, channel_axis=None, out=None):
    """Multi-dimensional Gaussian filter (skimage semantics: default
    ``mode='nearest'``, integer images converted with img_as_float unless
    ``preserve_range``)."""
    if out is None and channel_axis is None:
        # a float32 image through a recorded launch plan (same result as the
        # path below: no conversion, a new float32 output, the fused filter)
        r = ndi_filters._gaussian_fast(image, sigma, 0, mode, cval, truncate, None, (0, 1))
        if r is not None:
            return r
    sigma_msg = "Sigma values less than zero are not valid"
    if not isinstance(sigma, Iterable):
        if sigma < 0:
            raise ValueError(sigma_msg)
    elif any(s < 0 for s in sigma):
        raise ValueError(sigma_msg)
    image = _array.as_tensor(image)
    if channel_axis is not None:
        axes = tuple(ax for ax in range(image.dim()) if ax != channel_axis % image.dim())
    else:
        axes = tuple(range(image.dim()))
    image = convert_to_float(image, preserve_range)
    float_dtype = _supported_float_type(image.dtype)
    image = astype(image, float_dtype)
    if out is None:
        # a new output of the float type (through a recorded launch plan on
        # repeated float32 2-D calls)
        return ndi.gaussian_filter(image, sigma, mode=mode, cval=cval, truncate=truncate,
                                   axes=axes)
    out_t = _array.as_tensor(out)
    if _array.np_dtype(out_t.dtype).kind != "f":
        raise ValueError(f"dtype of `out` must be floating point; got {out_t.dtype!r}.")
    ndi.gaussian_filter(image, sigma, output=out_t, mode=mode, cval=cval,
                        truncate=truncate, axes=axes)
    return out
