"""``difference_of_gaussians`` (``filters/_gaussian.py:13-174`` of the
reference, paths relative to ``python/cucim/src/cucim/skimage/``)."""

import numpy as np

from ... import _array, _lib
from .._shared.filters import gaussian
from ..util.dtype import img_as_float


def difference_of_gaussians(image, low_sigma, high_sigma=None, *, mode="nearest", cval=0,
                            channel_axis=None, truncate=4.0):
    """Band-pass filter: the Gaussian blur at ``low_sigma`` minus the blur at
    ``high_sigma`` (default ``1.6 * low_sigma``), both of the image converted
    as ``img_as_float`` does; the subtraction is one float rounding per
    element (``out -= ...`` of the reference)."""
    image = img_as_float(image)
    spatial_dims = image.dim() - 1 if channel_axis is not None else image.dim()
    if np.isscalar(low_sigma):
        low_sigma = (low_sigma,) * spatial_dims
    low_sigma = tuple(map(float, low_sigma))
    if high_sigma is None:
        high_sigma = tuple(low * 1.6 for low in low_sigma)
    elif np.isscalar(high_sigma):
        high_sigma = (high_sigma,) * spatial_dims
    high_sigma = tuple(map(float, high_sigma))
    if len(low_sigma) != 1 and len(low_sigma) != spatial_dims:
        raise ValueError("low_sigma must have length equal to number of"
                         " spatial dimensions of input")
    if len(high_sigma) != 1 and len(high_sigma) != spatial_dims:
        raise ValueError("high_sigma must have length equal to number of"
                         " spatial dimensions of input")
    if any(s_hi < s_low for s_hi, s_low in zip(high_sigma, low_sigma)):
        raise ValueError("high_sigma must be equal to or larger than"
                         "low_sigma for all axes")
    out = gaussian(image, sigma=low_sigma, mode=mode, cval=cval, channel_axis=channel_axis,
                   truncate=truncate, preserve_range=False)
    high = gaussian(image, sigma=high_sigma, mode=mode, cval=cval, channel_axis=channel_axis,
                    truncate=truncate, preserve_range=False)
    return _array.binary(out, high, out, _lib.BIN_SUB)
