"""Sparse correlation (``filters/_sparse.py`` of the reference, paths
relative to ``python/cucim/src/cucim/skimage/``)."""

import numpy as np
import torch

from ... import _array, _lib
from .._shared.utils import _supported_float_type
from .._vendored._ndimage_util import _mode_code

__all__ = ["correlate_sparse"]

# ndimage names of ``_shared/utils.py:1095-1102`` (_to_np_mode) and the
# numpy.pad names that pass through it unchanged, as boundary modes
_PAD_MODES = {"nearest": "nearest", "edge": "nearest", "reflect": "reflect",
              "symmetric": "reflect", "mirror": "mirror", "wrap": "wrap",
              "constant": "constant"}


def _validate_window_size(axis_sizes):
    """``_sparse.py:12-22``."""
    for axis_size in axis_sizes:
        if axis_size % 2 == 0:
            raise ValueError(
                "Window size for `threshold_sauvola` or `threshold_niblack` must not be "
                f"even on any dimension. Got {tuple(axis_sizes)}")


def correlate_sparse(image, kernel, mode="reflect"):
    """Correlation with a sparse kernel (``_sparse.py:94-118``): the image is
    padded by kernel.shape // 2 on both sides of each axis (numpy.pad's
    boundary rules; 'valid' pads nothing and shrinks the output), then the
    kernel's nonzero taps are summed in C order, starting from the corner tap
    (value 0 when it is zero), one rounding per product and per add in the
    image's supported float type."""
    image = _array.as_tensor(image)
    float_dtype = _supported_float_type(image.dtype)
    k = kernel.detach().cpu().numpy() if isinstance(kernel, torch.Tensor) else np.asarray(kernel)
    if k.ndim != image.dim():
        raise ValueError("kernel and image must have the same number of dimensions")
    valid = mode == "valid"
    if not valid:
        if mode not in _PAD_MODES:
            raise NotImplementedError(f"padding mode {mode!r} is not supported")
        _validate_window_size(k.shape)
    idx = np.nonzero(k)
    values = k[idx].astype(float_dtype)
    taps = np.stack(idx, axis=1).astype(np.int32) if k.ndim else np.zeros((0, 0), np.int32)
    if not any(np.all(t == 0) for t in taps):
        taps = np.concatenate([np.zeros((1, k.ndim), np.int32), taps])
        values = np.concatenate([np.zeros(1, float_dtype), values])
    if valid:
        oshape = tuple(max(0, s - w + 1) for s, w in zip(image.shape, k.shape))
    else:
        oshape = tuple(image.shape)
    out = _array.empty(oshape, float_dtype)
    taps = np.ascontiguousarray(taps)
    w64 = np.ascontiguousarray(values.astype(np.float64))
    ks = np.ascontiguousarray(np.asarray(k.shape, np.int32))
    mode_code = 0 if valid else _mode_code(_PAD_MODES[mode])
    _lib.check(_lib.lib().hc_correlate_sparse(
        _array.hc_array(image), _array.hc_array(out), ks.ctypes.data, taps.ctypes.data,
        w64.ctypes.data, len(w64), 1 if valid else 0, mode_code, _array.stream()),
        "correlate_sparse")
    return out
