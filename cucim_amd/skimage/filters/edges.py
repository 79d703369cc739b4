"""Edge filters (``filters/edges.py`` of the reference, paths relative to
``python/cucim/src/cucim/skimage/``)."""

import math

import numpy as np
import torch

from ... import _array, _lib
from .._shared.utils import _supported_float_type, astype
from .._vendored import ndimage as ndi
from .._vendored._ndimage_util import _mode_code
from ..util.dtype import img_as_float

# ``edges.py:33-69``
SOBEL_EDGE = np.array([1, 0, -1])
SOBEL_SMOOTH = np.array([1, 2, 1]) / 4
SCHARR_EDGE = np.array([1, 0, -1])
SCHARR_SMOOTH = np.array([3, 10, 3]) / 16
PREWITT_EDGE = np.array([1, 0, -1])
PREWITT_SMOOTH = np.full((3,), 1 / 3)
farid_smooth = np.array([0.0376593171958126, 0.249153396177344, 0.426374573253687,
                         0.249153396177344, 0.0376593171958126])
farid_edge = np.array([0.109603762960254, 0.276690988455557, 0, -0.276690988455557,
                       -0.109603762960254])


def _mask_filter_result(result, mask):
    """``edges.py:75-85``: multiply by the mask eroded with a full structure
    (border value 0)."""
    if mask is None:
        return result
    m = _array.as_tensor(mask)
    if m.dtype != torch.bool and m.dtype != torch.uint8:
        m = m != 0
    m = m.contiguous()
    if tuple(m.shape) != tuple(result.shape):
        raise ValueError("mask must have the image's shape")
    r = result if result.is_contiguous() else result.contiguous()
    _lib.check(_lib.lib().hc_mask_eroded(_array.hc_array(r), _array.hc_array(m),
                                         _array.stream()), "mask")
    if r is not result:
        _array.copy_cast(r, result)
    return result


def _generic_edge_filter(image, *, smooth_weights, edge_weights=(1, 0, -1), axis=None,
                         mode="reflect", cval=0.0, mask=None):
    """``edges.py:148-241``."""
    image = _array.as_tensor(image)
    ndim = image.dim()
    if axis is None:
        axes = list(range(ndim))
    elif np.isscalar(axis):
        axes = [axis]
    else:
        axes = list(axis)
    return_magnitude = len(axes) > 1
    if _array.np_dtype(image.dtype).kind == "f":
        float_dtype = _supported_float_type(image.dtype)
        image = astype(image, float_dtype)
    else:
        image = img_as_float(image)
        float_dtype = _array.np_dtype(image.dtype)
    edge = np.asarray(edge_weights, dtype=float_dtype)
    smooth = np.asarray(smooth_weights, dtype=float_dtype)
    if return_magnitude:
        edge = (edge / math.sqrt(ndim)).astype(float_dtype)
    lib = _lib.lib()
    if (ndim == 2 and return_magnitude and sorted(axes) == [0, 1] and axes == [0, 1]
            and edge.size == 3 and smooth.size == 3 and isinstance(mode, str)):
        x = image.contiguous()
        out = _array.empty(x.shape, float_dtype)
        e3 = np.ascontiguousarray(edge[::-1], dtype=np.float64)
        s3 = np.ascontiguousarray(smooth[::-1], dtype=np.float64)
        import ctypes

        P = ctypes.POINTER(ctypes.c_double)
        rc = lib.hc_edge_magnitude(_array.hc_array(x), _array.hc_array(out),
                                   e3.ctypes.data_as(P), s3.ctypes.data_as(P), 3,
                                   _mode_code(mode), None, _array.stream())
        _lib.check(rc, "edge magnitude")
        return out
    output = None
    for i, edge_dim in enumerate(axes):
        ax_output = ndi.convolve1d(image, edge, axis=edge_dim, mode=mode, output=float_dtype)
        smooth_axes = list(set(range(ndim)) - {edge_dim})
        for smooth_dim in smooth_axes:
            ax_output = ndi.convolve1d(ax_output, smooth, axis=smooth_dim, mode=mode,
                                       output=float_dtype)
        if return_magnitude:
            if i == 0:
                output = _array.empty(ax_output.shape, float_dtype)
            _lib.check(lib.hc_square_acc(_array.hc_array(output), _array.hc_array(ax_output),
                                         1 if i == 0 else 0, _array.stream()), "edge")
        else:
            output = ax_output
    if return_magnitude:
        _lib.check(lib.hc_sqrt_inplace(_array.hc_array(output), _array.stream()), "edge")
    return output


def sobel(image, mask=None, *, axis=None, mode="reflect", cval=0.0):
    """Sobel edge map (``edges.py:244-299``)."""
    out = _generic_edge_filter(image, smooth_weights=SOBEL_SMOOTH, axis=axis, mode=mode,
                               cval=cval)
    return _mask_filter_result(out, mask)


def scharr(image, mask=None, *, axis=None, mode="reflect", cval=0.0):
    """Scharr edge map (``edges.py:362-420``)."""
    out = _generic_edge_filter(image, smooth_weights=SCHARR_SMOOTH, axis=axis, mode=mode,
                               cval=cval)
    return _mask_filter_result(out, mask)


def prewitt(image, mask=None, *, axis=None, mode="reflect", cval=0.0):
    """Prewitt edge map (``edges.py:480-540``)."""
    out = _generic_edge_filter(image, smooth_weights=PREWITT_SMOOTH, axis=axis, mode=mode,
                               cval=cval)
    return _mask_filter_result(out, mask)


def farid(image, mask=None, *, axis=None, mode="reflect", cval=0.0):
    """Farid edge map, 5-tap derivative / smoothing pair (``edges.py:783-850``)."""
    out = _generic_edge_filter(image, smooth_weights=farid_smooth, edge_weights=farid_edge,
                               axis=axis, mode=mode, cval=cval)
    return _mask_filter_result(out, mask)


def _edge_input(image):
    """Float conversion of the 2-D-only edge filters (``edges.py:680-684``):
    float images keep their supported float type, others go through
    img_as_float."""
    image = _array.as_tensor(image)
    if _array.np_dtype(image.dtype).kind == "f":
        return astype(image, _supported_float_type(image.dtype))
    return img_as_float(image)


def _check_2d(image):
    from .._shared.utils import check_nD

    check_nD(_array.as_tensor(image), 2)


def sobel_h(image, mask=None):
    """Horizontal edges: sobel along axis 0 (``edges.py:302-329``)."""
    _check_2d(image)
    return sobel(image, mask=mask, axis=0)


def sobel_v(image, mask=None):
    """Vertical edges: sobel along axis 1 (``edges.py:332-359``)."""
    _check_2d(image)
    return sobel(image, mask=mask, axis=1)


def scharr_h(image, mask=None):
    """Horizontal edges: scharr along axis 0 (``edges.py:425-457``)."""
    _check_2d(image)
    return scharr(image, mask=mask, axis=0)


def scharr_v(image, mask=None):
    """Vertical edges: scharr along axis 1 (``edges.py:460-491``)."""
    _check_2d(image)
    return scharr(image, mask=mask, axis=1)


def prewitt_h(image, mask=None):
    """Horizontal edges: prewitt along axis 0 (``edges.py:554-581``)."""
    _check_2d(image)
    return prewitt(image, mask=mask, axis=0)


def prewitt_v(image, mask=None):
    """Vertical edges: prewitt along axis 1 (``edges.py:584-611``)."""
    _check_2d(image)
    return prewitt(image, mask=mask, axis=1)


# ``edges.py:50-53``: 2-D Roberts cross kernels (convolution weights)
ROBERTS_PD_WEIGHTS = np.array([[1, 0], [0, -1]], dtype=np.float64)
ROBERTS_ND_WEIGHTS = np.array([[0, 1], [-1, 0]], dtype=np.float64)
# ``edges.py:71-72``: full 5x5 Farid kernels (edge along axis 0 for _h)
HFARID_WEIGHTS = farid_edge[:, None] * farid_smooth[None, :]
VFARID_WEIGHTS = np.copy(HFARID_WEIGHTS.T)


def _convolve_masked(image, weights, mask):
    """n-D convolution with the weights cast to the image's float dtype, then
    the mask (``edges.py:659-738, 853-927``)."""
    image = _edge_input(image)
    w = np.asarray(weights, dtype=_array.np_dtype(image.dtype))
    result = ndi.convolve(image, w)
    return _mask_filter_result(result, mask)


def roberts_pos_diag(image, mask=None):
    """Roberts cross along the positive diagonal (``edges.py:659-696``)."""
    _check_2d(image)
    return _convolve_masked(image, ROBERTS_PD_WEIGHTS, mask)


def roberts_neg_diag(image, mask=None):
    """Roberts cross along the negative diagonal (``edges.py:699-736``)."""
    _check_2d(image)
    return _convolve_masked(image, ROBERTS_ND_WEIGHTS, mask)


def roberts(image, mask=None):
    """Roberts edge magnitude (``edges.py:614-660``): each step rounded in
    the float dtype as the reference's in-place ufuncs do (``edges.py:614-656``):
    pos *= pos; out = neg * neg; out += pos; sqrt; out /= sqrt(2)."""
    _check_2d(image)
    pos = roberts_pos_diag(image, mask)
    out = roberts_neg_diag(image, mask)
    lib = _lib.lib()
    st = _array.stream()
    _lib.check(lib.hc_square_acc(_array.hc_array(out), _array.hc_array(out), 1, st), "roberts")
    _lib.check(lib.hc_square_acc(_array.hc_array(out), _array.hc_array(pos), 0, st), "roberts")
    _lib.check(lib.hc_sqrt_inplace(_array.hc_array(out), st), "roberts")
    _lib.check(lib.hc_div_scalar(_array.hc_array(out), math.sqrt(2), st), "roberts")
    return out


def _laplacian_kernel(ndim, dtype):
    """The spatial kernel of ``restoration/uft.py:429-478`` (``laplacian``
    with shape None): 3^ndim zeros, -1 at the two axial neighbours of the
    centre along every axis, 2*ndim at the centre."""
    k = np.zeros([3] * ndim)
    for d in range(ndim):
        idx = [1] * ndim
        for v in (0, 2):
            idx[d] = v
            k[tuple(idx)] = -1.0
    k[(1,) * ndim] = 2.0 * ndim
    return k.astype(dtype)


def laplace(image, ksize=3, mask=None):
    """Discrete Laplace operator (``edges.py:739-780``)."""
    image = _edge_input(image)
    if ksize != 3:
        raise NotImplementedError("only ksize=3 is supported")
    k = _laplacian_kernel(image.dim(), _array.np_dtype(image.dtype))
    result = ndi.convolve(image, k)
    return _mask_filter_result(result, mask)


def farid_h(image, *, mask=None):
    """Horizontal Farid edges: the full 5x5 kernel (``edges.py:853-890``)."""
    _check_2d(image)
    return _convolve_masked(image, HFARID_WEIGHTS, mask)


def farid_v(image, *, mask=None):
    """Vertical Farid edges (``edges.py:893-927``)."""
    _check_2d(image)
    return _convolve_masked(image, VFARID_WEIGHTS, mask)
