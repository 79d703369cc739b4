"""Hessian matrix and its eigenvalues (``feature/corner.py`` of the
reference, paths relative to ``python/cucim/src/cucim/skimage/``)."""

import math
from itertools import combinations_with_replacement
from warnings import warn

import numpy as np
import torch

from ... import _array, _lib, _trace
from .._shared.utils import _supported_float_type, astype
from .._vendored import ndimage as ndi
from .._vendored._ndimage_filters import _gaussian_kernel1d, _trimmed
from .._vendored._ndimage_util import _mode_code

__all__ = ["hessian_matrix", "hessian_matrix_eigvals", "hessian_kernels", "structure_tensor",
           "structure_tensor_eigenvalues"]


def hessian_kernels(sigma, ndim, dtype):
    """Per-axis (order-0, order-1) Gaussian kernels of
    ``_hessian_matrix_with_gaussian`` (``corner.py:218-232``): sigma/sqrt(2),
    truncate 8 when every sigma > 1 else 100, radius int(truncate*s + 0.5),
    cast to the image float type, trimmed to their common nonzero support.
    Returns [(w0, w1, anchor)] per axis."""
    if np.isscalar(sigma):
        sigma = (sigma,) * ndim
    truncate = 8 if all(s > 1 for s in sigma) else 100
    sq1_2 = 1 / math.sqrt(2)
    out = []
    for s in sigma:
        ss = sq1_2 * s
        r = int(truncate * float(ss) + 0.5)
        w0 = _gaussian_kernel1d(ss, 0, r).astype(dtype)
        w1 = _gaussian_kernel1d(ss, 1, r).astype(dtype)
        nz = np.flatnonzero((w0 != 0) | (w1 != 0))
        if nz.size == 0:
            out.append((w0[:1] * 0, w1[:1] * 0, 0))
            continue
        lo, hi = int(nz[0]), int(nz[-1]) + 1
        out.append((w0[lo:hi], w1[lo:hi], r - lo))
    return out


def _taps_arrays(kernels, f64):
    """hc_taps arrays (order 0 and order 1) for a list of per-axis kernels."""
    n = len(kernels)
    ts = (_lib.HCTaps * n)()
    td = (_lib.HCTaps * n)()
    keep = []
    for i, (w0, w1, anc) in enumerate(kernels):
        t0, k0 = _lib.make_taps(w0.astype(np.float64), anc, f64)
        t1, k1 = _lib.make_taps(w1.astype(np.float64), anc, f64)
        ts[i], td[i] = t0, t1
        keep += [k0, k1]
    return ts, td, keep


def _hessian_sequential(image, sigma, mode, cval, order):
    """The reference's pass sequence (``corner.py:218-254``) on the tiled
    line kernels; used when the fused schedule does not fit."""
    ndim = image.dim()
    if np.isscalar(sigma):
        sigma = (sigma,) * ndim
    truncate = 8 if all(s > 1 for s in sigma) else 100
    sq1_2 = 1 / math.sqrt(2)
    sigma_scaled = tuple(sq1_2 * s for s in sigma)
    kw = dict(sigma=sigma_scaled, mode=mode, cval=cval, truncate=truncate)
    orders = tuple([0] * d + [1] + [0] * (ndim - d - 1) for d in range(ndim))
    gradients = [ndi.gaussian_filter(image, order=orders[d], **kw) for d in range(ndim)]
    axes = range(ndim)
    if order == "xy":
        axes = reversed(axes)
    return [ndi.gaussian_filter(gradients[a0], order=orders[a1], **kw)
            for a0, a1 in combinations_with_replacement(axes, 2)]


def _hessian_matrix_with_gaussian(image, sigma=1, mode="reflect", cval=0, order="rc"):
    """``corner.py:161-254`` as one fused schedule (2-D / 3-D)."""
    from ..util.dtype import img_as_float

    image = img_as_float(image)
    float_dtype = _supported_float_type(image.dtype)
    image = astype(image, float_dtype)
    if image.dim() > 2 and order == "xy":
        raise ValueError("Order 'xy' is only supported for 2D images.")
    if order not in ["rc", "xy"]:
        raise ValueError(f"unrecognized order: {order}")
    ndim = image.dim()
    commuting = not (mode in ("constant", "grid-constant") and float(cval) != 0.0)
    if ndim in (2, 3) and image.numel() > 0 and commuting:
        x = image.contiguous()
        nH = ndim * (ndim + 1) // 2
        H = [_array.empty(x.shape, float_dtype) for _ in range(nH)]
        f64 = float_dtype == np.float64
        ts, td, keep = _taps_arrays(hessian_kernels(sigma, ndim, float_dtype), f64)
        scratch = torch.empty(64, dtype=torch.uint8, device=x.device)
        rc = _lib.lib().hc_hessian_gaussian(
            _array.hc_array(x), _array.hc_arrays(H), nH, ts, td, _mode_code(mode),
            float(cval), 1 if order == "xy" else 0, scratch.data_ptr(), _array.stream())
        if rc == _lib.OK:
            return H
        if rc != _lib.ERR_RESOURCE:
            _lib.check(rc, "hessian_matrix")
    return _hessian_sequential(image, sigma, mode, cval, order)


def _gradient(f, axis):
    """``_shared/_gradient.py`` along one axis."""
    f = _array.as_tensor(f).contiguous()
    out = _array.empty(f.shape, f.dtype)
    _lib.check(_lib.lib().hc_gradient(_array.hc_array(f), _array.hc_array(out), axis,
                                      _array.stream()), "gradient")
    return out


@_trace.annotate("hessian_matrix")
def hessian_matrix(image, sigma=1, mode="constant", cval=0, order="rc",
                   use_gaussian_derivatives=None):
    """Hessian matrix elements (``corner.py:257-384``)."""
    from ..filters import gaussian
    from ..util.dtype import img_as_float

    image = img_as_float(_array.as_tensor(image))
    float_dtype = _supported_float_type(image.dtype)
    image = astype(image, float_dtype)
    if image.dim() > 2 and order == "xy":
        raise ValueError("Order 'xy' is only supported for 2D images.")
    if order not in ["rc", "xy"]:
        raise ValueError(f"unrecognized order: {order}")
    if use_gaussian_derivatives is None:
        use_gaussian_derivatives = False
        warn("use_gaussian_derivatives currently defaults to False, but will "
             "change to True in a future version. Please specify this "
             "argument explicitly to maintain the current behavior",
             category=FutureWarning, stacklevel=2)
    if use_gaussian_derivatives:
        return _hessian_matrix_with_gaussian(image, sigma=sigma, mode=mode, cval=cval,
                                             order=order)
    channel_axis = -1 if (image.dim() == 3 and image.shape[-1] == 3) else None
    gaussian_filtered = gaussian(image, sigma=sigma, mode=mode, cval=cval,
                                 channel_axis=channel_axis)
    gradients = [_gradient(gaussian_filtered, a) for a in range(image.dim())]
    axes = range(image.dim())
    if order == "xy":
        axes = reversed(axes)
    return [_gradient(gradients[a0], a1) for a0, a1 in combinations_with_replacement(axes, 2)]


def _symmetric_compute_eigenvalues(S_elems, sort="descending", abs_sort=False):
    """``corner.py:699-755``: closed form for 2-D / 3-D, the eigvalsh branch
    (Jacobi rotations on the device) for other orders up to 8; returned as
    one (ndim,) + shape array."""
    S = [_array.as_tensor(s) for s in S_elems]
    ne = {3: 2, 6: 3}.get(len(S))
    if ne is None:
        # the eigvalsh branch builds matrices of the image's rank from the
        # first n(n+1)/2 elements (_symmetric_image, corner.py:758-783)
        ne = S[0].dim()
        if not 1 <= ne <= 8:
            raise ValueError("eigenvalues of matrices of order 1..8 only")
        if len(S) < ne * (ne + 1) // 2:
            raise ValueError(f"{ne}-D images need {ne * (ne + 1) // 2} matrix elements")
        S = S[:ne * (ne + 1) // 2]
    if _array.np_dtype(S[0].dtype).kind != "f":
        raise ValueError("expected real-valued floating point matrices")
    if sort not in ("ascending", "descending"):
        raise ValueError(f"unknown sort type: {sort}")
    dt = S[0].dtype
    S = [s.to(dt).contiguous() for s in S]
    if dt not in (_array.torch_dtype(np.float32), _array.torch_dtype(np.float64)):
        S = [astype(s, np.float32) for s in S]
        dt = S[0].dtype
    eigs = _array.empty((ne,) + tuple(S[0].shape), dt)
    E = [eigs[i] for i in range(ne)]
    _lib.check(_lib.lib().hc_sym_eigvals(_array.hc_arrays(S), len(S), _array.hc_arrays(E),
                                         0 if sort == "ascending" else 1,
                                         1 if abs_sort else 0, _array.stream()),
               "eigenvalues")
    return eigs


@_trace.annotate("hessian_matrix_eigvals")
def hessian_matrix_eigvals(H_elems):
    """Eigenvalues of the Hessian, decreasing (``corner.py:824-856``)."""
    return _symmetric_compute_eigenvalues(H_elems)


def _prepare_grayscale_input_nD(image):
    """``feature/util.py:19-24``: squeeze, 2-5 dimensions, img_as_float in
    the supported float type."""
    from .._shared.utils import check_nD
    from ..util.dtype import img_as_float

    image = _array.as_tensor(image).squeeze()
    check_nD(image, range(2, 6))
    image = img_as_float(image)
    return astype(image, _supported_float_type(_array.np_dtype(image.dtype)))


def _compute_derivatives(image, mode="constant", cval=0):
    """``corner.py:27-52``: ``ndi.sobel`` along every axis."""
    return [ndi.sobel(image, axis=i, mode=mode, cval=cval) for i in range(image.dim())]


@_trace.annotate("structure_tensor")
def structure_tensor(image, sigma=1, mode="constant", cval=0, order="rc"):
    """Structure tensor (``corner.py:55-158``): the Gaussian-smoothed
    products of the Sobel derivatives, in ``combinations_with_replacement``
    order of the axes (reversed for ``order='xy'``, 2-D only).  A 3-D image
    whose last axis has 3 entries is smoothed per channel, as the reference
    does."""
    from .._shared.filters import gaussian

    t = _array.as_tensor(image)
    if order == "xy" and t.dim() > 2:
        raise ValueError('Only "rc" order is supported for dim > 2.')
    if order not in ["rc", "xy"]:
        raise ValueError(f'order {order} is invalid. Must be either "rc" or "xy"')
    if not np.isscalar(sigma):
        sigma = tuple(sigma)
        if len(sigma) != t.dim():
            raise ValueError("sigma must have as many elements as image has axes")
    image = _prepare_grayscale_input_nD(t).contiguous()
    derivatives = _compute_derivatives(image, mode=mode, cval=cval)
    if order == "xy":
        derivatives = list(reversed(derivatives))
    channel_axis = -1 if (image.dim() == 3 and image.shape[-1] == 3) else None
    out = []
    for d0, d1 in combinations_with_replacement(derivatives, 2):
        prod = _array.binary(d0, d1, _array.empty(tuple(d0.shape), _array.np_dtype(d0.dtype)),
                             _lib.BIN_MUL)
        out.append(gaussian(prod, sigma=sigma, mode=mode, cval=cval, channel_axis=channel_axis))
    return out


@_trace.annotate("structure_tensor_eigenvalues")
def structure_tensor_eigenvalues(A_elems):
    """Eigenvalues of the structure tensor, decreasing (``corner.py:786-821``)."""
    return _symmetric_compute_eigenvalues(A_elems)
