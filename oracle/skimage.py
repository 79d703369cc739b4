"""Oracle: numpy restatement of the reference's skimage-level functions.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

``dtype`` arguments select the working float type.  float64 is "an FP64 run
of the reference": every pass, eigenvalue and Frangi step in float64.  The
separable passes use ``oracle.ndimage`` (``acc="f64"``) by default; passing
``ndi=scipy.ndimage`` runs the same algorithm on scipy's C filters, which is
what the timed CPU baseline does (scikit-image itself calls scipy.ndimage).

Reference files (relative to ``python/cucim/src/cucim/skimage/``):
  _shared/filters.py:21-166        gaussian
  _shared/utils.py:1019-1179       convert_to_float, _supported_float_type
  util/dtype.py:219-517            img_as_float32/64 (_convert)
  filters/edges.py:33-299          sobel / prewitt / _generic_edge_filter
  filters/edges.py:362-927         scharr, farid, roberts*, laplace, farid_h/v
  restoration/uft.py:429-478       laplacian (spatial kernel)
  filters/_sparse.py:34-118        correlate_sparse
  exposure/exposure.py:35-345      histogram
  filters/thresholding.py:298-782  threshold_otsu / yen / li (with scikit-image's
                                   histogram formulas, restated)
  measure/_label.py:21-124         label
  color/colorconv.py:160-180,1037-1095  rgb2gray
  feature/corner.py:161-856        hessian_matrix, eigenvalue kernels
  filters/ridges.py:228-476        frangi
  filters/ridges.py:50-225         meijering, sato
  filters/ridges.py:479-565        hessian (hybrid Hessian filter)
  _shared/_gradient.py             finite-difference gradient
  filters/_gaussian.py:13-174      difference_of_gaussians
  feature/corner.py:27-158         structure_tensor (_compute_derivatives)
  feature/util.py:19-24            _prepare_grayscale_input_nD
  feature/corner.py:786-821        structure_tensor_eigenvalues
"""

import math
import warnings
from itertools import combinations_with_replacement

import numpy as np

from . import ndimage as ondi
from .ndimage import supported_float_type

__all__ = [
    "img_as_float",
    "img_as_float32",
    "convert_to_float",
    "gaussian",
    "sobel",
    "prewitt",
    "scharr",
    "farid",
    "roberts",
    "roberts_pos_diag",
    "roberts_neg_diag",
    "laplace",
    "farid_h",
    "farid_v",
    "correlate_sparse",
    "rgb2gray",
    "gradient",
    "hessian_matrix",
    "eig22",
    "eig33",
    "symmetric_compute_eigenvalues",
    "hessian_matrix_eigvals",
    "frangi",
    "sato",
    "meijering",
    "hessian",
]


def _int_to_float(x, out_dtype):
    """``util/dtype.py:335-356``: uint -> x * (1/imax) in the computation float
    type (float32 for <= 16-bit ints, i.e. f32(x) * f32(1/imax) rounded once);
    signed ints additionally clip at -1."""
    x = np.asarray(x)
    out_dtype = np.dtype(out_dtype)
    info = np.iinfo(x.dtype)
    # _dtype_itemsize(itemsize_in, dtype_out, float32, float64) (dtype.py:94-114)
    comp = next(dt for dt in (out_dtype.type, np.float32, np.float64)
                if np.dtype(dt).itemsize >= x.dtype.itemsize)
    y = x.astype(comp) * comp(1.0 / info.max)
    if x.dtype.kind == "i":
        y = np.maximum(y, comp(-1.0))
    return y.astype(out_dtype)


def _convert_to(x, dtype):
    x = np.asarray(x)
    dtype = np.dtype(dtype)
    if x.dtype == dtype:
        return x
    if x.dtype.kind == "f":
        return x.astype(dtype)
    if x.dtype.kind == "b":
        return x.astype(dtype)
    return _int_to_float(x, dtype)


def img_as_float32(x):
    return _convert_to(x, np.float32)


def img_as_float(x):
    """``util/dtype.py:480-517``: ``_convert(image, _supported_float_type)``,
    so float16 -> float32, 8/16-bit ints and bool -> float32, others float64."""
    x = np.asarray(x)
    return _convert_to(x, supported_float_type(x.dtype))


def convert_to_float(image, preserve_range):
    """``_shared/utils.py:1019-1052``."""
    image = np.asarray(image)
    if image.dtype == np.float16:
        return image.astype(np.float32)
    if preserve_range:
        if image.dtype.char not in "df":
            image = image.astype(supported_float_type(image.dtype))
    else:
        image = img_as_float(image)
    return image


def gaussian(image, sigma=1, mode="nearest", cval=0, preserve_range=False,
             truncate=4.0, *, channel_axis=None, ndi=None, dtype=None):
    """``_shared/filters.py:21-166``."""
    ndi = ondi if ndi is None else ndi
    image = np.asarray(image)
    if np.isscalar(sigma):
        if sigma < 0:
            raise ValueError("Sigma values less than zero are not valid")
    elif any(s < 0 for s in sigma):
        raise ValueError("Sigma values less than zero are not valid")
    if channel_axis is not None:
        axes = tuple(a for a in range(image.ndim) if a != channel_axis % image.ndim)
    else:
        axes = tuple(range(image.ndim))
    image = convert_to_float(image, preserve_range)
    float_dtype = supported_float_type(image.dtype) if dtype is None else np.dtype(dtype)
    image = image.astype(float_dtype, copy=False)
    return ndi.gaussian_filter(image, sigma, output=float_dtype, mode=mode,
                               cval=cval, truncate=truncate, axes=axes)


def _generic_edge_filter(image, *, smooth_weights, edge_weights=(1, 0, -1),
                         axis=None, mode="reflect", cval=0.0, ndi=None,
                         dtype=None):
    """``filters/edges.py:148-241``: convolve1d(edge) along each edge axis,
    convolve1d(smooth) along the other axes (ascending), square, accumulate,
    sqrt.  cval is not forwarded (``edges.py:218-231``)."""
    ndi = ondi if ndi is None else ndi
    image = np.asarray(image)
    ndim = image.ndim
    if axis is None:
        axes = list(range(ndim))
    elif np.isscalar(axis):
        axes = [axis]
    else:
        axes = list(axis)
    return_magnitude = len(axes) > 1
    if image.dtype.kind == "f":
        float_dtype = supported_float_type(image.dtype)
        image = image.astype(float_dtype, copy=False)
    else:
        float_dtype = supported_float_type(image.dtype)
        image = _convert_to(image, float_dtype)
    if dtype is not None:
        float_dtype = np.dtype(dtype)
        image = image.astype(float_dtype)
    wdt = np.float64 if ndi is not ondi else float_dtype
    edge = np.asarray(edge_weights, dtype=float_dtype).astype(wdt)
    smooth = np.asarray(smooth_weights, dtype=float_dtype).astype(wdt)
    if return_magnitude:
        edge = edge / math.sqrt(ndim)
    output = None
    for i, edge_dim in enumerate(axes):
        ax = ndi.convolve1d(image, edge, axis=edge_dim, mode=mode, output=float_dtype)
        # edges.py:229 subtracts the axis as given: a negative axis is not
        # removed, so the edge axis is smoothed as well
        for sd in sorted(set(range(ndim)) - {edge_dim}):
            ax = ndi.convolve1d(ax, smooth, axis=sd, mode=mode, output=float_dtype)
        if return_magnitude:
            ax = ax * ax
        output = ax if i == 0 else output + ax
    if return_magnitude:
        output = np.sqrt(output)
    return output


def sobel(image, mask=None, *, axis=None, mode="reflect", cval=0.0, ndi=None,
          dtype=None):
    """``filters/edges.py:244-299`` (weights ``:33-35``; mask ``:75-85``)."""
    out = _generic_edge_filter(image, smooth_weights=np.array([1, 2, 1]) / 4,
                               axis=axis, mode=mode, cval=cval, ndi=ndi,
                               dtype=dtype)
    if mask is not None:
        import scipy.ndimage as sndi

        fp = sndi.generate_binary_structure(mask.ndim, mask.ndim)
        m = sndi.binary_erosion(np.asarray(mask), fp, border_value=0)
        out = out * m
    return out


def prewitt(image, mask=None, *, axis=None, mode="reflect", cval=0.0, ndi=None,
            dtype=None):
    """``filters/edges.py:42-44`` weights with ``_generic_edge_filter``."""
    out = _generic_edge_filter(image, smooth_weights=np.full((3,), 1 / 3),
                               axis=axis, mode=mode, cval=cval, ndi=ndi,
                               dtype=dtype)
    if mask is not None:
        import scipy.ndimage as sndi

        fp = sndi.generate_binary_structure(mask.ndim, mask.ndim)
        m = sndi.binary_erosion(np.asarray(mask), fp, border_value=0)
        out = out * m
    return out


def _mask_result(out, mask):
    """``filters/edges.py:75-85``: multiply by the mask eroded with a full
    structure (border value 0)."""
    if mask is None:
        return out
    import scipy.ndimage as sndi

    m = np.asarray(mask)
    fp = sndi.generate_binary_structure(m.ndim, m.ndim)
    m = sndi.binary_erosion(m, fp, border_value=0)
    return (out * m).astype(out.dtype)


def scharr(image, mask=None, *, axis=None, mode="reflect", cval=0.0, ndi=None,
           dtype=None):
    """``filters/edges.py:362-420`` (weights ``:37-40``)."""
    out = _generic_edge_filter(image, smooth_weights=np.array([3, 10, 3]) / 16,
                               axis=axis, mode=mode, cval=cval, ndi=ndi, dtype=dtype)
    return _mask_result(out, mask)


# ``filters/edges.py:60-69``: Farid & Simoncelli (2004) 5-tap pair
FARID_SMOOTH = np.array([0.0376593171958126, 0.249153396177344, 0.426374573253687,
                         0.249153396177344, 0.0376593171958126])
FARID_EDGE = np.array([0.109603762960254, 0.276690988455557, 0, -0.276690988455557,
                       -0.109603762960254])


def farid(image, mask=None, *, axis=None, mode="reflect", cval=0.0, ndi=None, dtype=None):
    """``filters/edges.py:783-850``."""
    out = _generic_edge_filter(image, smooth_weights=FARID_SMOOTH, edge_weights=FARID_EDGE,
                               axis=axis, mode=mode, cval=cval, ndi=ndi, dtype=dtype)
    return _mask_result(out, mask)


def _edge_float(image):
    """``filters/edges.py:680-684``: float images keep their supported float
    type, others go through img_as_float."""
    image = np.asarray(image)
    if image.dtype.kind == "f":
        return image.astype(supported_float_type(image.dtype), copy=False)
    return img_as_float(image)


def _conv_masked(image, weights, mask, acc):
    x = _edge_float(image)
    w = np.asarray(weights, dtype=x.dtype)
    return _mask_result(ondi.convolve(x, w, acc=acc), mask)


def roberts_pos_diag(image, mask=None, *, acc="f64"):
    """``filters/edges.py:659-696``: convolve with [[1, 0], [0, -1]]."""
    return _conv_masked(image, [[1, 0], [0, -1]], mask, acc)


def roberts_neg_diag(image, mask=None, *, acc="f64"):
    """``filters/edges.py:699-736``: convolve with [[0, 1], [-1, 0]]."""
    return _conv_masked(image, [[0, 1], [-1, 0]], mask, acc)


def roberts(image, mask=None, *, acc="f64"):
    """``filters/edges.py:614-656``: pos *= pos; out = neg * neg; out += pos;
    sqrt(out); out /= sqrt(2), each step in the float dtype."""
    pos = roberts_pos_diag(image, mask, acc=acc)
    neg = roberts_neg_diag(image, mask, acc=acc)
    F = pos.dtype.type
    out = (neg * neg).astype(F) + (pos * pos).astype(F)
    out = np.sqrt(out.astype(F))
    return (out / F(math.sqrt(2))).astype(F)


def laplacian_kernel(ndim, dtype=np.float64):
    """``restoration/uft.py:429-478`` with shape None: -1 at the axial
    neighbours of the centre of a 3^ndim block, 2*ndim at the centre."""
    k = np.zeros([3] * ndim)
    c = (1,) * ndim
    for d in range(ndim):
        for v in (0, 2):
            idx = list(c)
            idx[d] = v
            k[tuple(idx)] = -1.0
    k[c] = 2.0 * ndim
    return k.astype(dtype)


def laplace(image, ksize=3, mask=None, *, acc="f64"):
    """``filters/edges.py:739-780``."""
    x = _edge_float(image)
    if ksize != 3:
        raise NotImplementedError("only ksize=3 is supported")
    return _mask_result(ondi.convolve(x, laplacian_kernel(x.ndim, x.dtype), acc=acc), mask)


def farid_h(image, *, mask=None, acc="f64"):
    """``filters/edges.py:853-890``: the 5x5 kernel edge (axis 0) x smooth."""
    return _conv_masked(image, FARID_EDGE[:, None] * FARID_SMOOTH[None, :], mask, acc)


def farid_v(image, *, mask=None, acc="f64"):
    """``filters/edges.py:893-927``: the transposed kernel."""
    return _conv_masked(image, (FARID_EDGE[:, None] * FARID_SMOOTH[None, :]).T, mask, acc)


def correlate_sparse(image, kernel, mode="reflect"):
    """``filters/_sparse.py:94-118`` with ``_correlate_sparse`` :55-91 and
    ``_get_view`` :34-52: numpy.pad by kernel.shape // 2 (modes translated
    by ``_shared/utils.py:1095-1102``; 'valid' pads nothing), then the views
    at the nonzero taps (C order, led by the corner tap, value 0 if it is
    zero) scaled and summed with one rounding per step in the supported
    float type."""
    image = np.asarray(image)
    F = supported_float_type(image.dtype).type
    image = image.astype(F, copy=False)
    k = np.asarray(kernel)
    if mode == "valid":
        padded = image
    else:
        if any(w % 2 == 0 for w in k.shape):
            raise ValueError("Window size must not be even on any dimension")
        np_mode = {"nearest": "edge", "reflect": "symmetric", "mirror": "reflect"}.get(mode, mode)
        padded = np.pad(image, [(w // 2, w // 2) for w in k.shape], mode=np_mode)
    idx = np.nonzero(k)
    values = list(k[idx].astype(F))
    taps = list(zip(*idx))
    corner = (0,) * k.ndim
    if corner not in taps:
        taps = [corner] + taps
        values = [F(0.0)] + values

    def view(t, v):
        sl = tuple(slice(c, s - (w - 1 - c)) for c, w, s in zip(t, k.shape, padded.shape))
        x = padded[sl]
        if v == 1:
            return x.copy()
        if v == -1:
            return -x
        with np.errstate(invalid="ignore", over="ignore"):
            return (F(v) * x).astype(F)

    out = view(taps[0], values[0])
    for t, v in zip(taps[1:], values[1:]):
        with np.errstate(invalid="ignore", over="ignore"):
            out = (out + view(t, v)).astype(F)
    return out


def histogram(image, nbins=256, source_range="image"):
    """``exposure/exposure.py:35-345`` (no channel axis, not normalised).
    Integer images: counts of every value in [min, max] (or the dtype's
    range) and those values as centers.  Float images: cupy.histogram's
    edges -- the range widened by 0.5 per side when it is one value, then
    start + i * ((stop - start) / nbins) in float64, the last edge = stop,
    rounded to the image's float type -- with numpy's searchsorted rule (the
    last bin closed); centers (e[:-1] + e[1:]) / 2 in that type."""
    x = np.asarray(image).reshape(-1)
    if x.dtype.kind in "iub":
        if source_range == "image":
            lo, hi = int(x.min()), int(x.max())
        elif x.dtype.kind == "b":
            lo, hi = 0, 1
        else:
            lo, hi = int(np.iinfo(x.dtype).min), int(np.iinfo(x.dtype).max)
        counts = np.bincount((x.astype(np.int64) - lo), minlength=hi - lo + 1)
        return counts, np.arange(lo, hi + 1)
    F = x.dtype.type
    first, last = (float(x.min()), float(x.max())) if source_range == "image" else (-1.0, 1.0)
    if first == last:
        first, last = first - 0.5, last + 0.5
    first, last = float(F(first)), float(F(last))
    step = (last - first) / nbins
    e = first + np.arange(nbins + 1, dtype=np.float64) * step
    e[-1] = last
    e = e.astype(F)
    idx = np.searchsorted(e, x, side="right") - 1
    idx[x == e[-1]] = nbins - 1
    keep = (idx >= 0) & (idx < nbins) & ~np.isnan(x)
    counts = np.bincount(idx[keep], minlength=nbins)
    return counts, ((e[:-1] + e[1:]) / F(2.0)).astype(F)


def threshold_otsu(image, nbins=256):
    """``filters/thresholding.py:355-416`` + scikit-image's Otsu on the
    float32 counts: argmax over splits of w1 * w2 * (m1 - m2)^2."""
    x = np.asarray(image)
    if np.all(x == x.reshape(-1)[0]):
        return x.reshape(-1)[0]
    counts, centers = histogram(x, nbins)
    counts = counts.astype(np.float32)
    w1 = np.cumsum(counts)
    w2 = np.cumsum(counts[::-1])[::-1]
    with np.errstate(divide="ignore", invalid="ignore"):
        m1 = np.cumsum(counts * centers) / w1
        m2 = (np.cumsum((counts * centers)[::-1]) / w2[::-1])[::-1]
    v = w1[:-1] * w2[1:] * (m1[:-1] - m2[1:]) ** 2
    return centers[np.argmax(v)]


def threshold_yen(image, nbins=256):
    """``filters/thresholding.py:419-474`` + scikit-image's Yen criterion."""
    counts, centers = histogram(np.asarray(image), nbins)
    if centers.size == 1:
        return centers[0]
    counts = counts.astype(np.float32)
    pmf = counts / counts.sum()
    P1 = np.cumsum(pmf)
    P1_sq = np.cumsum(pmf ** 2)
    P2_sq = np.cumsum(pmf[::-1] ** 2)[::-1]
    with np.errstate(divide="ignore", invalid="ignore"):
        crit = np.log(((P1_sq[:-1] * P2_sq[1:]) ** -1) * (P1[:-1] * (1.0 - P1[:-1])) ** 2)
    return centers[crit.argmax()]


def threshold_li(image, tolerance=None):
    """``filters/thresholding.py:614-782`` in float64 (means exact)."""
    x = np.asarray(image).reshape(-1)
    x = x[~np.isnan(x)] if x.dtype.kind == "f" else x
    if x.size == 0:
        return np.nan
    if np.all(x == x[0]):
        return x[0]
    if x.dtype.kind == "f":
        x = x[np.isfinite(x)]
        if x.size == 0:
            return 0.0
    image_min = x.min()
    v = (x - image_min).astype(x.dtype)
    if v.dtype.kind in "iu":
        tolerance = tolerance or 0.5
    else:
        tolerance = tolerance or float(np.min(np.diff(np.unique(v))) / 2)
    t_next = float(np.mean(v.astype(np.float64)))
    t_curr = -2 * tolerance
    eps = 100 * np.finfo(float).eps
    while abs(t_next - t_curr) > tolerance:
        t_curr = t_next
        fg = v > t_curr
        mean_fore = float(np.mean(v[fg].astype(np.float64)))
        mean_back = float(np.mean(v[~fg].astype(np.float64)))
        if mean_back == 0:
            break
        t_next = (mean_back - mean_fore) / (math.log(mean_back + eps) - math.log(mean_fore + eps))
    return t_next + float(image_min)


def label(label_image, background=0, connectivity=None):
    """``measure/_label.py:21-124``: connected regions of equal value
    (scipy.ndimage.label on each value's mask with
    generate_binary_structure(ndim, connectivity)), background 0, regions
    numbered in raster order of their first element."""
    import scipy.ndimage as sndi

    x = np.asarray(label_image)
    if x.dtype.kind not in "bui":
        x = x.astype(np.intp)
    nd = x.ndim
    connectivity = nd if connectivity is None else connectivity
    st = sndi.generate_binary_structure(nd, connectivity)
    comp = np.zeros(x.shape, np.int64)
    nxt = 0
    for v in np.unique(x):
        if v == background:
            continue
        lab, k = sndi.label(x == v, structure=st)
        comp[lab > 0] = lab[lab > 0] + nxt
        nxt += k
    flat = comp.reshape(-1)
    order = np.zeros(nxt + 1, np.int64)
    seen = flat[flat > 0]
    _, first = np.unique(seen, return_index=True)
    ids = np.unique(seen)
    rank = np.argsort(np.argsort(first))
    order[ids] = rank + 1
    return order[flat].reshape(x.shape).astype(np.int32), nxt


def rgb2gray(rgb, *, channel_axis=-1):
    """``color/colorconv.py:1037-1095``.

    uint8/16 & co -> float32 by ``f32(x) * f32(1/imax)`` (``_prepare_colorarray``
    ``:160-180``), then ``0.2125 R + 0.7154 G + 0.0721 B`` with double literals,
    summed left to right in double, stored in the float type.
    """
    rgb = np.asarray(rgb)
    if rgb.shape[channel_axis] != 3:
        raise ValueError(
            f"the input array must have size 3 along `channel_axis`, got {rgb.shape}")
    rgb = np.moveaxis(rgb, channel_axis, -1)
    fdt = supported_float_type(rgb.dtype)
    f = _convert_to(rgb, fdt)
    r = f[..., 0].astype(np.float64)
    g = f[..., 1].astype(np.float64)
    b = f[..., 2].astype(np.float64)
    y = 0.2125 * r
    y = y + 0.7154 * g
    y = y + 0.0721 * b
    return y.astype(fdt)


def gradient(f, axis=None):
    """``_shared/_gradient.py``: central differences inside, one-sided first
    differences at the two ends; ints -> supported float type."""
    f = np.asarray(f)
    if f.dtype.kind not in "fc":
        f = f.astype(supported_float_type(f.dtype))
    axes = range(f.ndim) if axis is None else ([axis] if np.isscalar(axis) else axis)
    outs = []
    for ax in axes:
        ax = ax % f.ndim
        if f.shape[ax] < 2:
            raise ValueError("Shape of array too small to calculate a numerical gradient, "
                             "at least 2 elements are required.")
        fm = np.moveaxis(f, ax, -1)
        o = np.empty_like(fm)
        o[..., 1:-1] = (fm[..., 2:] - fm[..., :-2]) / f.dtype.type(2.0)
        o[..., 0] = fm[..., 1] - fm[..., 0]
        o[..., -1] = fm[..., -1] - fm[..., -2]
        outs.append(np.moveaxis(o, -1, ax))
    return outs[0] if len(outs) == 1 else outs


def _hessian_matrix_with_gaussian(image, sigma=1, mode="reflect", cval=0,
                                  order="rc", ndi=None):
    """``feature/corner.py:161-254``: sigma/sqrt(2), truncate 8 (all sigma > 1)
    else 100; gradients g_d = gaussian_filter(I, order=e_d); then
    H_ab = gaussian_filter(g_a, order=e_b) for (a, b) in
    combinations_with_replacement(axes, 2)."""
    ndi = ondi if ndi is None else ndi
    image = np.asarray(image)
    if np.isscalar(sigma):
        sigma = (sigma,) * image.ndim
    truncate = 8 if all(s > 1 for s in sigma) else 100
    sq1_2 = 1 / math.sqrt(2)
    sigma_scaled = tuple(sq1_2 * s for s in sigma)
    kw = dict(sigma=sigma_scaled, mode=mode, cval=cval, truncate=truncate)
    ndim = image.ndim
    orders = tuple([0] * d + [1] + [0] * (ndim - d - 1) for d in range(ndim))
    gradients = [ndi.gaussian_filter(image, order=orders[d], **kw) for d in range(ndim)]
    axes = range(ndim)
    if order == "xy":
        axes = reversed(axes)
    return [ndi.gaussian_filter(gradients[a0], order=orders[a1], **kw)
            for a0, a1 in combinations_with_replacement(axes, 2)]


def hessian_matrix(image, sigma=1, mode="constant", cval=0, order="rc",
                   use_gaussian_derivatives=None, *, ndi=None, dtype=None):
    """``feature/corner.py:257-384``."""
    image = img_as_float(np.asarray(image))
    float_dtype = supported_float_type(image.dtype) if dtype is None else np.dtype(dtype)
    image = image.astype(float_dtype, copy=False)
    if image.ndim > 2 and order == "xy":
        raise ValueError("Order 'xy' is only supported for 2D images.")
    if order not in ["rc", "xy"]:
        raise ValueError(f"unrecognized order: {order}")
    if use_gaussian_derivatives is None:
        use_gaussian_derivatives = False
        warnings.warn("use_gaussian_derivatives currently defaults to False, but will "
                      "change to True in a future version. Please specify this "
                      "argument explicitly to maintain the current behavior",
                      category=FutureWarning, stacklevel=2)
    if use_gaussian_derivatives:
        return _hessian_matrix_with_gaussian(image, sigma=sigma, mode=mode,
                                             cval=cval, order=order, ndi=ndi)
    channel_axis = -1 if (image.ndim == 3 and image.shape[-1] == 3) else None
    g = gaussian(image, sigma=sigma, mode=mode, cval=cval,
                 channel_axis=channel_axis, ndi=ndi, dtype=float_dtype)
    grads = gradient(g)
    if image.ndim == 1:
        grads = [grads]
    axes = range(image.ndim)
    if order == "xy":
        axes = reversed(axes)
    return [gradient(grads[a0], axis=a1) for a0, a1 in combinations_with_replacement(axes, 2)]


def eig22(M00, M01, M11, sort="descending", abs_sort=False):
    """``feature/corner.py:473-527``.  Temporaries are of the input float type F
    (the products are formed in double and rounded to F), so for float32 input
    the arithmetic is float32."""
    F = np.asarray(M00).dtype.type
    m00 = np.asarray(M00, np.float64)
    m01 = np.asarray(M01, np.float64)
    m11 = np.asarray(M11, np.float64)
    tmp1 = (m01 * m01).astype(F)
    tmp1 = (tmp1 * F(4)).astype(F)
    tmp2 = (m00 - m11).astype(F)
    tmp2 = (tmp2 * tmp2).astype(F)
    tmp2 = (tmp2 + tmp1).astype(F)
    tmp2 = np.sqrt(tmp2).astype(F)
    tmp2 = (tmp2 / F(2)).astype(F)
    tmp1 = (m00 + m11).astype(F)
    tmp1 = (tmp1 / F(2)).astype(F)
    if sort == "ascending":
        l1 = (tmp1 - tmp2).astype(F)
        l2 = (tmp1 + tmp2).astype(F)
        if abs_sort:
            sw = np.abs(l1) > np.abs(l2)
            l1, l2 = np.where(sw, l2, l1), np.where(sw, l1, l2)
    elif sort == "descending":
        l1 = (tmp1 + tmp2).astype(F)
        l2 = (tmp1 - tmp2).astype(F)
        if abs_sort:
            sw = np.abs(l1) < np.abs(l2)
            l1, l2 = np.where(sw, l2, l1), np.where(sw, l1, l2)
    else:
        raise ValueError(f"unknown sort type: {sort}")
    return np.stack([l1, l2])


def eig33(a, d, f, b, e, c, sort="descending", abs_sort=False):
    """``feature/corner.py:566-696`` (Deledalle et al. 2017 closed form).

    Matrix [[a, d, f], [d, b, e], [f, e, c]].  The core runs in double;
    x1 >= 0 and 4 x1^3 - x2^2 >= 0 are clamped; phi = pi/2 when x2 == 0.
    The eigenvalues are rounded to the input float type F and then put through
    the reference's three compare-swap steps.  With ``abs_sort`` the keys are
    the absolute values taken BEFORE any swap (``abs_lam*`` are never
    updated, ``corner.py:620-640``); that is reproduced here.
    """
    F = np.asarray(a).dtype.type
    a, b, c, d, e, f = (np.asarray(v, np.float64) for v in (a, b, c, d, e, f))
    d_sq, e_sq, f_sq = d * d, e * e, f * f
    tmpa = 2 * a - b - c
    tmpb = 2 * b - a - c
    tmpc = 2 * c - a - b
    x2 = -tmpa * tmpb * tmpc
    x2 = x2 + 9 * (tmpc * d_sq + tmpb * f_sq + tmpa * e_sq)
    x2 = x2 - 54 * (d * e * f)
    x1 = a * a + b * b + c * c - a * b - a * c - b * c + 3 * (d_sq + e_sq + f_sq)
    # CUDA's max(double, double) is fmax (a NaN operand is dropped)
    x1 = np.fmax(x1, 0.0)
    arg = np.fmax(4 * x1 * x1 * x1 - x2 * x2, 0.0)
    with np.errstate(divide="ignore", invalid="ignore"):
        phi = np.arctan(np.sqrt(arg) / x2)
    phi = np.where(x2 < 0, phi + np.pi, phi)
    phi = np.where(x2 == 0.0, np.pi / 2.0, phi)
    x1_term = (2.0 / 3.0) * np.sqrt(x1)
    abc = (a + b + c) / 3.0
    lam1 = (abc - x1_term * np.cos(phi / 3.0)).astype(F)
    lam2 = (abc + x1_term * np.cos((phi - np.pi) / 3.0)).astype(F)
    lam3 = (abc + x1_term * np.cos((phi + np.pi) / 3.0)).astype(F)
    if sort == "ascending":
        v1, v2, v3 = lam1, lam2, lam3
    elif sort == "descending":
        v1, v2, v3 = lam3, lam2, lam1
    else:
        raise ValueError(f"unknown sort type: {sort}")
    if abs_sort:
        k1, k2, k3 = np.abs(v1), np.abs(v2), np.abs(v3)  # stale keys
        s = k1 > k2
        v1, v2 = np.where(s, v2, v1), np.where(s, v1, v2)
        s = k1 > k3
        v1, v3 = np.where(s, v3, v1), np.where(s, v1, v3)
        s = k2 > k3
        v2, v3 = np.where(s, v3, v2), np.where(s, v2, v3)
    else:
        s = v1 > v2
        v1, v2 = np.where(s, v2, v1), np.where(s, v1, v2)
        s = v1 > v3
        v1, v3 = np.where(s, v3, v1), np.where(s, v1, v3)
        s = v2 > v3
        v2, v3 = np.where(s, v3, v2), np.where(s, v2, v3)
    if sort == "ascending":
        return np.stack([v1, v2, v3])
    return np.stack([v3, v2, v1])


def symmetric_compute_eigenvalues(S_elems, sort="descending", abs_sort=False):
    """``feature/corner.py:699-755``; nD falls back to eigvalsh."""
    if len(S_elems) == 3:
        if np.asarray(S_elems[0]).dtype.kind != "f":
            raise ValueError("expected real-valued floating point matrices")
        return eig22(*S_elems, sort=sort, abs_sort=abs_sort)
    if len(S_elems) == 6:
        if np.asarray(S_elems[0]).dtype.kind != "f":
            raise ValueError("expected real-valued floating point matrices")
        return eig33(*S_elems, sort=sort, abs_sort=abs_sort)
    image = np.asarray(S_elems[0])
    nd = image.ndim
    mat = np.zeros(image.shape + (nd, nd), dtype=image.dtype)
    for idx, (r, cc) in enumerate(combinations_with_replacement(range(nd), 2)):
        mat[..., r, cc] = S_elems[idx]
        mat[..., cc, r] = S_elems[idx]
    eigs = np.linalg.eigvalsh(mat)
    eigs = np.moveaxis(eigs, -1, 0)
    if abs_sort:
        # CuPy's argsort is a stable (radix) sort: equal magnitudes keep the
        # ascending order of eigvalsh
        eigs = np.take_along_axis(eigs, np.abs(eigs).argsort(0, kind="stable"), 0)
    if sort == "descending":
        eigs = eigs[::-1]
    return eigs


def hessian_matrix_eigvals(H_elems):
    """``feature/corner.py:824-856``: descending, by value."""
    return symmetric_compute_eigenvalues(H_elems)


def frangi(image, sigmas=range(1, 10, 2), scale_range=None, scale_step=None,
           alpha=0.5, beta=0.5, gamma=None, black_ridges=True, mode="reflect",
           cval=0, *, ndi=None, dtype=None, return_gamma=False):
    """``filters/ridges.py:322-476`` with the inner kernels ``:228-319``.

    Per sigma: Hessian (Gaussian derivatives), eigenvalues ascending by
    |value| (stale-key sort), r_g = sum lambda^2 (float type F); gamma from
    the FIRST sigma only: s = max(r_g), gamma_sq = 2.0 if s == 0 else
    max(2 (sqrt(s)/2)^2, 1e-10); vesselness per the 2-D/3-D inner kernel;
    output = running maximum over sigmas (NaN-propagating).
    """
    if scale_range is not None and scale_step is not None:
        sigmas = np.arange(scale_range[0], scale_range[1], scale_step)
    image = np.asarray(image)
    if image.ndim not in (2, 3):
        raise ValueError("The parameter `image` must be a 2-dimensional or 3-dimensional array")
    F = supported_float_type(image.dtype) if dtype is None else np.dtype(dtype)
    image = image.astype(F, copy=False)
    Ft = F.type
    if not black_ridges:
        image = -image
    alpha_sq = 2 * alpha * alpha
    beta_sq = 2 * beta * beta
    gamma_sq = None
    if gamma is not None:
        gamma_sq = 2 * gamma * gamma
    ndim = image.ndim
    filtered_max = None
    for i, sigma in enumerate(sigmas):
        H = hessian_matrix(image, sigma, mode=mode, cval=cval,
                           use_gaussian_derivatives=True, ndi=ndi, dtype=F)
        H = [np.asarray(h).astype(F) for h in H]
        ev = symmetric_compute_eigenvalues(H, sort="ascending", abs_sort=True)
        ev = ev.astype(F)
        r_g = (ev[0] * ev[0]).astype(F)
        for k in range(1, ndim):
            r_g = (r_g + ev[k] * ev[k]).astype(F)
        if gamma_sq is None:
            s_max = float(r_g.max())
            g = math.sqrt(s_max) / 2.0
            gamma_sq = 2.0 if s_max == 0 else max(2 * g * g, 1e-10)
            gamma_used = gamma_sq
        with np.errstate(divide="ignore", over="ignore", invalid="ignore"):
            if ndim == 2:
                lam2 = np.fmax(ev[1], Ft(1e-10))
                r_b = (np.abs(ev[0]) / lam2).astype(F)
                r_b = (r_b * r_b).astype(F)
                res = np.exp(-r_b.astype(np.float64) / beta_sq).astype(F)
                res = (res.astype(np.float64)
                       * (1.0 - np.exp(-r_g.astype(np.float64) / gamma_sq))).astype(F)
            else:
                lam2 = np.fmax(ev[1], Ft(1e-10))
                lam3 = np.fmax(ev[2], Ft(1e-10))
                r_a = (lam2 / lam3).astype(F)
                r_a = (r_a * r_a).astype(F)
                r_b = ((ev[0] * ev[0]).astype(F) / (lam2 * lam3).astype(F)).astype(F)
                res = (1.0 - np.exp(-r_a.astype(np.float64) / alpha_sq)).astype(F)
                res = (res.astype(np.float64) * np.exp(-r_b.astype(np.float64) / beta_sq)).astype(F)
                res = (res.astype(np.float64)
                       * (1.0 - np.exp(-r_g.astype(np.float64) / gamma_sq))).astype(F)
        filtered_max = res.copy() if i == 0 else np.maximum(filtered_max, res)
    if return_gamma:
        return filtered_max, gamma_sq
    return filtered_max


def _ridge_eigvals(image, sigma, mode, cval, ndi, F):
    """hessian_matrix (Gaussian derivatives) + hessian_matrix_eigvals
    (descending), eigenvalues below 1e-12 in magnitude zeroed
    (``ridges.py:123-133, 205-215``)."""
    H = hessian_matrix(image, sigma, mode=mode, cval=cval, use_gaussian_derivatives=True,
                       ndi=ndi, dtype=F)
    ev = hessian_matrix_eigvals([np.asarray(h).astype(F) for h in H]).astype(F)
    ev[np.abs(ev) < 1e-12] = 0
    return ev


def sato(image, sigmas=range(1, 10, 2), black_ridges=True, mode="reflect", cval=0, *,
         ndi=None, dtype=None):
    """``filters/ridges.py:148-225``: per sigma, the geometric mean of the
    eigenvalues other than the smallest, each clipped at 0, times sigma^2;
    pixel-wise maximum over sigmas, starting from 0."""
    image = np.asarray(image)
    if image.ndim not in (2, 3):
        raise ValueError("The parameter `image` must be a 2-dimensional or 3-dimensional array")
    F = supported_float_type(image.dtype) if dtype is None else np.dtype(dtype)
    image = image.astype(F, copy=False)
    if not black_ridges:
        image = -image
    out = np.zeros(image.shape, F)
    for sigma in sigmas:
        ev = _ridge_eigvals(image, sigma, mode, cval, ndi, F)[:-1]
        with np.errstate(invalid="ignore"):
            vals = np.prod(np.maximum(ev, 0), 0) ** (1 / len(ev))
        vals = (vals * F.type(sigma ** 2)).astype(F)
        out = np.maximum(out, vals)
    return out


def meijering(image, sigmas=range(1, 10, 2), alpha=None, black_ridges=True, mode="reflect",
              cval=0, *, ndi=None, dtype=None):
    """``filters/ridges.py:50-145``: normalised eigenvalues
    l_i = e_i + alpha * sum_{j != i} e_j, the one of largest magnitude,
    clipped at 0, divided by its image maximum (when positive); pixel-wise
    maximum over sigmas, starting from 0."""
    image = np.asarray(image)
    F = supported_float_type(image.dtype) if dtype is None else np.dtype(dtype)
    image = image.astype(F, copy=False)
    if not black_ridges:
        image = -image
    if alpha is None:
        alpha = 1 / (image.ndim + 1)
    mtx = np.full((image.ndim, image.ndim), alpha, dtype=F)
    np.fill_diagonal(mtx, 1)
    out = np.zeros(image.shape, F)
    for sigma in sigmas:
        ev = _ridge_eigvals(image, sigma, mode, cval, ndi, F)
        vals = np.tensordot(mtx, ev, 1)
        vals = np.take_along_axis(vals, np.abs(vals).argmax(0)[None], 0).squeeze(0)
        vals = np.maximum(vals, 0)
        mx = vals.max()
        if mx > 0:
            vals = vals / mx
        out = np.maximum(out, vals.astype(F))
    return out


def hessian(image, sigmas=range(1, 10, 2), scale_range=None, scale_step=None, alpha=0.5,
            beta=0.5, gamma=15, black_ridges=True, mode="reflect", cval=0, *, ndi=None,
            dtype=None):
    """``filters/ridges.py:479-565``: frangi (gamma 15 by default), then
    every value <= 0 set to 1."""
    out = frangi(image, sigmas=sigmas, scale_range=scale_range, scale_step=scale_step,
                 alpha=alpha, beta=beta, gamma=gamma, black_ridges=black_ridges, mode=mode,
                 cval=cval, ndi=ndi, dtype=dtype)
    out[out <= 0] = 1
    return out


def difference_of_gaussians(image, low_sigma, high_sigma=None, *, mode="nearest", cval=0,
                            channel_axis=None, truncate=4.0, ndi=None, dtype=None):
    """``filters/_gaussian.py:13-174``: img_as_float, sigma normalisation and
    checks, gaussian(low) - gaussian(high) with preserve_range=False."""
    image = img_as_float(image)
    spatial_dims = image.ndim - 1 if channel_axis is not None else image.ndim
    if np.isscalar(low_sigma):
        low_sigma = (low_sigma,) * spatial_dims
    low_sigma = tuple(map(float, low_sigma))
    if high_sigma is None:
        high_sigma = tuple(low * 1.6 for low in low_sigma)
    elif np.isscalar(high_sigma):
        high_sigma = (high_sigma,) * spatial_dims
    high_sigma = tuple(map(float, high_sigma))
    if len(low_sigma) != 1 and len(low_sigma) != spatial_dims:
        raise ValueError("low_sigma must have length equal to number of spatial dimensions")
    if len(high_sigma) != 1 and len(high_sigma) != spatial_dims:
        raise ValueError("high_sigma must have length equal to number of spatial dimensions")
    if any(s_hi < s_low for s_hi, s_low in zip(high_sigma, low_sigma)):
        raise ValueError("high_sigma must be equal to or larger than low_sigma for all axes")
    kw = dict(mode=mode, cval=cval, channel_axis=channel_axis, truncate=truncate,
              preserve_range=False, ndi=ndi, dtype=dtype)
    out = gaussian(image, sigma=low_sigma, **kw)
    return out - gaussian(image, sigma=high_sigma, **kw)


def _prepare_grayscale_input_nD(image):
    """``feature/util.py:19-24``: squeeze, 2-5 dimensions, img_as_float."""
    image = np.squeeze(np.asarray(image))
    if image.ndim not in range(2, 6):
        raise ValueError("The parameter `image` must be a 2-5-dimensional array")
    return img_as_float(image)


def structure_tensor(image, sigma=1, mode="constant", cval=0, order="rc", *, ndi=None,
                     dtype=None):
    """``feature/corner.py:55-158``: Sobel derivatives along every axis
    (``_compute_derivatives``, ``:27-52``), their pairwise products in
    ``combinations_with_replacement`` order, each Gaussian-smoothed with the
    skimage gaussian (a 3-D image whose last axis has 3 entries is smoothed
    per channel)."""
    ndi_mod = ondi if ndi is None else ndi
    image = np.asarray(image)
    if order == "xy" and image.ndim > 2:
        raise ValueError('Only "rc" order is supported for dim > 2.')
    if order not in ["rc", "xy"]:
        raise ValueError(f'order {order} is invalid. Must be either "rc" or "xy"')
    if not np.isscalar(sigma):
        sigma = tuple(sigma)
        if len(sigma) != image.ndim:
            raise ValueError("sigma must have as many elements as image has axes")
    image = _prepare_grayscale_input_nD(image)
    if dtype is not None:
        image = image.astype(dtype)
    derivatives = [ndi_mod.sobel(image, axis=i, mode=mode, cval=cval)
                   for i in range(image.ndim)]
    if order == "xy":
        derivatives = list(reversed(derivatives))
    channel_axis = -1 if (image.ndim == 3 and image.shape[-1] == 3) else None
    return [gaussian(d0 * d1, sigma=sigma, mode=mode, cval=cval, channel_axis=channel_axis,
                     ndi=ndi, dtype=dtype)
            for d0, d1 in combinations_with_replacement(derivatives, 2)]


def structure_tensor_eigenvalues(A_elems):
    """``feature/corner.py:786-821``: descending, by value."""
    return symmetric_compute_eigenvalues(A_elems)
