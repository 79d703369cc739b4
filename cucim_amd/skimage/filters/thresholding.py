"""Global thresholds (``filters/thresholding.py`` of the reference, paths
relative to ``python/cucim/src/cucim/skimage/``): threshold_otsu,
threshold_yen and threshold_li.

As in the reference, the histogram (or, for Li on float images, each
iteration's class means) is computed on the device; the per-bin arithmetic on
the 256-entry histogram runs on the host, where the reference hands it to
scikit-image's CPU functions (scikit-image >= 0.23, whose formulas are
restated here: Otsu's between-class variance over cumulative sums, Yen's
maximum-correlation criterion)."""

import math
from warnings import warn

import numpy as np
import torch

from ... import _array, _lib, _trace
from ..exposure.exposure import _range, histogram

__all__ = ["threshold_otsu", "threshold_yen", "threshold_li"]


def _scalar(v, like_dtype=None):
    """A 0-d device tensor, as the reference's cp.asarray(threshold)."""
    a = np.asarray(v)
    if like_dtype is not None:
        a = a.astype(like_dtype)
    return torch.from_numpy(np.array(a)).to(_array.device())


def _validate_image_histogram(image, hist, nbins=None, normalize=False):
    """``thresholding.py:298-352``: counts (float32) and bin centers, from
    ``hist`` (trimmed of leading / trailing empty bins) or the image's
    histogram; returned as host arrays (the reference moves them to the host
    before the per-bin arithmetic)."""
    if image is None and hist is None:
        raise Exception("Either image or hist must be provided.")
    if hist is not None:
        if isinstance(hist, (tuple, list)):
            counts, centers = hist
            counts, centers = _host(counts), _host(centers)
        else:
            counts = _host(hist)
            centers = np.arange(counts.size)
        if counts[0] == 0 or counts[-1] == 0:
            cond = counts > 0
            start = int(np.argmax(cond))
            end = cond.size - int(np.argmax(cond[::-1]))
            counts, centers = counts[start:end], centers[start:end]
    else:
        counts, centers = histogram(_array.as_tensor(image).reshape(-1), nbins,
                                    source_range="image", normalize=normalize)
        counts, centers = _host(counts), _host(centers)
    return counts.astype(np.float32, copy=False), centers


def _host(a):
    if isinstance(a, torch.Tensor):
        return a.detach().cpu().numpy()
    return np.asarray(a)


def _single_value(image):
    """(True, value) when every element equals the first one (the
    reference's ``cp.all(image == first_pixel)``), via the device range."""
    x = _array.as_tensor(image).contiguous().reshape(-1)
    lo, hi = _range(x)
    if lo == hi:  # NaN compares unequal
        return True, x[0].clone()
    return False, None


def _otsu_from_hist(counts, centers):
    """Otsu (scikit-image's threshold_otsu on a histogram): the bin center
    that maximises weight1 * weight2 * (mean1 - mean2)^2 over the split
    after each bin."""
    weight1 = np.cumsum(counts)
    weight2 = np.cumsum(counts[::-1])[::-1]
    mean1 = np.cumsum(counts * centers) / weight1
    mean2 = (np.cumsum((counts * centers)[::-1]) / weight2[::-1])[::-1]
    variance12 = weight1[:-1] * weight2[1:] * (mean1[:-1] - mean2[1:]) ** 2
    return centers[np.argmax(variance12)]


@_trace.annotate("threshold_otsu")
def threshold_otsu(image=None, nbins=256, *, hist=None):
    """Otsu's threshold (``thresholding.py:355-416``)."""
    if image is not None:
        image = _array.as_tensor(image)
        if image.dim() > 2 and image.shape[-1] in (3, 4):
            warn(f"threshold_otsu is expected to work correctly only for grayscale images; "
                 f"image shape {tuple(image.shape)} looks like that of an RGB image.",
                 stacklevel=2)
        same, v = _single_value(image)
        if same:
            return v
    counts, centers = _validate_image_histogram(image, hist, nbins)
    with np.errstate(divide="ignore", invalid="ignore"):
        return _scalar(_otsu_from_hist(counts, centers))


@_trace.annotate("threshold_yen")
def threshold_yen(image=None, nbins=256, *, hist=None):
    """Yen's threshold (``thresholding.py:419-474``)."""
    counts, centers = _validate_image_histogram(image, hist, nbins)
    if centers.size == 1:
        return _scalar(centers[0])
    pmf = counts.astype(np.float32, copy=False) / counts.sum()
    P1 = np.cumsum(pmf)
    P1_sq = np.cumsum(pmf ** 2)
    P2_sq = np.cumsum(pmf[::-1] ** 2)[::-1]
    with np.errstate(divide="ignore", invalid="ignore"):
        crit = np.log(((P1_sq[:-1] * P2_sq[1:]) ** -1) * (P1[:-1] * (1.0 - P1[:-1])) ** 2)
    return _scalar(centers[crit.argmax()])


@_trace.annotate("threshold_li")
def threshold_li(image, *, tolerance=None, initial_guess=None, iter_callback=None):
    """Li's iterative minimum cross entropy threshold
    (``thresholding.py:614-782``).  NaN values are ignored; a single-valued
    image returns that value; with both infinities only, 0.  Integer images
    iterate on their histogram; float images on the device's foreground /
    background means of each step."""
    x = _array.as_tensor(image).contiguous().reshape(-1)
    dt = _array.np_dtype(x.dtype)
    if dt.kind == "f" and dt != np.float64 and dt != np.float32:
        x = x.to(torch.float32)
        dt = np.dtype(np.float32)
    xs = _host_finite_stats(x, dt)
    if xs is None:
        return _scalar(np.nan)
    kind, val = xs
    if kind == "single":
        return _scalar(val)
    if kind == "no_finite":
        return 0.0
    image_min, image_max, uniq_gap = val
    is_int = dt.kind in "iub"
    if is_int:
        tolerance = tolerance or 0.5
    else:
        tolerance = tolerance or float(uniq_gap / 2)
    vmax = image_max - image_min
    if is_int:
        # histogram of the shifted image = counts of x over [min, max] with
        # centers 0 .. max - min
        from ..exposure.exposure import _counts

        hist = _host(_counts(x, image_max - image_min + 1, None, image_min)).astype(np.float32)
        centers = np.arange(0, image_max - image_min + 1)
    if initial_guess is None:
        if is_int:
            t_next = float(np.dot(centers.astype(np.float64), hist.astype(np.float64))
                           / hist.astype(np.float64).sum())
        else:
            t_next = _mean_shifted(x, dt, image_min)
    elif callable(initial_guess):
        t_next = initial_guess(_shifted(x, dt, image_min))
    elif np.isscalar(initial_guess):
        t_next = initial_guess - float(image_min)
        if not 0 < t_next < vmax:
            raise ValueError(
                f"The initial guess for threshold_li must be within the range of the image. "
                f"Got {initial_guess} for image min {float(image_min)} and max "
                f"{float(image_max)}.")
    else:
        raise TypeError("Incorrect type for `initial_guess`; should be a floating point "
                        "value, or a function mapping an array to a floating point value.")
    t_curr = -2 * tolerance
    if iter_callback is not None:
        iter_callback(t_next + image_min)
    eps = 100 * np.finfo(float).eps
    if is_int:
        while abs(t_next - t_curr) > tolerance:
            t_curr = t_next
            fg = centers > t_curr
            bg = ~fg
            mean_fore = np.average(centers[fg], weights=hist[fg])
            mean_back = np.average(centers[bg], weights=hist[bg])
            if mean_back == 0:
                break
            mean_back, mean_fore = float(mean_back), float(mean_fore)
            t_next = (mean_back - mean_fore) / (math.log(mean_back + eps) -
                                                math.log(mean_fore + eps))
            if iter_callback is not None:
                iter_callback(t_next + image_min)
    else:
        work = torch.empty(32768, dtype=torch.uint8, device=x.device)
        out4 = torch.empty(4, dtype=torch.float64, device=x.device)
        lib = _lib.lib()
        while abs(t_next - t_curr) > tolerance:
            t_curr = t_next
            _lib.check(lib.hc_li_means(_array.hc_array(x), float(image_min), float(t_curr),
                                       out4.data_ptr(), work.data_ptr(), work.numel(),
                                       _array.stream()), "threshold_li")
            sf, cf, sb, cb = (float(v) for v in out4.cpu().numpy())
            mean_fore = _rounded_mean(sf, cf, dt)
            mean_back = _rounded_mean(sb, cb, dt)
            if mean_back == 0:
                break
            t_next = (mean_back - mean_fore) / (math.log(mean_back + eps) -
                                                math.log(mean_fore + eps))
            if iter_callback is not None:
                iter_callback(t_next + image_min)
    return t_next + image_min


def _rounded_mean(s, c, dt):
    """cp.mean of the selected values, as a Python float of the image dtype
    (the reference converts the float32 mean with float())."""
    if c == 0:
        return float("nan")
    return float(np.dtype(dt).type(s / c))


def _shifted(x, dt, image_min):
    """The finite values minus the minimum (the reference's ``image -=
    image_min`` after dropping non-finite values), for a callable
    ``initial_guess``: float images in their dtype, integer images as int64."""
    if dt.kind == "f":
        v = x[torch.isfinite(x)]
        return (v - torch.tensor(image_min, dtype=v.dtype, device=v.device)).contiguous()
    return x.to(torch.int64) - int(image_min)


def _mean_shifted(x, dt, image_min):
    if dt.kind == "f":
        work = torch.empty(32768, dtype=torch.uint8, device=x.device)
        out4 = torch.empty(4, dtype=torch.float64, device=x.device)
        # every finite value is background for t = +inf
        _lib.check(_lib.lib().hc_li_means(_array.hc_array(x), float(image_min), math.inf,
                                          out4.data_ptr(), work.data_ptr(), work.numel(),
                                          _array.stream()), "threshold_li")
        _, _, sb, cb = (float(v) for v in out4.cpu().numpy())
        return _rounded_mean(sb, cb, dt)
    raise AssertionError("integer images take the histogram mean")


def _host_finite_stats(x, dt):
    """The special cases at the top of threshold_li and the values it needs:
    None (no non-NaN values); ('single', v); ('no_finite', None);
    ('range', (min, max, smallest gap between distinct finite values))."""
    if dt.kind == "f":
        v = x[~torch.isnan(x)]
        if v.numel() == 0:
            return None
        lo, hi = _range(v.contiguous())
        if lo == hi:
            return "single", lo
        v = v[torch.isfinite(v)]
        if v.numel() == 0:
            return "no_finite", None
        lo, hi = _range(v.contiguous())
        u = torch.unique(v)
        gap = float((u[1:] - u[:-1]).min()) if u.numel() > 1 else 0.0
        return "range", (lo, hi, gap)
    lo, hi = _range(x)
    if lo == hi:
        return "single", lo
    return "range", (lo, hi, 1.0)
