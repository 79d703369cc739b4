"""Image histogram (``exposure/exposure.py:35-345`` of the reference, paths
relative to ``python/cucim/src/cucim/skimage/``) on the device: a range
reduction, then LDS-privatised bin counts."""

from warnings import warn

import numpy as np
import torch

from ... import _array, _lib

__all__ = ["histogram"]


def _range(x):
    """(min, max) of a C-contiguous tensor, as Python numbers (one host
    synchronisation, as the reference's ``int(image.min())`` / ``float(...)``)."""
    dt = _array.np_dtype(x.dtype)
    work = torch.empty(16384, dtype=torch.uint8, device=x.device)
    res = torch.empty(2, dtype=torch.int64, device=x.device)
    _lib.check(_lib.lib().hc_minmax(_array.hc_array(x), res.data_ptr(), work.data_ptr(),
                                    work.numel(), _array.stream()), "min/max")
    h = res.cpu().numpy()
    if dt.kind == "f":
        v = h.view(np.float64)
        return float(v[0]), float(v[1])
    if dt.kind == "u" or dt.kind == "b":
        v = h.view(np.uint64)
        return int(v[0]), int(v[1])
    return int(h[0]), int(h[1])


def _counts(x, nbins, edges=None, lo=0):
    counts = torch.empty(int(nbins), dtype=torch.int64, device=x.device)
    ed = None
    if edges is not None:
        ed = torch.from_numpy(np.ascontiguousarray(edges, dtype=np.float64)).to(x.device)
    _lib.check(_lib.lib().hc_histogram(_array.hc_array(x), int(nbins),
                                       None if ed is None else ed.data_ptr(), int(lo),
                                       counts.data_ptr(), _array.stream()), "histogram")
    return counts


def _float_edges(first, last, nbins, F):
    """cupy.histogram's edges for an int ``bins`` and no range: the image
    range, widened by 0.5 on both sides when it is a single value, then
    ``linspace`` computed as start + i * ((stop - start) / nbins) in float64,
    the last edge set to stop, all rounded to the image's float type."""
    if not (np.isfinite(first) and np.isfinite(last)):
        raise ValueError(f"autodetected range of [{first}, {last}] is not finite")
    if first == last:
        first, last = first - 0.5, last + 0.5
    F = np.dtype(F).type
    first, last = float(F(first)), float(F(last))
    step = (last - first) / nbins
    e = first + np.arange(nbins + 1, dtype=np.float64) * step
    e[-1] = last
    return e.astype(F)


def _dtype_limits(dt):
    dt = np.dtype(dt)
    if dt.kind == "b":
        return 0, 1
    info = np.iinfo(dt)
    return int(info.min), int(info.max)


def _histogram_flat(x, nbins, source_range, normalize):
    """``_histogram`` (:308-345) of a flattened C-contiguous tensor."""
    dt = _array.np_dtype(x.dtype)
    if source_range not in ("image", "dtype"):
        raise ValueError(f"Incorrect value for `source_range` argument: {source_range}")
    if dt.kind in "iub":
        # bincount over [min, max] (image) or the dtype's range (:51-105)
        lo, hi = _range(x) if source_range == "image" else _dtype_limits(dt)
        counts = _counts(x, hi - lo + 1, None, lo)
        centers = torch.arange(lo, hi + 1, dtype=torch.int64, device=x.device)
    else:
        if source_range == "image":
            first, last = _range(x) if x.numel() else (0.0, 1.0)
        else:
            first, last = (-1.0, 1.0)
        e = _float_edges(first, last, int(nbins), dt)
        counts = _counts(x, int(nbins), e)
        ce = (e[:-1] + e[1:]) / dt.type(2.0)
        centers = torch.from_numpy(np.ascontiguousarray(ce.astype(dt))).to(x.device)
    if normalize:
        counts = counts.to(torch.float64) / counts.sum()
    return counts, centers


def histogram(image, nbins=256, source_range="image", normalize=False, *, channel_axis=None):
    """Histogram of an image (``exposure.py:213-305``): integer images count
    every integer value in the range (nbins is ignored), float images use
    ``nbins`` equal bins over the image range.  Returns (counts,
    bin_centers) as device tensors; with ``channel_axis`` the counts of each
    channel over the bins of the whole image."""
    x = _array.as_tensor(image)
    sh = tuple(x.shape)
    if len(sh) == 3 and sh[-1] < 4 and channel_axis is None:
        warn("This might be a color image. The histogram will be computed on the flattened "
             "image. You can instead apply this function to each color channel, or set "
             "channel_axis.", stacklevel=2)
    if channel_axis is None:
        return _histogram_flat(x.contiguous().reshape(-1), nbins, source_range, normalize)
    x = torch.movedim(x, channel_axis, -1)
    flat = x.contiguous().reshape(-1)
    dt = _array.np_dtype(x.dtype)
    # bins of the raveled array, then per-channel counts over them
    if dt.kind in "iub":
        lo, hi = _range(flat) if source_range == "image" else _dtype_limits(dt)
        centers = torch.arange(lo, hi + 1, dtype=torch.int64, device=x.device)
        hists = [_counts(x[..., c].contiguous().reshape(-1), hi - lo + 1, None, lo)
                 for c in range(x.shape[-1])]
    else:
        first, last = _range(flat) if source_range == "image" else (-1.0, 1.0)
        e = np.linspace(first, last, int(nbins) + 1, endpoint=True,
                        dtype=np.result_type(first, last, dt))
        centers = torch.from_numpy(((e[:-1] + e[1:]) / 2.0).astype(dt)).to(x.device)
        hists = [_counts(x[..., c].contiguous().reshape(-1), int(nbins), e)
                 for c in range(x.shape[-1])]
    hist = torch.stack(hists, 0)
    if normalize:
        hist = hist.to(torch.float64) / hist.sum(dim=1, keepdim=True)
    return hist, centers
