"""Connected-component labelling (``measure/_label.py`` and
``measure/_label_kernels.py`` of the reference, paths relative to
``python/cucim/src/cucim/skimage/``) on the device: union-find with atomic
hooking, then the components ranked in raster order by a block scan."""

import numpy as np
import torch

from ... import _array, _lib, _trace

__all__ = ["label"]


def _background_key(background, dt):
    """The integer key of the background value in the image's element type:
    label_image - background == 0 holds exactly where the element equals
    background wrapped into the dtype (integers) or truncated (floats, which
    the reference converts with astype(intp) first)."""
    dt = np.dtype(dt)
    if dt.kind == "f":
        return int(np.trunc(background))
    if dt.kind == "b":
        return int(background)
    bits = dt.itemsize * 8
    v = int(background) % (1 << bits)
    if dt.kind == "i" and v >= 1 << (bits - 1):
        v -= 1 << bits
    if dt == np.uint64 and v >= 1 << 63:  # keys hold uint64 bit patterns
        v -= 1 << 64
    return v


@_trace.annotate("label")
def label(label_image, background=None, return_num=False, connectivity=None):
    """Label connected regions of equal value (``_label.py:21-124``):
    neighbours within ``connectivity`` (1 .. ndim, default ndim) that hold
    the same value share a label; ``background`` (default 0) is labelled 0;
    labels 1..num follow the raster order of each region's first voxel.
    Returns int32 labels (and num with ``return_num``)."""
    x = _array.as_tensor(label_image)
    ndim = x.dim()
    if connectivity is None:
        connectivity = ndim
    if not 1 <= connectivity <= ndim:
        raise ValueError(f"Connectivity below 1 or above {ndim} is illegal.")
    if background is None:
        background = 0
    dt = _array.np_dtype(x.dtype)
    if dt == np.float16:
        x = x.to(torch.float32)
        dt = np.dtype(np.float32)
    x = x.contiguous()
    out = _array.empty(tuple(x.shape), np.int32)
    lib = _lib.lib()
    n = x.numel()
    work = torch.empty(max(1, int(lib.hc_label_work_elems(n))), dtype=torch.int32,
                       device=x.device)
    count = torch.zeros(1, dtype=torch.int32, device=x.device)
    _lib.check(lib.hc_label(_array.hc_array(x), _background_key(background, dt), int(connectivity),
                            _array.hc_array(out), work.data_ptr(), work.numel(),
                            count.data_ptr(), _array.stream()), "label")
    if return_num:
        return out, int(count.item())
    return out
