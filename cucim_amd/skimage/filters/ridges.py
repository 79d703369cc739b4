"""Ridge filters: Frangi vesselness, Sato tubeness, Meijering neuriteness
and the hybrid Hessian filter (``filters/ridges.py`` of the reference, paths
relative to ``python/cucim/src/cucim/skimage/``)."""

import functools
from warnings import warn

import numpy as np
import torch

from ... import _array, _lib, _trace
from .._shared.utils import _supported_float_type, astype, check_nD
from .._vendored._ndimage_util import _mode_code
from ..feature.corner import (
    _hessian_sequential,
    hessian_matrix,
    _symmetric_compute_eigenvalues,
    _taps_arrays,
    hessian_kernels,
)

__all__ = ["frangi", "hessian", "meijering", "sato"]


class FrangiPlan:
    """Device buffers and per-sigma kernels of one Frangi configuration.

    Re-using a plan across calls of the same shape avoids re-allocating the
    work arrays and re-deriving the kernels (the batched drivers and the
    benchmark do this).
    """

    def __init__(self, shape, float_dtype, sigmas, mode="reflect"):
        self.shape = tuple(shape)
        self.ndim = len(shape)
        self.dtype = np.dtype(float_dtype)
        self.sigmas = list(sigmas)
        self.mode = mode
        nH = 6 if self.ndim == 3 else 3
        self.work = [_array.empty(shape, self.dtype) for _ in range(nH)]
        self.work_hc = _array.hc_arrays(self.work)
        self.scratch = torch.zeros(64, dtype=torch.uint8, device=_array.device())
        self.ts, self.td, self._keep = _plan_taps(self.ndim, self.dtype.str, tuple(self.sigmas))

    def run(self, image, out, alpha=0.5, beta=0.5, gamma=None, black_ridges=True, cval=0,
            center=None):
        """Fused schedule; returns an hc status code.  center: an optional
        one-element float64 device tensor subtracted from the image before
        the first pass instead of the image's own midrange."""
        x = image.contiguous()
        return _lib.lib().hc_frangi(
            _array.hc_array(x), _array.hc_array(out), self.ts, self.td, len(self.sigmas),
            float(alpha), float(beta), 0.0 if gamma is None else float(gamma),
            0 if gamma is None else 1, 1 if black_ridges else 0, _mode_code(self.mode),
            float(cval), self.work_hc, len(self.work), self.scratch.data_ptr(),
            None if center is None else center.data_ptr(), _array.stream())


@functools.lru_cache(maxsize=32)
def _plan_taps(ndim, dtype_str, sigmas):
    """The host-built Gaussian (derivative) tap arrays of every sigma and
    axis, shared by the plans of one (ndim, dtype, sigmas): repeated frangi()
    calls of one configuration then build them once."""
    dtype = np.dtype(dtype_str)
    kern = []
    for s in sigmas:
        kern += hessian_kernels(s, ndim, dtype)
    return _taps_arrays(kern, dtype == np.float64)


def _frangi_sequential(image, sigmas, alpha, beta, gamma, black_ridges, mode, cval, out):
    """Reference pass order per sigma (Hessian by 1-D passes, closed-form
    eigenvalues, vesselness), for images the fused schedule cannot hold."""
    float_dtype = _array.np_dtype(out.dtype)
    img = astype(image, float_dtype)
    if not black_ridges:
        img = -img
    scratch = torch.zeros(64, dtype=torch.uint8, device=_array.device())
    lib = _lib.lib()
    for i, sigma in enumerate(sigmas):
        H = _hessian_sequential(img, sigma, mode, cval, "rc")
        eig = _symmetric_compute_eigenvalues(H, sort="ascending", abs_sort=True)
        E = [eig[k] for k in range(img.dim())]
        rc = lib.hc_frangi_from_eigvals(
            _array.hc_arrays(E), img.dim(), _array.hc_array(out), float(alpha), float(beta),
            0.0 if gamma is None else float(gamma), 0 if gamma is None else 1,
            1 if i == 0 else 0,
            scratch.data_ptr(), _array.stream())
        _lib.check(rc, "frangi")
    return out


@_trace.annotate("frangi")
def frangi(image, sigmas=range(1, 10, 2), scale_range=None, scale_step=None, alpha=0.5,
           beta=0.5, gamma=None, black_ridges=True, mode="reflect", cval=0):
    """Frangi vesselness filter (2-D / 3-D), maximum over ``sigmas``."""
    if scale_range is not None and scale_step is not None:
        warn("Use keyword parameter `sigmas` instead of `scale_range` and "
             "`scale_range` which will be removed in version 0.17.", stacklevel=2)
        sigmas = np.arange(scale_range[0], scale_range[1], scale_step)
    image = _array.as_tensor(image)
    check_nD(image, [2, 3])
    float_dtype = _supported_float_type(image.dtype)
    sigmas = [float(s) for s in sigmas]
    if not sigmas:
        raise ValueError("sigmas must contain at least one value")
    _mode_code(mode)
    out = _array.empty(image.shape, float_dtype)
    if _array.np_dtype(image.dtype) == np.float16:
        image = astype(image, np.float32)
    if mode in ("constant", "grid-constant") and float(cval) != 0.0:
        # the per-axis schedule reorders passes; with nonzero constant
        # padding only the reference order is exact
        return _frangi_sequential(image, sigmas, alpha, beta, gamma, black_ridges, mode,
                                  cval, out)
    plan = FrangiPlan(image.shape, float_dtype, sigmas, mode)
    rc = plan.run(image, out, alpha, beta, gamma, black_ridges, cval)
    if rc == _lib.OK:
        return out
    if rc != _lib.ERR_RESOURCE:
        _lib.check(rc, "frangi")
    del plan
    return _frangi_sequential(image, sigmas, alpha, beta, gamma, black_ridges, mode, cval, out)


def _ridge_input(image, black_ridges):
    image = _array.as_tensor(image)
    float_dtype = _supported_float_type(image.dtype)
    image = astype(image, float_dtype)
    if not black_ridges:
        image = -image
    return image, float_dtype


@_trace.annotate("sato")
def sato(image, sigmas=range(1, 10, 2), black_ridges=True, mode="reflect", cval=0):
    """Sato tubeness filter (``ridges.py:148-225``): per sigma, the geometric
    mean of the positive Hessian eigenvalues other than the smallest, times
    sigma^2; maximum over sigmas (from 0)."""
    check_nD(_array.as_tensor(image), [2, 3])
    sigmas = [float(s) for s in sigmas]  # one pass over a one-shot iterable
    image, float_dtype = _ridge_input(image, black_ridges)
    out = _array.empty(tuple(image.shape), float_dtype)
    lib = _lib.lib()
    for i, sigma in enumerate(sigmas):
        H = hessian_matrix(image, sigma, mode=mode, cval=cval, use_gaussian_derivatives=True)
        _lib.check(lib.hc_sato_fold(_array.hc_arrays(H), len(H), _array.hc_array(out),
                                    float(sigma), 1 if i == 0 else 0, _array.stream()), "sato")
    if not sigmas:
        out.zero_()
    return out


@_trace.annotate("meijering")
def meijering(image, sigmas=range(1, 10, 2), alpha=None, black_ridges=True, mode="reflect",
              cval=0):
    """Meijering neuriteness filter (``ridges.py:50-145``): per sigma, the
    normalised eigenvalue combination of largest magnitude, clipped at 0 and
    scaled by its image maximum; maximum over sigmas (from 0).  2-D and 3-D
    images (the closed-form eigenvalues)."""
    image, float_dtype = _ridge_input(image, black_ridges)
    if image.dim() not in (2, 3):
        raise NotImplementedError("meijering is implemented for 2-D and 3-D images")
    if alpha is None:
        alpha = 1 / (image.dim() + 1)
    out = _array.empty(tuple(image.shape), float_dtype)
    tmp = _array.empty(tuple(image.shape), float_dtype)
    scratch = torch.zeros(64, dtype=torch.uint8, device=_array.device())
    lib = _lib.lib()
    n = 0
    for i, sigma in enumerate(sigmas):
        H = hessian_matrix(image, sigma, mode=mode, cval=cval, use_gaussian_derivatives=True)
        _lib.check(lib.hc_meijering_fold(_array.hc_arrays(H), len(H), _array.hc_array(out),
                                         _array.hc_array(tmp), float(alpha),
                                         1 if i == 0 else 0, scratch.data_ptr(),
                                         _array.stream()), "meijering")
        n += 1
    if n == 0:
        out.zero_()
    return out


@_trace.annotate("hessian")
def hessian(image, sigmas=range(1, 10, 2), scale_range=None, scale_step=None, alpha=0.5,
            beta=0.5, gamma=15, black_ridges=True, mode="reflect", cval=0):
    """Hybrid Hessian filter (``ridges.py:479-565``): Frangi with gamma = 15
    by default, then every value <= 0 set to 1."""
    filtered = frangi(image, sigmas=sigmas, scale_range=scale_range, scale_step=scale_step,
                      alpha=alpha, beta=beta, gamma=gamma, black_ridges=black_ridges,
                      mode=mode, cval=cval)
    _lib.check(_lib.lib().hc_fill_le(_array.hc_array(filtered), 0.0, 1.0, _array.stream()),
               "hessian")
    return filtered
