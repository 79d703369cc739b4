"""cucim_amd: MI355X-native separable n-D filtering and Hessian-ridge path.

A drop-in for cuCIM's ``skimage.filters.gaussian / sobel / frangi``,
``skimage.feature.hessian_matrix / hessian_matrix_eigvals``,
``skimage.color.rgb2gray`` and the vendored ``ndimage`` correlate core, on
hand-written HIP kernels for gfx950 behind the C ABI of
``include/cucim_amd.h``.
"""

from . import skimage  # noqa: F401

__version__ = "0.1.0"
