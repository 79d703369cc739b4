"""Drop-in subset of ``cucim.skimage`` on hand-written gfx950 kernels."""

from . import color, exposure, feature, filters, measure  # noqa: F401
