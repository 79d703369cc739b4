"""Drop-in subset of ``cucim.skimage`` on hand-written gfx950 kernels."""

from . import color, feature, filters  # noqa: F401
