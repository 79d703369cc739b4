"""``cucim.skimage.exposure`` subset: the histogram behind the thresholds."""

from .exposure import histogram  # noqa: F401
