"""``cucim.skimage.filters`` subset on the MI355X kernels."""

from .._shared.filters import gaussian  # noqa: F401
from ._gaussian import difference_of_gaussians  # noqa: F401
from .edges import farid, prewitt, scharr, sobel  # noqa: F401
from .ridges import frangi, hessian, meijering, sato  # noqa: F401
