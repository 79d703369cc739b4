"""``cucim.skimage.filters`` subset on the MI355X kernels."""

from .._shared.filters import gaussian  # noqa: F401
from ._gaussian import difference_of_gaussians  # noqa: F401
from ._sparse import correlate_sparse  # noqa: F401
from .edges import (  # noqa: F401
    farid,
    farid_h,
    farid_v,
    laplace,
    prewitt,
    prewitt_h,
    prewitt_v,
    roberts,
    roberts_neg_diag,
    roberts_pos_diag,
    scharr,
    scharr_h,
    scharr_v,
    sobel,
    sobel_h,
    sobel_v,
)
from .ridges import frangi, hessian, meijering, sato  # noqa: F401
from .thresholding import threshold_li, threshold_otsu, threshold_yen  # noqa: F401
