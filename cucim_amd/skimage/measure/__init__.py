"""``cucim.skimage.measure`` subset: connected-component labelling."""

from ._label import label  # noqa: F401
