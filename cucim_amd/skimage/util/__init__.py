from .dtype import img_as_float, img_as_float32, img_as_float64  # noqa: F401
