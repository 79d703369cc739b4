from .colorconv import rgb2gray  # noqa: F401
