"""Write tests/golden/camera_crop.npy: the 100 x 100 crop
camera()[200:300, 100:200] used by the reference's ridge-filter tests
(filters/tests/test_ridges.py, ``crop(camera(), ((200, 212), (100, 312)))``).

camera.png is scikit-image's sample image; it is read from a local
scikit-image installation (data only, no code of the reference runs).
"""

import os
import sys

import numpy as np


def main(src=None):
    from PIL import Image

    src = src or "/opt/conda/lib/python3.9/site-packages/skimage/data/camera.png"
    img = np.asarray(Image.open(src))
    assert img.shape == (512, 512) and img.dtype == np.uint8
    crop = np.ascontiguousarray(img[200:300, 100:200])
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "camera_crop.npy")
    np.save(out, crop)
    print(out, crop.shape, crop.dtype)


if __name__ == "__main__":
    main(*sys.argv[1:])
