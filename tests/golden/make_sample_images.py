"""Write tests/golden/camera.npy, coins.npy and retina_crop.npz: scikit-image
sample images used by the reference's threshold tests
(filters/tests/test_thresholding.py) and its vesselness notebook
(notebooks/vesselness_example.ipynb: rgb2gray(retina)[200:-200, 200:-200];
the uint8 RGB crop is stored, the conversion is part of the measured path).

The images are read from a local scikit-image installation (data only; no
code of the reference runs)."""

import os
import sys

import numpy as np

SRC = "/opt/conda/lib/python3.9/site-packages/skimage/data"


def main(src=SRC):
    from PIL import Image

    here = os.path.dirname(os.path.abspath(__file__))
    cam = np.asarray(Image.open(os.path.join(src, "camera.png")))
    assert cam.shape == (512, 512) and cam.dtype == np.uint8
    np.save(os.path.join(here, "camera.npy"), cam)
    coins = np.asarray(Image.open(os.path.join(src, "coins.png")))
    assert coins.shape == (303, 384) and coins.dtype == np.uint8
    np.save(os.path.join(here, "coins.npy"), coins)
    ret = np.asarray(Image.open(os.path.join(src, "retina.jpg")))
    assert ret.shape == (1411, 1411, 3) and ret.dtype == np.uint8
    crop = np.ascontiguousarray(ret[200:-200, 200:-200])
    np.savez_compressed(os.path.join(here, "retina_crop.npz"), rgb=crop)
    print("wrote", cam.shape, coins.shape, crop.shape)


if __name__ == "__main__":
    main(*sys.argv[1:])
