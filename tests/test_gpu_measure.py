"""GPU parity of the steps after the filter: exposure.histogram,
threshold_otsu / threshold_yen / threshold_li and measure.label.

Histogram counts equal the oracle's (numpy bincount / searchsorted on
cupy.histogram's edges) exactly; thresholds equal the oracle's on the same
histogram exactly (Li: within its own tolerance); labels equal the oracle's
(scipy.ndimage.label per value, renumbered in raster order) exactly, plus the
reference's known answers (filters/tests/test_thresholding.py,
measure/tests/test_ccomp.py).
"""

import os

import numpy as np
import pytest
import torch

from oracle import skimage as osk
from tests import measure_kats as K

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _t(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def _np(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _img(name):
    return np.load(os.path.join(HERE, "golden", name + ".npy"))


def _sk():
    import cucim_amd.skimage as sk

    return sk


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.uint8, np.int16, np.int32,
                                   np.uint16, np.int64])
@pytest.mark.parametrize("n", [1, 1000, 3_000_000])
def test_histogram_vs_oracle(gpu, dtype, n):
    rng = np.random.default_rng(n)
    if np.dtype(dtype).kind == "f":
        x = (rng.standard_normal(n) * 5).astype(dtype)
    else:
        info = np.iinfo(dtype)
        x = rng.integers(max(info.min, -300), min(info.max, 5000), n).astype(dtype)
    c, ce = _sk().exposure.histogram(_t(x), 256)
    ec, ece = osk.histogram(x, 256)
    np.testing.assert_array_equal(_np(c), ec)
    np.testing.assert_array_equal(_np(ce), ece)


def test_histogram_dtype_range_and_channels(gpu):
    x = np.random.default_rng(2).integers(0, 256, (40, 30, 3)).astype(np.uint8)
    c, ce = _sk().exposure.histogram(_t(x), source_range="dtype", channel_axis=-1)
    assert _np(c).shape == (3, 256) and _np(ce).tolist() == list(range(256))
    for k in range(3):
        np.testing.assert_array_equal(_np(c)[k], np.bincount(x[..., k].ravel(), minlength=256))


def test_threshold_kats(gpu):
    """test_thresholding.py:55-125, 301-460 on the device path."""
    f = _sk().filters
    s = K.SIMPLE
    assert float(f.threshold_otsu(_t(s))) == 2
    assert float(f.threshold_otsu(_t(s - 2))) == 0
    assert 2 <= float(f.threshold_otsu(_t(s.astype(np.float64)))) < 3
    assert float(f.threshold_yen(_t(s))) == 2
    assert float(f.threshold_yen(_t(s - 2))) == 0
    assert 2 <= float(f.threshold_yen(_t(s.astype(np.float64)))) < 3
    assert float(f.threshold_yen(_t(np.arange(256)))) == 127
    b = np.zeros([2, 256], np.uint8)
    b[0] = 255
    assert float(f.threshold_yen(_t(b))) < 1
    assert float(f.threshold_yen(_t(np.zeros((5, 5), np.uint8)))) == 0
    assert float(f.threshold_yen(_t(np.full((5, 5), 255, np.uint8)))) == 255
    assert 2 < float(f.threshold_li(_t(s))) < 3
    assert 0 < float(f.threshold_li(_t(s - 2))) < 1
    assert 2 < float(f.threshold_li(_t(s.astype(float)))) < 3
    assert float(f.threshold_li(_t(np.ones((10, 10))))) == 1.0
    cam, coins = _img("camera"), _img("coins")
    assert 101 < float(f.threshold_otsu(_t(cam))) < 103
    assert 106 < float(f.threshold_otsu(_t(coins))) < 108
    cf = _sk().util.img_as_float(_t(coins))
    assert 0.41 < float(f.threshold_otsu(cf)) < 0.42
    assert 145 < float(f.threshold_yen(_t(cam))) < 147
    assert 78 < float(f.threshold_li(_t(cam))) < 79
    assert 94 < float(f.threshold_li(_t(coins))) < 95
    assert 94 / 255 < float(f.threshold_li(cf)) < 95 / 255
    assert float(f.threshold_otsu(_t(np.ones((10, 10))))) == 1
    # test_thresholding.py:394-412
    assert np.isnan(float(f.threshold_li(_t(np.full((5, 5), np.nan)))))
    assert float(f.threshold_li(_t(np.array([np.inf, np.nan])))) == np.inf
    assert float(f.threshold_li(_t(np.array([np.inf, -np.inf])))) == 0
    assert float(f.threshold_li(_t(np.array([8, 8, 8, 8, np.nan])))) == 8


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.uint8, np.uint16])
def test_thresholds_vs_oracle(gpu, dtype):
    """Otsu and Yen equal the oracle's on the same data; Li within its
    tolerance."""
    rng = np.random.default_rng(7)
    base = np.concatenate([rng.normal(60, 10, 50000), rng.normal(170, 25, 30000)])
    x = (base if np.dtype(dtype).kind == "f" else np.clip(base, 0, 255)).astype(dtype)
    f = _sk().filters
    assert float(f.threshold_otsu(_t(x))) == float(osk.threshold_otsu(x))
    assert float(f.threshold_yen(_t(x))) == float(osk.threshold_yen(x))
    li = float(f.threshold_li(_t(x)))
    ref = float(osk.threshold_li(x))
    tol = 0.5 if np.dtype(dtype).kind in "iu" else float(np.min(np.diff(np.unique(x))) / 2)
    assert abs(li - ref) <= 2 * tol + 1e-6 * abs(ref), (li, ref)


def test_otsu_on_vesselness(gpu):
    """The intended use: Otsu on a Frangi response, equal to the oracle's
    Otsu of the same float32 map."""
    from cucim_amd.synthetic import tube_volume

    v = _sk().filters.frangi(tube_volume((48, 48, 48), ntubes=6, seed=3), sigmas=(1, 2),
                             black_ridges=False)
    t = float(_sk().filters.threshold_otsu(v))
    assert t == float(osk.threshold_otsu(_np(v)))


def test_label_kats(gpu):
    label = _sk().measure.label
    for x, kw, exp in K.label_cases():
        got = _np(label(_t(x), **kw))
        assert got.dtype == np.int32
        np.testing.assert_array_equal(got, exp)
    for x, kw, num in K.NUM_CASES:
        assert label(_t(x), return_num=True, **kw)[1] == num


@pytest.mark.parametrize("dtype", [np.uint8, np.int32, np.int64, np.float32, bool])
@pytest.mark.parametrize("shape", [(300, 257), (40, 50, 60), (6, 7, 8, 9), (100000,)])
def test_label_vs_oracle(gpu, dtype, shape):
    rng = np.random.default_rng(len(shape))
    x = rng.integers(0, 3, shape)
    x = (x > 0) if dtype is bool else x.astype(dtype)
    if np.dtype(dtype).kind == "f":
        x = x + np.float32(0.25)  # truncated to the integer key, as astype(intp)
    for c in range(1, len(shape) + 1):
        got, n = _sk().measure.label(_t(x), return_num=True, connectivity=c)
        exp, m = osk.label(x, 0, c)
        np.testing.assert_array_equal(_np(got), exp)
        assert n == m


def test_label_large_mask(gpu):
    """A thresholded vesselness-sized mask (a million voxels, long
    components) against the oracle, and an empty mask."""
    rng = np.random.default_rng(9)
    import scipy.ndimage as sndi

    v = sndi.gaussian_filter(rng.standard_normal((100, 100, 100)), 2.0)
    m = v > 0.02
    got, n = _sk().measure.label(_t(m), return_num=True)
    exp, k = osk.label(m, 0, 3)
    np.testing.assert_array_equal(_np(got), exp)
    assert n == k
    got, n = _sk().measure.label(_t(np.zeros((10, 10), np.uint8)), return_num=True)
    assert n == 0 and not _np(got).any()
