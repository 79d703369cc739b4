"""Pins the oracle's histogram, thresholds and labelling (CPU only):
numpy.histogram / numpy.bincount for the counts, the known answers of the
reference's filters/tests/test_thresholding.py and measure/tests/
test_ccomp.py, and a brute-force flood fill for the labels."""

import os
from collections import deque
from itertools import product

import numpy as np
import pytest

from oracle import skimage as osk
from tests import measure_kats as K

HERE = os.path.dirname(os.path.abspath(__file__))


def _img(name):
    return np.load(os.path.join(HERE, "golden", name + ".npy"))


def test_histogram_vs_numpy():
    rng = np.random.default_rng(1)
    for F in (np.float32, np.float64):
        x = (rng.standard_normal(10000) * 3).astype(F)
        c, ce = osk.histogram(x, 256)
        lo, hi = float(x.min()), float(x.max())
        e = np.linspace(lo, hi, 257).astype(F)
        ref, _ = np.histogram(x, bins=e)
        np.testing.assert_array_equal(c, ref)
        assert c.sum() == x.size and ce.dtype == F
    x = rng.integers(-7, 40, 5000).astype(np.int16)
    c, ce = osk.histogram(x)
    np.testing.assert_array_equal(ce, np.arange(x.min(), x.max() + 1))
    np.testing.assert_array_equal(c, np.bincount(x - x.min()))
    c, ce = osk.histogram(np.full(10, 2.5, np.float32), 4)
    assert c.tolist() == [0, 0, 10, 0] and ce[0] == np.float32(2.0) + np.float32(0.125)


def test_threshold_kats():
    """test_thresholding.py:55-125, 301-460."""
    s = K.SIMPLE
    assert osk.threshold_otsu(s) == 2
    assert osk.threshold_otsu(s - 2) == 0
    assert 2 <= osk.threshold_otsu(s.astype(np.float64)) < 3
    assert osk.threshold_yen(s) == 2
    assert osk.threshold_yen(s - 2) == 0
    assert 2 <= osk.threshold_yen(s.astype(np.float64)) < 3
    assert osk.threshold_yen(np.arange(256)) == 127
    b = np.zeros([2, 256], np.uint8)
    b[0] = 255
    assert osk.threshold_yen(b) < 1
    assert osk.threshold_yen(np.zeros((5, 5), np.uint8)) == 0
    assert osk.threshold_yen(np.full((5, 5), 255, np.uint8)) == 255
    assert 2 < osk.threshold_li(s) < 3
    assert 0 < osk.threshold_li(s - 2) < 1
    assert 2 < osk.threshold_li(s.astype(float)) < 3
    assert osk.threshold_li(np.ones((10, 10))) == 1.0
    cam, coins = _img("camera"), _img("coins")
    assert 101 < osk.threshold_otsu(cam) < 103
    assert 106 < osk.threshold_otsu(coins) < 108
    assert 0.41 < osk.threshold_otsu(osk.img_as_float(coins)) < 0.42
    assert 145 < osk.threshold_yen(cam) < 147
    assert 78 < osk.threshold_li(cam) < 79
    assert 94 < osk.threshold_li(coins) < 95
    assert 94 / 255 < osk.threshold_li(osk.img_as_float(coins)) < 95 / 255
    assert osk.threshold_otsu(np.ones((10, 10))) == 1


def _flood(x, background, connectivity):
    """Brute force: BFS over the structure's neighbours, labels in raster
    order of the first element."""
    nd = x.ndim
    offs = [o for o in product((-1, 0, 1), repeat=nd)
            if 0 < sum(map(abs, o)) <= connectivity]
    lab = np.zeros(x.shape, np.int32)
    n = 0
    for idx in np.ndindex(*x.shape):
        if x[idx] == background or lab[idx]:
            continue
        n += 1
        lab[idx] = n
        q = deque([idx])
        while q:
            p = q.popleft()
            for o in offs:
                r = tuple(a + b for a, b in zip(p, o))
                if all(0 <= r[d] < x.shape[d] for d in range(nd)) and not lab[r] \
                        and x[r] == x[p]:
                    lab[r] = n
                    q.append(r)
    return lab, n


def test_label_kats():
    for x, kw, exp in K.label_cases():
        got, _ = osk.label(x, kw.get("background", 0), kw.get("connectivity"))
        np.testing.assert_array_equal(got, exp)
    for x, kw, num in K.NUM_CASES:
        assert osk.label(x, kw.get("background", 0))[1] == num


@pytest.mark.parametrize("shape", [(9, 11), (5, 6, 7), (3, 4, 3, 4)])
def test_label_vs_flood_fill(shape):
    rng = np.random.default_rng(sum(shape))
    x = rng.integers(0, 3, shape)
    for c in range(1, len(shape) + 1):
        got, n = osk.label(x, 0, c)
        exp, m = _flood(x, 0, c)
        np.testing.assert_array_equal(got, exp)
        assert n == m
