"""Pins the oracle's edge filters (CPU only): the known answers of the
reference's filters/tests/test_edges.py, scipy.ndimage.convolve (an
independent implementation of n-D convolution) and the weight tables'
defining properties."""

import math

import numpy as np
import pytest
import scipy.ndimage as sndi

from oracle import ndimage as on
from oracle import skimage as osk

MODES = ["reflect", "mirror", "nearest", "wrap", "constant"]


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("wshape", [(2, 2), (3, 5), (4, 1), (5, 5)])
@pytest.mark.parametrize("origin", [0, -1])
def test_convolve_vs_scipy(mode, wshape, origin):
    if origin and min(wshape) < 2:
        pytest.skip("origin -1 is invalid on a 1-tap axis")
    rng = np.random.default_rng(1)
    x = rng.standard_normal((13, 17))
    w = rng.standard_normal(wshape)
    np.testing.assert_allclose(on.convolve(x, w, mode=mode, cval=0.3, origin=origin),
                               sndi.convolve(x, w, mode=mode, cval=0.3, origin=origin),
                               atol=1e-12)


def test_laplace_kat():
    """test_edges.py:377-395 (test_laplace_zeros)."""
    img = np.zeros((9, 9))
    img[3:-3, 3:-3] = 1
    exp = np.zeros((9, 9))
    exp[2, 3:6] = exp[6, 3:6] = exp[3:6, 2] = exp[3:6, 6] = -1
    exp[3, 3] = exp[3, 5] = exp[5, 3] = exp[5, 5] = 2
    exp[3, 4] = exp[4, 3] = exp[4, 5] = exp[5, 4] = 1
    np.testing.assert_array_equal(osk.laplace(img), exp)
    assert np.all(osk.laplace(img, mask=np.zeros((9, 9), bool)) == 0)
    # the 3-D kernel: 6 at the centre, -1 at its 6 face neighbours
    k = osk.laplacian_kernel(3)
    assert k.sum() == 0 and k[1, 1, 1] == 6 and (k == -1).sum() == 6


def test_roberts_kats():
    """test_edges.py:19-69: diagonal edges give diagonal lines."""
    image = np.tri(10, 10, 0)
    expected = ~(np.tri(10, 10, -1).astype(bool) | np.tri(10, 10, -2).astype(bool).T)
    expected[-1, -1] = 0
    np.testing.assert_array_equal(osk.roberts(image).astype(bool), expected)
    image = np.rot90(np.tri(10, 10, 0), 3)
    expected = ~np.rot90(np.tri(10, 10, -1).astype(bool) |
                         np.tri(10, 10, -2).astype(bool).T)
    np.testing.assert_array_equal(osk.roberts(image).astype(bool), expected)
    assert np.all(osk.roberts(np.zeros((10, 10)), np.ones((10, 10), bool)) == 0)


def test_roberts_vs_scipy():
    x = np.random.default_rng(2).random((20, 23))
    p = sndi.convolve(x, np.array([[1, 0], [0, -1]], float))
    n = sndi.convolve(x, np.array([[0, 1], [-1, 0]], float))
    np.testing.assert_allclose(osk.roberts(x), np.sqrt(p * p + n * n) / math.sqrt(2),
                               rtol=1e-14)


def test_farid_weights_and_kats():
    """The Farid pair (Farid & Simoncelli 2004, Table 1, 5-tap rows, given to
    6 digits there): smoothing sums to 1 and is symmetric, the derivative is
    antisymmetric with a first moment just under 1 (0.9918; the pair is
    optimised for rotation invariance, not for the moment);
    test_edges.py:409-515 line and zero checks."""
    s, e = osk.FARID_SMOOTH, osk.FARID_EDGE
    np.testing.assert_allclose(s, [0.037659, 0.249153, 0.426375, 0.249153, 0.037659],
                               atol=5e-7)
    np.testing.assert_allclose(e, [0.109604, 0.276691, 0, -0.276691, -0.109604], atol=5e-7)
    assert abs(s.sum() - 1) < 1e-14 and np.array_equal(s, s[::-1])
    assert np.array_equal(e, -e[::-1])
    assert 0.99 < -(np.arange(-2, 3) * e).sum() < 1.0
    i, j = np.mgrid[-5:6, -5:6]
    for img, sel in (((i >= 0).astype(float), i), ((j >= 0).astype(float), j)):
        r = osk.farid(img) * math.sqrt(2)
        assert np.all(r[sel == 0] == r[sel == 0][0])
        np.testing.assert_allclose(r[np.abs(sel) > 2], 0, atol=1e-10)
    hi, vj = (i >= 0).astype(float), (j >= 0).astype(float)
    r = osk.farid_h(hi)
    assert np.all(r[i == 0] == r[i == 0][0])
    np.testing.assert_allclose(r[np.abs(i) > 2], 0, atol=1e-10)
    np.testing.assert_allclose(osk.farid_h(vj * math.sqrt(2)), 0, atol=1e-10)
    r = osk.farid_v(vj)
    assert np.all(r[j == 0] == r[j == 0][0])
    np.testing.assert_allclose(osk.farid_v(hi), 0, atol=1e-10)
    x = np.random.default_rng(3).random((12, 14))
    k = e[:, None] * s[None, :]
    np.testing.assert_allclose(osk.farid_h(x), sndi.convolve(x, k), atol=1e-14)
    np.testing.assert_allclose(osk.farid_v(x), sndi.convolve(x, k.T), atol=1e-14)


@pytest.mark.parametrize("name", ["sobel", "scharr", "prewitt"])
def test_directional_kats(name):
    """test_edges.py:89-370: an edge gives a unit line (the smoothing
    weights sum to 1), the orthogonal direction gives 0; the mask-line
    checks of :518-555 (constant 0.2 gradient, masked line and its
    neighbours 0)."""
    fn = getattr(osk, name)
    i, j = np.mgrid[-5:6, -5:6]
    hi, vj = (i >= 0).astype(float), (j >= 0).astype(float)
    r = fn(hi) * math.sqrt(2)
    np.testing.assert_allclose(r[i == 0], 1)
    assert np.all(np.abs(r[np.abs(i) > 1]) < 1e-12)
    r = fn(hi, axis=0)
    np.testing.assert_allclose(r[i == 0], 1)
    np.testing.assert_allclose(fn(vj, axis=0), 0, atol=1e-12)
    np.testing.assert_allclose(fn(hi, axis=1), 0, atol=1e-12)
    vgrad, _ = np.mgrid[:1:11j, :1:11j]
    vgrad[5, :] = 1
    mask = np.ones_like(vgrad)
    mask[5, :] = 0
    exp = np.zeros_like(vgrad)
    exp[1:-1, 1:-1] = 0.2
    exp[4:7, 1:-1] = 0
    np.testing.assert_allclose(fn(vgrad, mask, axis=0), exp, atol=1e-12)
    np.testing.assert_allclose(fn(vgrad.T.copy(), mask.T.copy(), axis=1), exp.T, atol=1e-12)


_SPARSE_IMG = np.array([[0, 0, 1, 3, 5], [0, 1, 4, 3, 4], [1, 2, 5, 4, 1], [2, 4, 5, 2, 1],
                        [4, 5, 1, 0, 0]])
_SPARSE_K = np.array([0, 1, 2, 4, 8, 16, 32, 64, 128]).reshape((3, 3))


def test_correlate_sparse_kats():
    """filters/tests/test_correlate.py:14-76: equal to ndimage.correlate
    ('valid' = the 'wrap' result without its border), for each padding mode
    and input dtype; even kernels raise ValueError."""
    out = osk.correlate_sparse(_SPARSE_IMG.astype(float), _SPARSE_K, mode="valid")
    ref = sndi.correlate(_SPARSE_IMG.astype(float), _SPARSE_K, mode="wrap")[1:4, 1:4]
    np.testing.assert_array_equal(out, ref)
    for mode in ("nearest", "reflect", "mirror"):
        for dt in (np.uint16, np.int32, np.float16, np.float32, np.float64):
            img = _SPARSE_IMG.astype(dt)
            out = osk.correlate_sparse(img, _SPARSE_K, mode=mode)
            assert out.dtype == on.supported_float_type(img.dtype)
            np.testing.assert_array_equal(
                out, sndi.correlate(img.astype(float), _SPARSE_K, mode=mode))
        with pytest.raises(ValueError):
            osk.correlate_sparse(_SPARSE_IMG.astype(float), np.array([[0, 1], [2, 4]]), mode)


@pytest.mark.parametrize("mode", ["nearest", "reflect", "mirror", "wrap", "constant"])
def test_correlate_sparse_vs_scipy(mode):
    """Random sparse kernels in 3-D (corner tap zero or not), in float64:
    scipy.ndimage.correlate with the same boundary mode (constant: 0)."""
    rng = np.random.default_rng(11)
    x = rng.standard_normal((9, 11, 7))
    for corner in (0.0, 1.5):
        k = np.where(rng.random((3, 5, 3)) < 0.3, rng.standard_normal((3, 5, 3)), 0.0)
        k[0, 0, 0] = corner
        np.testing.assert_allclose(osk.correlate_sparse(x, k, mode=mode),
                                   sndi.correlate(x, k, mode=mode, cval=0.0), atol=1e-12)
