"""GPU parity of the skimage edge filters beyond sobel (filters/edges.py:
scharr, farid, the directional *_h / *_v forms, roberts*, laplace,
farid_h / farid_v).

float32 results are bit-exact against the oracle's float32 emulation of the
reference's arithmetic (each pass an FMA chain over the taps in order, every
in-place ufunc step rounded in float32); float64 within 1e-12 of the FP64
oracle; plus the known answers of the reference's
filters/tests/test_edges.py.
"""

import math

import numpy as np
import pytest
import torch

from oracle import ndimage as on
from oracle import skimage as osk

pytestmark = pytest.mark.gpu


def _t(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def _np(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _f(name):
    from cucim_amd.skimage import filters

    return getattr(filters, name)


def _bits(a):
    return a.view(f"u{a.itemsize}")


def _magnitude_emulated(x, edge, smooth, mode, axes):
    """edges.py:218-241 in float32: convolve1d(edge / sqrt(ndim)) along the
    edge axis, convolve1d(smooth) along the others, square, accumulate, sqrt,
    each rounded to float32; one axis: the pass results unsquared."""
    nd = x.ndim
    e = np.asarray(edge, np.float32)
    if len(axes) > 1:
        e = (e / math.sqrt(nd)).astype(np.float32)
    sm = np.asarray(smooth, np.float32)
    out = None
    for i, ea in enumerate(axes):
        a = on.convolve1d(x, e, axis=ea, output=np.float32, mode=mode, acc="emulate")
        for sa in sorted(set(range(nd)) - {ea}):
            a = on.convolve1d(a, sm, axis=sa, output=np.float32, mode=mode, acc="emulate")
        if len(axes) > 1:
            a = a * a
        out = a if out is None else out + a
    return np.sqrt(out) if len(axes) > 1 else out


SMOOTH = {"sobel": np.array([1, 2, 1]) / 4, "scharr": np.array([3, 10, 3]) / 16,
          "prewitt": np.full(3, 1 / 3), "farid": osk.FARID_SMOOTH}
EDGE = {"sobel": [1, 0, -1], "scharr": [1, 0, -1], "prewitt": [1, 0, -1],
        "farid": osk.FARID_EDGE}


@pytest.mark.parametrize("name", ["scharr", "farid", "sobel", "prewitt"])
@pytest.mark.parametrize("mode", ["reflect", "nearest", "constant", "wrap", "mirror"])
@pytest.mark.parametrize("shape,axis", [((37, 64), None), ((70, 1029), None),
                                        ((9, 10, 11), None), ((37, 64), 0), ((37, 64), 1),
                                        ((9, 10, 11), 2)])
def test_edge_filters_bit_exact(gpu, name, mode, shape, axis):
    x = np.random.default_rng(4).standard_normal(shape).astype(np.float32)
    axes = list(range(len(shape))) if axis is None else [axis]
    got = _np(_f(name)(_t(x), mode=mode, axis=axis))
    exp = _magnitude_emulated(x, EDGE[name], SMOOTH[name], mode, axes)
    np.testing.assert_array_equal(_bits(got), _bits(exp))


@pytest.mark.parametrize("name", ["scharr", "farid"])
@pytest.mark.parametrize("dtype", [np.float64, np.uint8, np.float16])
def test_edge_filters_vs_fp64(gpu, name, dtype):
    rng = np.random.default_rng(5)
    x = (rng.random((33, 40)) * (255 if dtype == np.uint8 else 1)).astype(dtype)
    m = rng.random((33, 40)) > 0.2
    got = _np(_f(name)(_t(x), _t(m)))
    exp = getattr(osk, name)(x, m)
    assert got.dtype == exp.dtype
    tol = 1e-12 if got.dtype == np.float64 else 3e-6
    np.testing.assert_allclose(got, exp, rtol=tol, atol=tol)


@pytest.mark.parametrize("base", ["sobel", "scharr", "prewitt", "farid"])
@pytest.mark.parametrize("suffix,axis", [("_h", 0), ("_v", 1)])
def test_directional_forms(gpu, base, suffix, axis):
    """*_h / *_v = the single-axis filter along axis 0 / 1 (2-D only)."""
    x = np.random.default_rng(6).standard_normal((40, 52)).astype(np.float32)
    m = np.random.default_rng(7).random((40, 52)) > 0.1
    fn = _f(base + suffix)
    if base == "farid":
        got = _np(fn(_t(x), mask=_t(m)))
        k = osk.FARID_EDGE[:, None] * osk.FARID_SMOOTH[None, :]
        exp = osk._mask_result(on.convolve(x, (k if axis == 0 else k.T).astype(np.float32),
                                           acc="emulate"), m)
    else:
        got = _np(fn(_t(x), _t(m)))
        exp = osk._mask_result(_magnitude_emulated(x, EDGE[base], SMOOTH[base], "reflect",
                                                   [axis]), m)
    np.testing.assert_array_equal(_bits(got), _bits(exp))
    with pytest.raises(ValueError):
        fn(_t(np.zeros((3, 4, 5), np.float32)))


@pytest.mark.parametrize("name", ["sobel_h", "scharr_h", "prewitt_h"])
def test_horizontal_mask_line(gpu, name):
    """test_edges.py:515-531."""
    vgrad, _ = np.mgrid[:1:11j, :1:11j]
    vgrad[5, :] = 1
    mask = np.ones_like(vgrad)
    mask[5, :] = 0
    exp = np.zeros_like(vgrad)
    exp[1:-1, 1:-1] = 0.2
    exp[4:7, 1:-1] = 0
    np.testing.assert_allclose(_np(_f(name)(_t(vgrad), _t(mask))), exp, atol=1e-12)
    vname = name[:-1] + "v"
    np.testing.assert_allclose(_np(_f(vname)(_t(vgrad.T.copy()), _t(mask.T.copy()))), exp.T,
                               atol=1e-12)


@pytest.mark.parametrize("name", ["roberts", "roberts_pos_diag", "roberts_neg_diag",
                                  "laplace", "farid_h", "farid_v"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.uint8, np.float16])
@pytest.mark.parametrize("shape", [(10, 10), (37, 301)])
def test_convolution_filters(gpu, name, dtype, shape):
    """Filters that are one n-D convolution (plus roberts' element steps):
    float32 bit-exact to the oracle's emulation, float64 within 1e-12."""
    rng = np.random.default_rng(8)
    x = (rng.random(shape) * (255 if dtype == np.uint8 else 1)).astype(dtype)
    m = rng.random(shape) > 0.15
    kw = {"mask": _t(m)}
    got = _np(_f(name)(_t(x), **kw))
    exp = getattr(osk, name)(x, mask=m, acc="emulate")
    assert got.dtype == exp.dtype
    if got.dtype == np.float32:
        np.testing.assert_array_equal(_bits(got), _bits(exp))
    else:
        np.testing.assert_allclose(got, exp, rtol=1e-12, atol=1e-12)


def test_laplace_3d_and_kats(gpu):
    """test_edges.py:377-406, and the 3-D 7-point kernel."""
    img = np.zeros((9, 9))
    img[3:-3, 3:-3] = 1
    np.testing.assert_array_equal(_np(_f("laplace")(_t(img))), osk.laplace(img))
    for dt in (np.float16, np.float32, np.float64):
        r = _f("laplace")(_t(img.astype(dt)), ksize=3, mask=_t(np.zeros((9, 9), bool)))
        assert (_np(r) == 0).all()
        assert r.dtype == (torch.float64 if dt == np.float64 else torch.float32)
    x = np.random.default_rng(9).standard_normal((7, 8, 9)).astype(np.float32)
    np.testing.assert_array_equal(_bits(_np(_f("laplace")(_t(x)))),
                                  _bits(osk.laplace(x, acc="emulate")))
    with pytest.raises(NotImplementedError):
        _f("laplace")(_t(x), ksize=5)


def test_roberts_kats(gpu):
    """test_edges.py:19-69."""
    for dt in (np.float16, np.float32, np.float64):
        r = _f("roberts")(_t(np.zeros((10, 10), dt)), _t(np.ones((10, 10), bool)))
        assert (_np(r) == 0).all()
        image = np.tri(10, 10, 0, dtype=dt)
        expected = ~(np.tri(10, 10, -1).astype(bool) | np.tri(10, 10, -2).astype(bool).T)
        expected[-1, -1] = 0
        np.testing.assert_array_equal(_np(_f("roberts")(_t(image))).astype(bool), expected)
    image = np.rot90(np.tri(10, 10, 0), 3)
    expected = ~np.rot90(np.tri(10, 10, -1).astype(bool) | np.tri(10, 10, -2).astype(bool).T)
    np.testing.assert_array_equal(_np(_f("roberts")(_t(image))).astype(bool), expected)


@pytest.mark.parametrize("name", ["farid", "laplace", "prewitt", "roberts", "scharr", "sobel"])
def test_int_rescaling(gpu, name):
    """test_edges.py:43-57: uint8 input is rescaled to [0, 1] first."""
    crop = np.load(__file__.replace("test_gpu_edges.py", "golden/camera_crop.npy"))
    assert _np(_f(name)(_t(crop))).max() <= 2.0


_SPARSE_IMG = np.array([[0, 0, 1, 3, 5], [0, 1, 4, 3, 4], [1, 2, 5, 4, 1], [2, 4, 5, 2, 1],
                        [4, 5, 1, 0, 0]])
_SPARSE_K = np.array([0, 1, 2, 4, 8, 16, 32, 64, 128]).reshape((3, 3))


def test_correlate_sparse_reference_kats(gpu):
    """filters/tests/test_correlate.py:14-76."""
    import scipy.ndimage as sndi

    cs = _f("correlate_sparse")
    out = _np(cs(_t(_SPARSE_IMG.astype(float)), _SPARSE_K, mode="valid"))
    np.testing.assert_array_equal(
        out, sndi.correlate(_SPARSE_IMG.astype(float), _SPARSE_K, mode="wrap")[1:4, 1:4])
    for mode in ("nearest", "reflect", "mirror"):
        for dt in (np.uint16, np.int32, np.float16, np.float32, np.float64):
            img = _SPARSE_IMG.astype(dt)
            out = _np(cs(_t(img), _SPARSE_K, mode=mode))
            assert out.dtype == on.supported_float_type(img.dtype)
            np.testing.assert_array_equal(
                out, sndi.correlate(img.astype(float), _SPARSE_K, mode=mode))
        with pytest.raises(ValueError):
            cs(_t(_SPARSE_IMG.astype(float)), np.array([[0, 1], [2, 4]]), mode=mode)


@pytest.mark.parametrize("mode", ["nearest", "reflect", "mirror", "wrap", "constant", "valid"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.uint8])
@pytest.mark.parametrize("shape,kshape,density", [((64, 80), (5, 7), 0.3),
                                                  ((9, 11, 13), (3, 5, 3), 0.5),
                                                  ((40, 300), (21, 31), 0.25)])
def test_correlate_sparse_bit_exact(gpu, mode, dtype, shape, kshape, density):
    """Against the oracle's restatement, bit for bit: one rounding per
    product and per add in the float type, taps in C order from the corner
    (the third case has more taps than one launch holds)."""
    rng = np.random.default_rng(12)
    x = (rng.random(shape) * (255 if dtype == np.uint8 else 1)).astype(dtype)
    k = np.where(rng.random(kshape) < density, rng.standard_normal(kshape), 0.0)
    k.flat[rng.integers(0, k.size, 3)] = [1.0, -1.0, 0.5]
    for corner in (0.0, 2.0):
        k[(0,) * len(kshape)] = corner
        got = _np(_f("correlate_sparse")(_t(x), k, mode=mode))
        exp = osk.correlate_sparse(x, k, mode=mode)
        assert got.dtype == exp.dtype and got.shape == exp.shape
        np.testing.assert_array_equal(_bits(got), _bits(exp))


def test_correlate_sparse_strided_input(gpu):
    x = np.random.default_rng(13).standard_normal((30, 41)).astype(np.float32)
    big = _t(np.repeat(x, 2, axis=1))
    view = big[:, ::2]
    k = np.zeros((3, 3))
    k[0, 0], k[2, 1], k[1, 2] = -1.0, 1.0, 0.25
    np.testing.assert_array_equal(_bits(_np(_f("correlate_sparse")(view, k))),
                                  _bits(osk.correlate_sparse(x, k)))
