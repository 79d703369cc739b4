"""Pins the oracle's ndimage layer (CPU only): scipy.ndimage is an
independent C implementation of the same filters; numpy.pad gives each
boundary mode independently; closed forms fix the kernel math."""

from fractions import Fraction

import numpy as np
import pytest
import scipy.ndimage as sndi

from oracle import ndimage as on

MODES = ["reflect", "mirror", "nearest", "wrap", "constant"]
PAD = {"reflect": "symmetric", "mirror": "reflect", "nearest": "edge", "wrap": "wrap",
       "constant": "constant"}


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("n", [1, 2, 3, 7])
def test_remap_matches_numpy_pad(mode, n):
    # widths of several periods exercise the multi-fold extension
    x = np.arange(n) + 10
    width = 3 * n + 2
    if mode == "constant":
        padded = np.pad(x, width, mode="constant", constant_values=-1)
    elif mode == "mirror" and n == 1:
        padded = np.full(n + 2 * width, x[0])
    else:
        padded = np.pad(x, width, mode=PAD[mode])
    pos = np.arange(-width, n + width)
    idx = on.remap_index(pos, n, mode)
    got = np.where(idx >= 0, x[np.where(idx >= 0, idx, 0)], -1)
    assert np.array_equal(got, padded)


@pytest.mark.parametrize("mode", MODES)
def test_correlate1d_vs_scipy(mode):
    rng = np.random.default_rng(0)
    for shape in [(1,), (2,), (7,), (5, 13), (31, 4), (3, 4, 5)]:
        x = rng.standard_normal(shape)
        for K in [1, 2, 3, 5, 8, 17, 40]:
            w = rng.standard_normal(K)
            for origin in range(-(K // 2), K - K // 2):
                for axis in range(x.ndim):
                    a = on.correlate1d(x, w, axis, mode=mode, cval=1.5, origin=origin)
                    b = sndi.correlate1d(x, w, axis, mode=mode, cval=1.5, origin=origin)
                    np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12)


def test_convolve1d_vs_scipy():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((9, 11))
    for K in [1, 2, 3, 4, 7]:
        w = rng.standard_normal(K)
        for origin in range(-(K // 2), K - K // 2):
            for mode in MODES:
                a = on.convolve1d(x, w, 1, mode=mode, origin=origin)
                b = sndi.convolve1d(x, w, 1, mode=mode, origin=origin)
                np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12)


def test_correlate1d_closed_forms():
    w = np.array([0.5, -1.25, 3.0, 2.0, 0.125])
    x = np.zeros(21)
    x[10] = 1.0
    # a delta returns the reversed kernel around it
    y = on.correlate1d(x, w, mode="constant")
    np.testing.assert_array_equal(y[8:13], w[::-1])
    # a constant input returns c * sum(w) in every non-constant mode
    for mode in ["reflect", "mirror", "nearest", "wrap"]:
        y = on.correlate1d(np.full(6, 3.0), w, mode=mode)
        np.testing.assert_allclose(y, 3.0 * w.sum())


def test_correlate_nd_vs_scipy():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((6, 7, 5))
    w = rng.standard_normal((3, 2, 4))
    w[1, 0, 2] = 0.0
    for mode in MODES:
        a = on.correlate(x, w, mode=mode, cval=0.5, origin=(0, -1, 1))
        b = sndi.correlate(x, w, mode=mode, cval=0.5, origin=(0, -1, 1))
        np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("sigma", [0.3, 0.7071, 1.0, 2.5])
@pytest.mark.parametrize("order", [0, 1, 2, 3])
def test_gaussian_kernel1d(sigma, order):
    import scipy.ndimage._filters as sf

    r = int(4 * sigma + 0.5) + 3
    k = on.gaussian_kernel1d(sigma, order, r)
    np.testing.assert_allclose(k, sf._gaussian_kernel1d(sigma, order, r)[::-1],
                               rtol=1e-12, atol=1e-15)
    if order == 1:
        # closed form: d/dx of the normalised gaussian, reversed
        x = np.arange(-r, r + 1)
        phi = np.exp(-0.5 * x**2 / sigma**2)
        phi /= phi.sum()
        np.testing.assert_allclose(k, (-x / sigma**2 * phi)[::-1], rtol=1e-12, atol=1e-15)


def test_gaussian_filter_vs_scipy():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((12, 9, 7))
    for order in [0, 1, (1, 0, 2)]:
        for mode in MODES:
            a = on.gaussian_filter(x, 1.3, order=order, mode=mode, cval=0.25)
            b = sndi.gaussian_filter(x, 1.3, order=order, mode=mode, cval=0.25)
            np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12)


def test_gaussian_docstring_kats(kats):
    from oracle import skimage as osk

    kat = kats["gaussian_delta3x3"]
    a = np.zeros((3, 3))
    a[1, 1] = 1
    for case in kat["cases"]:
        g = osk.gaussian(a, sigma=case["sigma"], mode=case["mode"])
        exp = np.array([[case["corner"], case["edge"], case["corner"]],
                        [case["edge"], case["center"], case["edge"]],
                        [case["corner"], case["edge"], case["corner"]]])
        np.testing.assert_allclose(g, exp, atol=0.5 * 10 ** -kat["decimals"])


def test_sobel_prewitt_vs_scipy():
    rng = np.random.default_rng(4)
    x = rng.standard_normal((8, 10))
    for axis in (0, 1):
        for mode in MODES:
            np.testing.assert_allclose(on.sobel(x, axis, mode=mode),
                                       sndi.sobel(x, axis, mode=mode), rtol=1e-12, atol=1e-12)
            np.testing.assert_allclose(on.prewitt(x, axis, mode=mode),
                                       sndi.prewitt(x, axis, mode=mode), rtol=1e-12, atol=1e-12)


def test_uniform_vs_scipy():
    rng = np.random.default_rng(5)
    x = rng.standard_normal((8, 10))
    for mode in MODES:
        np.testing.assert_allclose(on.uniform_filter(x, 4, mode=mode),
                                   sndi.uniform_filter(x, 4, mode=mode), rtol=1e-12,
                                   atol=1e-12)


def test_fma_f32_exact():
    rng = np.random.default_rng(6)
    n = 4000
    a = rng.standard_normal(n).astype(np.float32)
    b = rng.standard_normal(n).astype(np.float32)
    # near-cancelling addends and exact ties exercise the rounding boundary
    c = (-(a.astype(np.float64) * b)).astype(np.float32)
    c[: n // 2] += rng.standard_normal(n // 2).astype(np.float32) * np.float32(1e-6)
    r = on.fma_f32(a, b, c)
    for i in range(n):
        ex = Fraction(float(a[i])) * Fraction(float(b[i])) + Fraction(float(c[i]))
        f = np.float32(float(ex))
        cand = [np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))]
        best = min(cand, key=lambda v: (abs(Fraction(float(v)) - ex),
                                        int(np.array(v, np.float32).view(np.uint32)) & 1))
        assert best == r[i], i


def test_cast_to_rules():
    v = np.array([-300.7, -1.5, -0.5, 0.0, 0.5, 1.5, 254.9, 300.2, np.nan, np.inf])
    np.testing.assert_array_equal(
        on.cast_to(v, np.int8), [-128, -1, 0, 0, 0, 1, 127, 127, 0, 127])
    # float -> unsigned: a >= 0 ? (B)a : -(B)(-a)  (saturating conversion)
    np.testing.assert_array_equal(
        on.cast_to(v, np.uint8), [1, 255, 0, 0, 0, 1, 254, 255, 0, 255])
    np.testing.assert_array_equal(on.cast_to(v, np.bool_), v != 0)


def test_emulated_accumulation_is_float32_sequential():
    # acc="emulate" must equal an explicit float32 FMA chain in tap order
    rng = np.random.default_rng(7)
    x = rng.standard_normal(40).astype(np.float32)
    w = rng.standard_normal(5).astype(np.float32)
    y = on.correlate1d(x, w, mode="nearest", acc="emulate")
    for i in range(40):
        s = np.float32(0)
        for k in range(5):
            j = min(max(i + k - 2, 0), 39)
            s = on.fma_f32(x[j], w[k], s)
        assert y[i] == s


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("axes", [None, (0, 2), (1,)])
def test_laplace_family_vs_scipy(mode, axes):
    """laplace / gaussian_laplace / gaussian_gradient_magnitude
    (_ndimage_filters.py:806-1158) against scipy.ndimage's own (an
    independent implementation of the same per-axis definitions)."""
    x = np.random.default_rng(5).standard_normal((14, 17, 11))
    kw = dict(mode=mode, cval=0.5, axes=axes)
    np.testing.assert_allclose(on.laplace(x, **kw), sndi.laplace(x, **kw), atol=1e-13)
    np.testing.assert_allclose(on.gaussian_laplace(x, 1.3, **kw),
                               sndi.gaussian_laplace(x, 1.3, **kw), atol=1e-13)
    np.testing.assert_allclose(on.gaussian_gradient_magnitude(x, 0.9, **kw),
                               sndi.gaussian_gradient_magnitude(x, 0.9, **kw), atol=1e-13)


def test_laplace_family_closed_forms():
    """A quadratic has Laplacian 2*sum(a) away from the border; the gradient
    magnitude of a plane is the norm of its slope.  With the kernels
    truncated at 12 sigma, the discrete Gaussian moments equal the continuous
    ones up to the aliasing term exp(-2 pi^2 sigma^2) (< 1e-19 at sigma 1.5)."""
    z, y = np.mgrid[0:60, 0:70].astype(np.float64)
    q = 0.5 * z * z + 1.5 * y * y
    lap = on.laplace(q)
    np.testing.assert_allclose(lap[2:-2, 2:-2], 4.0, atol=1e-9)
    plane = 0.3 * z - 0.4 * y
    g = on.gaussian_gradient_magnitude(plane, 1.5, truncate=12.0)
    np.testing.assert_allclose(g[19:-19, 19:-19], 0.5, rtol=1e-9)
    gl = on.gaussian_laplace(q, 1.5, truncate=12.0)
    np.testing.assert_allclose(gl[19:-19, 19:-19], 4.0, rtol=1e-9)


def test_laplace_integer_output_vs_scipy():
    """Integer outputs: the per-axis results are cast to the output dtype and
    added with wrap-around.  The per-axis values are integers of magnitude
    < 128 here, where the reference's float -> unsigned cast (negatives
    wrap) and scipy's C cast agree; beyond the destination range the device
    conversion saturates, which scipy does not model."""
    x = np.random.default_rng(6).integers(0, 60, (30, 41)).astype(np.uint8)
    for dt in (np.uint8, np.int16, np.int32):
        np.testing.assert_array_equal(on.laplace(x, output=dt), sndi.laplace(x, output=dt))
    # gradient magnitude with integer outputs (squares and sums in the output
    # dtype, then the unsafe in-place sqrt): Sobel derivatives are integers,
    # in range for signed 32/64-bit outputs (unsigned ones would wrap the
    # negative first-pass values, where scipy's cast and the device's part)
    y = np.random.default_rng(7).integers(0, 30, (25, 26)).astype(np.uint16)
    for dt in (np.int32, np.int64):
        got = on.generic_gradient_magnitude_arrays(
            [on.sobel(y, a, output=dt) for a in range(2)])
        ref = sndi.generic_gradient_magnitude(y.astype(dt), sndi.sobel, output=dt)
        np.testing.assert_array_equal(got, ref)
    np.testing.assert_allclose(on.gaussian_gradient_magnitude(y, 1.0, output=np.float64),
                               sndi.gaussian_gradient_magnitude(y, 1.0, output=np.float64),
                               rtol=1e-13)


@pytest.mark.parametrize("dt", [np.uint8, np.int8, np.uint16, np.int16, np.uint32, np.int32])
def test_sqrt_unsafe_exhaustive(dt):
    """The unsafe in-place sqrt of integers equals numpy's own
    sqrt(a, a, casting='unsafe') for every non-negative value (all values of
    the 8/16-bit types; a sample of the wider ones), and maps negatives
    (NaN) to 0."""
    info = np.iinfo(dt)
    if info.bits <= 16:
        a = np.arange(info.min, info.max + 1, dtype=np.int64).astype(dt)
    else:
        r = np.random.default_rng(8)
        a = np.concatenate([r.integers(info.min, info.max, 100000, dtype=np.int64),
                            np.arange(0, 70000), np.array([info.max])]).astype(dt)
    got = on._sqrt_unsafe(a)
    nonneg = a >= 0
    ref = a.copy()
    with np.errstate(invalid="ignore"):
        np.sqrt(ref, out=ref, casting="unsafe")
    np.testing.assert_array_equal(got[nonneg], ref[nonneg])
    assert np.all(got[~nonneg] == 0)
    # floor(sqrt(v)) exactly: the float16 step of the 8-bit types truncates
    # to the integer square root
    v = a[nonneg].astype(np.int64)
    isq = np.floor(np.sqrt(v.astype(np.float64))).astype(np.int64)
    np.testing.assert_array_equal(got[nonneg].astype(np.int64), isq)
