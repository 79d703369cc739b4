"""GPU parity of the ndimage layer against the oracle.

float32 weights: bit-exact against the oracle's emulation of the
reference's arithmetic (float32 fused multiply-add per tap, tap order,
zero taps skipped, float32 intermediates between passes).  float64
weights: within 1e-12 relative (the oracle adds separately rounded products
while the kernels fuse them).  Against the FP64 oracle: within the stated
tolerances.
"""

import numpy as np
import pytest
import torch

from oracle import ndimage as on

pytestmark = pytest.mark.gpu

MODES = ["reflect", "mirror", "nearest", "wrap", "constant"]


def _ndi():
    from cucim_amd.skimage._vendored import ndimage as ndi

    return ndi


def _img(shape, dtype, seed=123):
    rng = np.random.default_rng(seed)
    dt = np.dtype(dtype)
    if dt.kind == "b":
        return rng.integers(0, 2, shape).astype(bool)
    if dt.kind in "iu":
        return rng.integers(0, 128, shape).astype(dt)
    return rng.standard_normal(shape).astype(dt)


def _run(fn, x, *a, **k):
    out = fn(torch.from_numpy(np.ascontiguousarray(x)).cuda(), *a, **k)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _exact(got, exp):
    assert got.dtype == exp.dtype, (got.dtype, exp.dtype)
    if got.dtype.kind == "f":
        np.testing.assert_array_equal(got.view(f"u{got.itemsize}"), exp.view(f"u{exp.itemsize}"))
    else:
        np.testing.assert_array_equal(got, exp)


@pytest.mark.parametrize("shape", [(64, 57), (130, 301)])
@pytest.mark.parametrize("axis", [0, 1])
@pytest.mark.parametrize("origin", ["min", 0, "max"])
@pytest.mark.parametrize("K", [1, 2, 3, 4, 5, 8, 13, 16, 17])
@pytest.mark.parametrize("conv", [False, True])
def test_kernel_sizes_and_origins(gpu, shape, axis, origin, K, conv):
    ndi = _ndi()
    o = 0 if K == 1 else (-(K // 2) if origin == "min" else
                          (0 if origin == 0 else K // 2 - (1 if K % 2 == 0 else 0)))
    x = _img(shape, np.float32)
    w = _img((K,), np.float32, seed=7)
    fn = ndi.convolve1d if conv else ndi.correlate1d
    ofn = on.convolve1d if conv else on.correlate1d
    exp = ofn(x, w, axis, mode="nearest", origin=o, acc="emulate")
    for algo in (None, "shared_memory", "elementwise"):
        _exact(_run(fn, x, w, axis, mode="nearest", origin=o, algorithm=algo), exp)


@pytest.mark.parametrize("shape", [(1, 1), (5, 500), (1500, 5), (257, 1029)])
@pytest.mark.parametrize("axis", [-1, -2])
@pytest.mark.parametrize("K", [1, 38, 129])
@pytest.mark.parametrize("mode", MODES + [("constant", 1.0)])
def test_shapes_and_modes(gpu, shape, axis, K, mode):
    ndi = _ndi()
    cval = 0.0
    if isinstance(mode, tuple):
        mode, cval = mode
    x = _img(shape, np.float32)
    w = _img((K,), np.float32, seed=9)
    exp = on.correlate1d(x, w, axis, mode=mode, cval=cval, acc="emulate")
    for algo in (None, "elementwise"):
        _exact(_run(ndi.correlate1d, x, w, axis, mode=mode, cval=cval, algorithm=algo), exp)


@pytest.mark.parametrize("K", [145, 274, 901])
@pytest.mark.parametrize("axis", [0, 1])
def test_long_kernels_fall_back(gpu, K, axis):
    ndi = _ndi()
    x = _img((70, 90), np.float64)
    w = _img((K,), np.float64, seed=3)
    exp = on.correlate1d(x, w, axis, mode="reflect", acc="emulate")
    got = _run(ndi.correlate1d, x, w, axis, mode="reflect")
    np.testing.assert_allclose(got, exp, rtol=1e-12, atol=1e-12)
    if K > 256:
        with pytest.warns(UserWarning):
            got2 = _run(ndi.correlate1d, x, w, axis, mode="reflect", algorithm="shared_memory")
        np.testing.assert_array_equal(got2, got)


@pytest.mark.parametrize("shape", [(16, 24, 32), (9, 5, 33), (4, 70, 3)])
@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("K", [1, 4, 9, 16])
@pytest.mark.parametrize("mode", MODES)
def test_3d(gpu, shape, axis, K, mode):
    ndi = _ndi()
    x = _img(shape, np.float32)
    w = _img((K,), np.float32, seed=11)
    exp = on.correlate1d(x, w, axis, mode=mode, cval=0.5, acc="emulate")
    for algo in (None, "shared_memory", "elementwise"):
        _exact(_run(ndi.correlate1d, x, w, axis, mode=mode, cval=0.5, algorithm=algo), exp)


IMAGE_DTYPES = [np.float16, np.float32, np.float64, bool, np.int8, np.uint8, np.int16,
                np.uint16, np.int32, np.uint32, np.int64, np.uint64]


@pytest.mark.parametrize("axis", [0, 1])
@pytest.mark.parametrize("image_dtype", IMAGE_DTYPES)
@pytest.mark.parametrize("kernel_dtype", [None, np.float32, np.uint8])
def test_image_and_kernel_dtypes(gpu, axis, image_dtype, kernel_dtype):
    ndi = _ndi()
    x = _img((64, 32), image_dtype)
    kd = image_dtype if kernel_dtype is None else kernel_dtype
    w = _img((3,), kd, seed=5)
    exp = on.convolve1d(x, w, axis, mode="nearest", acc="emulate")
    got = _run(ndi.convolve1d, x, w, axis, mode="nearest")
    if np.dtype(on.weights_dtype(x.dtype, w.dtype)) == np.float64 and got.dtype.kind == "f":
        np.testing.assert_allclose(got, exp, rtol=1e-12, atol=1e-12)
    else:
        _exact(got, exp)


@pytest.mark.parametrize("image_dtype", IMAGE_DTYPES)
@pytest.mark.parametrize("output_dtype", [None, np.float32, np.int32, np.uint8])
@pytest.mark.parametrize("prealloc", [False, True])
def test_input_and_output_dtypes(gpu, image_dtype, output_dtype, prealloc):
    ndi = _ndi()
    x = _img((64, 32), image_dtype)
    w = _img((3,), image_dtype, seed=5)
    odt = x.dtype if output_dtype is None else np.dtype(output_dtype)
    exp = on.convolve1d(x, w, 0, output=odt, mode="nearest", acc="emulate")
    xt = torch.from_numpy(x).cuda()
    if prealloc:
        from cucim_amd import _array

        out = _array.empty(x.shape, odt)
        r = ndi.convolve1d(xt, w, 0, output=out, mode="nearest")
        assert r is out
    else:
        r = ndi.convolve1d(xt, w, 0, output=output_dtype, mode="nearest")
    got = r.cpu().numpy()
    if np.dtype(on.weights_dtype(x.dtype, w.dtype)) == np.float64 and got.dtype.kind == "f":
        np.testing.assert_allclose(got, exp, rtol=1e-12, atol=1e-12)
    else:
        _exact(got, exp)


def test_noncontiguous_input_and_output(gpu):
    ndi = _ndi()
    x = _img((40, 60), np.float32)
    w = _img((7,), np.float32, seed=2)
    xt = torch.from_numpy(x).cuda()
    xs = xt[:, ::2]  # strided view
    exp = on.correlate1d(x[:, ::2], w, 1, mode="wrap", acc="emulate")
    for algo in (None, "elementwise"):
        _exact(ndi.correlate1d(xs, w, 1, mode="wrap", algorithm=algo).cpu().numpy(), exp)
    out = torch.zeros((60, 40), dtype=torch.float32, device="cuda").t()  # non-contiguous
    ndi.correlate1d(xt, w, 0, output=out, mode="mirror")
    _exact(out.cpu().numpy(), on.correlate1d(x, w, 0, mode="mirror", acc="emulate"))
    # output aliasing the input
    y = xt.clone()
    ndi.correlate1d(y, w, 1, output=y, mode="reflect")
    _exact(y.cpu().numpy(), on.correlate1d(x, w, 1, mode="reflect", acc="emulate"))


@pytest.mark.parametrize("fn,kw", [
    ("gaussian_filter", dict(sigma=1.5)),
    ("gaussian_filter", dict(sigma=(2.0, 0.7), order=(1, 0))),
    ("gaussian_filter1d", dict(sigma=1.5, axis=0)),
    ("gaussian_filter1d", dict(sigma=1.5, axis=1, order=2)),
    ("uniform_filter", dict(size=7)),
    ("uniform_filter1d", dict(size=4, axis=0)),
    ("sobel", dict(axis=0)),
    ("sobel", dict(axis=1)),
    ("prewitt", dict(axis=0)),
])
@pytest.mark.parametrize("mode", ["nearest", "reflect", "constant", "wrap", "mirror"])
@pytest.mark.parametrize("shape", [(64, 57), (300, 200), (7, 9, 11)])
@pytest.mark.parametrize("fused", [False, True, None])
def test_separable_filters(gpu, fn, kw, mode, shape, fused, monkeypatch):
    from cucim_amd.skimage._vendored import _ndimage_filters as F

    monkeypatch.setattr(F, "FUSED_2D", fused)
    ndi = _ndi()
    x = _img(shape, np.float32)
    if len(shape) == 3 and "sigma" in kw and isinstance(kw["sigma"], tuple):
        pytest.skip("per-axis sigma given for 2-D")
    exp = getattr(on, fn)(x, mode=mode, cval=0.25, acc="emulate", **kw)
    for algo in (None, "elementwise"):
        _exact(_run(getattr(ndi, fn), x, mode=mode, cval=0.25, algorithm=algo, **kw), exp)


def test_correlate_nd(gpu):
    ndi = _ndi()
    x = _img((20, 17, 9), np.float32)
    w = _img((3, 4, 2), np.float32, seed=4)
    w[1, 2, 0] = 0
    for mode in MODES:
        exp = on.correlate(x, w, mode=mode, cval=1.0, origin=(0, -1, 0), acc="emulate")
        _exact(_run(ndi.correlate, x, w, mode=mode, cval=1.0, origin=(0, -1, 0)), exp)
    # convolution: flipped weights, origin -> -origin - (size even)
    exp = on.correlate(x, w[::-1, ::-1, ::-1], mode="reflect", origin=(0, -1, -1),
                       acc="emulate")
    _exact(_run(ndi.convolve, x, w, mode="reflect"), exp)


def test_gaussian_2048_within_1e6_of_fp64(gpu):
    """The 2-D Gaussian workload (2048^2 U[0,1), sigma 2, reflect, truncate 4)
    against an FP64 run: max-abs <= 1e-6."""
    import scipy.ndimage as sndi

    rng = np.random.default_rng(0)
    x = rng.random((2048, 2048), dtype=np.float32)
    ref = sndi.gaussian_filter(x.astype(np.float64), 2.0, mode="reflect", truncate=4.0)
    got = _run(_ndi().gaussian_filter, x, 2.0, mode="reflect", truncate=4.0)
    assert np.abs(got.astype(np.float64) - ref).max() <= 1e-6


def test_errors(gpu):
    ndi = _ndi()
    x = torch.zeros((4, 5), device="cuda")
    with pytest.raises(RuntimeError):
        ndi.correlate1d(x, [1.0, 2.0], mode="bogus")
    with pytest.raises(ValueError):
        ndi.correlate1d(x, [1.0, 2.0, 3.0], origin=2)
    with pytest.raises(RuntimeError):
        ndi.correlate1d(x, np.ones((2, 2)))
    with pytest.raises(ValueError):
        ndi.correlate1d(x, [1.0], algorithm="bogus")
    with pytest.raises(NotImplementedError):
        ndi.correlate1d(torch.zeros((4, 5, 6, 7), device="cuda"), [1.0, 2.0],
                        algorithm="shared_memory")
    with pytest.raises(NotImplementedError):
        ndi.correlate1d(torch.zeros((4, 5), dtype=torch.int32, device="cuda"), [1.0, 2.0],
                        mode="constant", cval=float("nan"))


def _same_with_nan(got, exp):
    """Bit-exact on finite values; NaN where the oracle has NaN (NaN
    payloads differ between the host and the GPU)."""
    assert got.dtype == exp.dtype
    gn, en = np.isnan(got), np.isnan(exp)
    np.testing.assert_array_equal(gn, en)
    u = f"u{got.itemsize}"
    np.testing.assert_array_equal(got[~gn].view(u), exp[~en].view(u))


@pytest.mark.parametrize("shape", [(40, 64), (33, 57)])
@pytest.mark.parametrize("fn,kw", [
    ("sobel", dict(axis=0)),
    ("sobel", dict(axis=1)),
    ("correlate1d", dict(weights=[0.5, 0.0, -0.25, 0.0, 1.0], axis=1)),
    ("correlate1d", dict(weights=[0.5, 0.0, -0.25, 0.0, 1.0], axis=0)),
    ("gaussian_filter", dict(sigma=1.0)),
])
def test_non_finite_inputs(gpu, shape, fn, kw):
    """Zero taps are skipped as in the reference, so an inf/NaN input only
    reaches the outputs whose non-zero taps cover it."""
    ndi = _ndi()
    x = _img(shape, np.float32)
    x[5, 7] = np.inf
    x[20, 30] = -np.inf
    x[11, 3] = np.nan
    exp = getattr(on, fn)(x, mode="reflect", acc="emulate", **kw)
    from cucim_amd.skimage._vendored import _ndimage_filters as F

    for fused in (False, True):
        F.FUSED_2D = fused
        try:
            _same_with_nan(_run(getattr(ndi, fn), x, mode="reflect", **kw), exp)
        finally:
            F.FUSED_2D = None


@pytest.mark.parametrize("seg", [256, 512, 2048, 4096])
def test_tuned_contig_segment(gpu, seg):
    """Every segment length the tuning knob accepts gives the same bits."""
    from cucim_amd import _lib

    ndi = _ndi()
    lib = _lib.lib()
    x = _img((5, 9000), np.float32)
    w = _img((17,), np.float32, seed=9)
    exp = on.correlate1d(x, w, 1, mode="mirror", acc="emulate")
    _lib.check(lib.hc_tune_set(_lib.TUNE_CONTIG_SEG, seg))
    try:
        _exact(_run(ndi.correlate1d, x, w, 1, mode="mirror"), exp)
    finally:
        lib.hc_tune_set(_lib.TUNE_CONTIG_SEG, 0)


@pytest.mark.parametrize("ty", [8, 16, 32, 64])
@pytest.mark.parametrize("shape", [(300, 260), (257, 131)])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_tuned_sep2d_rows(gpu, ty, shape, dtype, monkeypatch):
    """float32: the register-blocked kernel; float64: the LDS tile kernel."""
    from cucim_amd import _lib
    from cucim_amd.skimage._vendored import _ndimage_filters as F

    monkeypatch.setattr(F, "FUSED_2D", True)

    ndi = _ndi()
    lib = _lib.lib()
    x = _img(shape, dtype)
    exp = on.gaussian_filter(x, 2.0, mode="constant", cval=0.5, acc="emulate")
    _lib.check(lib.hc_tune_set(_lib.TUNE_SEP2D_TY, ty))
    try:
        got = _run(ndi.gaussian_filter, x, 2.0, mode="constant", cval=0.5)
        if dtype == np.float32:
            _exact(got, exp)
        else:
            np.testing.assert_allclose(got, exp, rtol=1e-12, atol=1e-12)
    finally:
        lib.hc_tune_set(_lib.TUNE_SEP2D_TY, 0)


def test_tuning_knob_errors(gpu):
    from cucim_amd import _lib

    lib = _lib.lib()
    assert lib.hc_tune_set(_lib.TUNE_CONTIG_SEG, 300) == _lib.ERR_VALUE
    assert lib.hc_tune_set(_lib.TUNE_SEP2D_TY, 24) == _lib.ERR_VALUE
    assert lib.hc_tune_set(99, 1) == _lib.ERR_VALUE
    assert lib.hc_tune_get(_lib.TUNE_CONTIG_SEG) == 0


@pytest.mark.parametrize("shape", [(1, 1), (1, 300), (300, 1), (5, 7), (17, 250), (33, 241),
                                   (64, 480), (100, 1001)])
@pytest.mark.parametrize("sigma", [(1.0, 3.0), (4.0, 0.5), 2.0, (7.9, 7.9)])
@pytest.mark.parametrize("mode", MODES)
def test_register_blocked_2d(gpu, shape, sigma, mode, monkeypatch):
    """The register-blocked fused kernel: uneven kernel lengths (zero-padded
    to a compiled length), all modes, shapes smaller than a tile."""
    from cucim_amd.skimage._vendored import _ndimage_filters as F

    monkeypatch.setattr(F, "FUSED_2D", None)
    ndi = _ndi()
    x = _img(shape, np.float32, seed=5)
    exp = on.gaussian_filter(x, sigma, mode=mode, cval=-0.75, acc="emulate")
    _exact(_run(ndi.gaussian_filter, x, sigma, mode=mode, cval=-0.75), exp)


@pytest.mark.parametrize("size,origin", [((3, 9), 0), ((9, 3), (-4, 1)), ((1, 5), (0, 2)),
                                         ((4, 6), (-2, 2)), ((5, 1), (1, 0))])
@pytest.mark.parametrize("mode", ["wrap", "constant", "mirror"])
def test_register_blocked_uniform(gpu, size, origin, mode):
    ndi = _ndi()
    x = _img((70, 90), np.float32, seed=6)
    exp = on.uniform_filter(x, size, mode=mode, cval=2.0, origin=origin, acc="emulate")
    _exact(_run(ndi.uniform_filter, x, size, mode=mode, cval=2.0, origin=origin), exp)


@pytest.mark.parametrize("fn,kw", [
    ("laplace", {}),
    ("gaussian_laplace", dict(sigma=1.3)),
    ("gaussian_laplace", dict(sigma=(0.8, 2.0), axes=(0, -1))),
    ("gaussian_gradient_magnitude", dict(sigma=1.1)),
    ("gaussian_gradient_magnitude", dict(sigma=2.0, axes=(1,))),
])
@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("shape", [(64, 57), (7, 9, 11)])
def test_laplace_family(gpu, fn, kw, mode, shape):
    """laplace / gaussian_laplace / gaussian_gradient_magnitude
    (_ndimage_filters.py:806-1158): float32 bit-exact to the oracle's
    emulation (per-axis passes, then adds / squares rounded in float32);
    float64 within 1e-12 relative."""
    ndi = _ndi()
    if len(shape) == 3 and isinstance(kw.get("sigma"), tuple):
        pytest.skip("two sigmas for two of three axes")
    x = _img(shape, np.float32)
    exp = getattr(on, fn)(x, mode=mode, cval=0.25, acc="emulate", **kw)
    for algo in (None, "elementwise"):
        _exact(_run(getattr(ndi, fn), x, mode=mode, cval=0.25, algorithm=algo, **kw), exp)
    x64 = _img(shape, np.float64)
    exp64 = getattr(on, fn)(x64, mode=mode, cval=0.25, **kw)
    got64 = _run(getattr(ndi, fn), x64, mode=mode, cval=0.25, **kw)
    np.testing.assert_allclose(got64, exp64, rtol=1e-12, atol=1e-12 * np.abs(exp64).max())


@pytest.mark.parametrize("out_dtype", [None, np.int16, np.int32, np.float32])
def test_laplace_integer_input_and_output(gpu, out_dtype):
    """uint8 input: the output defaults to uint8 (wrapping adds, saturating
    float -> integer conversion of each axis's pass), as in the reference."""
    ndi = _ndi()
    x = np.random.default_rng(31).integers(0, 256, (40, 33)).astype(np.uint8)
    _exact(_run(ndi.laplace, x, output=out_dtype),
           on.laplace(x, output=out_dtype, acc="emulate"))
    _exact(_run(ndi.gaussian_gradient_magnitude, x, 1.5, output=out_dtype),
           on.gaussian_gradient_magnitude(x, 1.5, output=out_dtype, acc="emulate"))


def test_laplace_family_preallocated_strided_output(gpu):
    """A preallocated non-contiguous output receives the same values."""
    ndi = _ndi()
    x = _img((30, 40), np.float32)
    exp = on.gaussian_laplace(x, 1.2, acc="emulate")
    big = torch.zeros((30, 80), dtype=torch.float32, device="cuda")
    out = big[:, ::2]
    r = ndi.gaussian_laplace(torch.from_numpy(x).cuda(), 1.2, output=out)
    assert r is out
    _exact(out.cpu().numpy(), exp)
    exp2 = on.gaussian_gradient_magnitude(x, 1.2, acc="emulate")
    ndi.gaussian_gradient_magnitude(torch.from_numpy(x).cuda(), 1.2, output=out)
    _exact(out.cpu().numpy(), exp2)


@pytest.mark.parametrize("dt", [np.uint8, np.int8, np.int16, np.uint16, np.int32, np.int64])
def test_sqrt_and_square_acc_integers(gpu, dt):
    """The integer element kernels: squares and sums wrap modulo 2^bits, the
    in-place sqrt truncates (negative -> 0), against the oracle's numpy
    restatement over every value of the 8/16-bit types."""
    from cucim_amd import _array, _lib

    info = np.iinfo(dt)
    if info.bits <= 16:
        a = np.arange(info.min, info.max + 1, dtype=np.int64).astype(dt)
    else:
        a = np.random.default_rng(9).integers(info.min // 2, info.max // 2, 70000).astype(dt)
    b = np.roll(a, 7)
    ta, tb = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    lib = _lib.lib()
    _lib.check(lib.hc_square_acc(_array.hc_array(ta), _array.hc_array(ta), 1,
                                 _array.stream()), "sq")
    _lib.check(lib.hc_square_acc(_array.hc_array(ta), _array.hc_array(tb), 0,
                                 _array.stream()), "sq")
    exp = on._wrap_or_round(on._wrap_or_round(a, a, "mul"), on._wrap_or_round(b, b, "mul"),
                            "add")
    _exact(ta.cpu().numpy(), exp)
    tc = torch.from_numpy(a).cuda()
    _lib.check(lib.hc_sqrt_inplace(_array.hc_array(tc), _array.stream()), "sqrt")
    _exact(tc.cpu().numpy(), on._sqrt_unsafe(a))
