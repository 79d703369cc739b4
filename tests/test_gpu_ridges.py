"""GPU parity of sato / meijering / the hybrid Hessian filter.

Against the FP64 oracle (scipy's C passes, numpy eigvalsh): float32 input
within rel-L2 1e-4 and max-abs 1e-4 * max|out| (sato) or 1e-2 (meijering,
whose per-pixel argmax may flip where two normalised eigenvalues tie in
magnitude); float64 input within 1e-9 relative.  Plus the invariants of the
reference's filters/tests/test_ridges.py.
"""

import os

import numpy as np
import pytest
import scipy.ndimage as sndi
import torch

from oracle import skimage as osk

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _f(name):
    from cucim_amd.skimage import filters

    return getattr(filters, name)


def _np(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _t(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def _vol(shape, seed):
    rng = np.random.default_rng(seed)
    x = sndi.gaussian_filter(rng.standard_normal(shape), 1.5)
    return (x / np.abs(x).max()).astype(np.float32)


@pytest.mark.parametrize("name", ["sato", "meijering"])
@pytest.mark.parametrize("shape", [(64, 80), (24, 30, 28)])
@pytest.mark.parametrize("mode", ["reflect", "nearest", "constant"])
@pytest.mark.parametrize("black", [True, False])
def test_vs_fp64_oracle(gpu, name, shape, mode, black):
    x = _vol(shape, 11)
    sig = (1.0, 2.0, 3.0)
    got = _np(_f(name)(_t(x), sigmas=sig, black_ridges=black, mode=mode))
    exp = getattr(osk, name)(x.astype(np.float64), sigmas=sig, black_ridges=black, mode=mode,
                             ndi=sndi)
    assert got.dtype == np.float32
    err = np.abs(got - exp)
    rel = np.linalg.norm(got - exp) / max(np.linalg.norm(exp), 1e-30)
    assert rel <= 1e-4, rel
    lim = 1e-4 * np.abs(exp).max() if name == "sato" else 1e-2
    assert err.max() <= lim, (err.max(), lim)


@pytest.mark.parametrize("name", ["sato", "meijering", "hessian"])
def test_float64_input(gpu, name):
    x = _vol((40, 36), 12).astype(np.float64)
    got = _np(_f(name)(_t(x), sigmas=(1.5, 2.5), mode="reflect"))
    exp = getattr(osk, name)(x, sigmas=(1.5, 2.5), mode="reflect", ndi=sndi)
    assert got.dtype == np.float64
    np.testing.assert_allclose(got, exp, rtol=1e-9, atol=1e-9 * np.abs(exp).max())


def test_hessian_filter_vs_oracle(gpu):
    """The hybrid filter is frangi(gamma=15) with every value <= 0 set to 1:
    the Frangi part within the Frangi tolerance of the FP64 oracle, the
    replacement rule bit-exact on the GPU's own Frangi (the rule flips
    voxels whose Frangi value is 0 in one computation and 1e-12 in another,
    so the two filters are not compared directly there)."""
    x = _vol((30, 26, 22), 13)
    fr = _np(_f("frangi")(_t(x), sigmas=(1, 2), black_ridges=False, gamma=15))
    got = _np(_f("hessian")(_t(x), sigmas=(1, 2), black_ridges=False))
    np.testing.assert_array_equal(got, np.where(fr <= 0, np.float32(1), fr))
    exp = osk.frangi(x.astype(np.float64), sigmas=(1, 2), black_ridges=False, gamma=15,
                     ndi=sndi)
    np.testing.assert_allclose(fr, exp, atol=2e-4)
    assert np.all(got > 0)


@pytest.mark.parametrize("shape", [(3, 3), (3, 3, 5)])
def test_null_matrix(gpu, shape):
    black = torch.zeros(shape, dtype=torch.uint8, device="cuda")
    white = 255 - black
    zeros = np.zeros(shape)
    atol = 1e-1 if len(shape) == 3 else 0
    np.testing.assert_allclose(_np(_f("meijering")(black, black_ridges=True)), zeros, atol=atol)
    np.testing.assert_allclose(_np(_f("meijering")(white, black_ridges=False)), zeros,
                               atol=atol)
    np.testing.assert_array_equal(_np(_f("sato")(black, black_ridges=True)), zeros)
    np.testing.assert_array_equal(_np(_f("sato")(white, black_ridges=False)), zeros)
    np.testing.assert_array_equal(_np(_f("hessian")(black, black_ridges=False)), 1)
    np.testing.assert_array_equal(_np(_f("hessian")(white, black_ridges=True)), 1)


@pytest.mark.parametrize("shape", [(5, 5), (5, 5, 5)])
@pytest.mark.parametrize("name", ["sato", "meijering", "hessian"])
def test_energy_decrease(gpu, shape, name):
    black = np.zeros(shape, np.uint8)
    black[(2,) * len(shape)] = 255
    white = 255 - black
    assert _np(_f(name)(_t(black), black_ridges=True)).std() < black.std()
    assert _np(_f(name)(_t(white), black_ridges=False)).std() < white.std()


@pytest.mark.parametrize("name", ["sato", "meijering", "hessian"])
def test_linearity(gpu, name):
    black = np.ones((3, 3, 5), np.uint8)
    for x, br in ((black, True), (255 - black, False)):
        np.testing.assert_allclose(_np(_f(name)(_t(1 * x), black_ridges=br)),
                                   _np(_f(name)(_t(10 * x), black_ridges=br)), atol=1e-3)


@pytest.mark.parametrize("name", ["sato", "meijering", "frangi"])
@pytest.mark.parametrize("dtype", ["float64", "uint8"])
def test_camera_black_white(gpu, name, dtype):
    """test_ridges.py::test_2d_cropped_camera_image."""
    crop = np.load(os.path.join(HERE, "golden", "camera_crop.npy"))
    if dtype == "float64":
        black = crop.astype(np.float64) / 255
        white = 1 - black
        tol = 1e-7
    else:
        # uint8 keeps the 0..255 range (no rescaling); the reference's
        # tolerance.  The fused Hessian subtracts the image's midrange before
        # its first pass (Hessian elements annihilate constants), so the
        # image and its 255-complement stage the same values.
        black, white = crop, 255 - crop
        tol = 1e-5
    a = _np(_f(name)(_t(black), black_ridges=True))
    b = _np(_f(name)(_t(white), black_ridges=False))
    np.testing.assert_allclose(a, b, atol=tol, rtol=tol)
    h = _np(_f("hessian")(_t(black), black_ridges=True))
    np.testing.assert_allclose(h, np.ones_like(h), atol=1 - 1e-7)


@pytest.mark.parametrize("name", ["sato", "meijering", "frangi", "hessian"])
@pytest.mark.parametrize("dtype", [torch.float16, torch.float32, torch.float64])
def test_output_dtype(gpu, name, dtype):
    x = torch.rand((20, 24), device="cuda").to(dtype)
    exp = torch.float64 if dtype == torch.float64 else torch.float32
    assert _f(name)(x, sigmas=(1,)).dtype == exp


@pytest.mark.parametrize("name", ["sato", "meijering", "frangi"])
def test_sigmas_from_generator(gpu, name):
    """sigmas may be any iterable of floats, a one-shot generator included
    (filters/ridges.py:148-225 iterates it once)."""
    x = _vol((40, 36), 14)
    a = _np(_f(name)(_t(x), sigmas=(s for s in (1.0, 2.0)), mode="reflect"))
    b = _np(_f(name)(_t(x), sigmas=[1.0, 2.0], mode="reflect"))
    np.testing.assert_array_equal(a, b)
    assert np.abs(b).max() > 0
