"""Pins of the oracle's sato / meijering / hybrid Hessian filters.

Invariants transcribed from the reference's filters/tests/test_ridges.py
(null matrices, energy decrease, linearity, black/white ridge symmetry on
the camera crop), a closed form (sato of a quadratic ridge), and a
brute-force per-pixel recomputation of the eigenvalue formulas on a tiny
image from scipy's Gaussian derivatives and numpy's eigvalsh.
"""

import itertools
import os

import numpy as np
import pytest
import scipy.ndimage as sndi

from oracle import skimage as osk

HERE = os.path.dirname(os.path.abspath(__file__))
FUNCS = {"sato": osk.sato, "meijering": osk.meijering, "frangi": osk.frangi,
         "hessian": osk.hessian}


def _call(name, x, **kw):
    if name in ("sato", "hessian"):
        kw.setdefault("mode", "reflect")
    return FUNCS[name](x, **kw)


@pytest.mark.parametrize("shape", [(3, 3), (3, 3, 5)])
@pytest.mark.parametrize("name", ["sato", "meijering", "frangi", "hessian"])
def test_null_matrix(shape, name):
    black = np.zeros(shape, np.uint8)
    white = 255 - black
    exp = np.ones(shape) if name == "hessian" else np.zeros(shape)
    atol = 1e-3 if name == "frangi" else (1e-1 if name == "meijering" and len(shape) == 3
                                          else 0)
    if name == "hessian":
        np.testing.assert_array_equal(_call(name, black, black_ridges=False), exp)
        np.testing.assert_array_equal(_call(name, white, black_ridges=True), exp)
    else:
        np.testing.assert_allclose(_call(name, black, black_ridges=True), exp, atol=atol)
        np.testing.assert_allclose(_call(name, white, black_ridges=False), exp, atol=atol)


@pytest.mark.parametrize("shape", [(5, 5), (5, 5, 5)])
@pytest.mark.parametrize("name", ["sato", "meijering", "frangi", "hessian"])
def test_energy_decrease(shape, name):
    black = np.zeros(shape, np.uint8)
    black[(2,) * len(shape)] = 255
    white = 255 - black
    assert _call(name, black, black_ridges=True).std() < black.std()
    assert _call(name, white, black_ridges=False).std() < white.std()


@pytest.mark.parametrize("shape", [(3, 3), (3, 3, 5)])
@pytest.mark.parametrize("name", ["sato", "meijering", "frangi", "hessian"])
def test_linearity(shape, name):
    black = np.ones(shape, np.uint8)
    white = 255 - black
    for x, br in ((black, True), (white, False)):
        np.testing.assert_allclose(_call(name, 1 * x, black_ridges=br),
                                   _call(name, 10 * x, black_ridges=br), atol=1e-3)


@pytest.mark.parametrize("name", ["sato", "meijering", "frangi"])
def test_camera_black_white_symmetry(name):
    """test_ridges.py::test_2d_cropped_camera_image (float64 case)."""
    crop = np.load(os.path.join(HERE, "golden", "camera_crop.npy"))
    black = crop.astype(np.float64) / 255
    white = 1 - black
    kw = dict(ndi=sndi)
    np.testing.assert_allclose(_call(name, black, black_ridges=True, **kw),
                               _call(name, white, black_ridges=False, **kw),
                               atol=1e-7, rtol=1e-7)
    h = osk.hessian(black, black_ridges=True, mode="reflect", ndi=sndi)
    np.testing.assert_allclose(h, np.ones_like(h), atol=1 - 1e-7)


def test_sato_quadratic_ridge():
    """f = a * y^2 / 2: the Hessian is diag(a, 0) away from the borders, so
    sato = a * sigma^2 there (2-D: the geometric mean of one eigenvalue)."""
    a = 0.02
    y = np.arange(64, dtype=np.float64)[:, None] - 31.5
    img = np.repeat(a * y * y / 2, 64, axis=1)
    # sigma <= 1 samples the derivative kernels at sigma/sqrt(2) ~ 0.7, whose
    # discrete second moment is ~0.2 % off; from sigma 2 on it is exact
    for sigma, rtol in ((1.0, 1e-2), (2.0, 1e-9), (3.0, 1e-9)):
        out = osk.sato(img, sigmas=[sigma], black_ridges=True, mode="nearest", ndi=sndi)
        np.testing.assert_allclose(out[24:40, 24:40], a * sigma ** 2, rtol=rtol)


def _brute(img, sigma, kind, alpha=None):
    """Per-pixel restatement: Hessian as two first-derivative Gaussian passes
    of sigma/sqrt(2) (scipy), numpy's eigvalsh per pixel, the formulas of
    ridges.py:114-143 / 205-224 written out per pixel."""
    nd = img.ndim
    s = sigma / np.sqrt(2)
    tr = 8 if sigma > 1 else 100
    grads = []
    for i in range(nd):
        order = [0] * nd
        order[i] = 1
        grads.append(sndi.gaussian_filter(img, s, order=order, mode="reflect", truncate=tr))
    H = {}
    for i, j in itertools.combinations_with_replacement(range(nd), 2):
        order = [0] * nd
        order[j] = 1
        H[i, j] = H[j, i] = sndi.gaussian_filter(grads[i], s, order=order, mode="reflect",
                                                 truncate=tr)
    out = np.zeros(img.shape)
    vals = np.zeros(img.shape)
    for idx in np.ndindex(img.shape):
        M = np.array([[H[i, j][idx] for j in range(nd)] for i in range(nd)])
        ev = np.linalg.eigvalsh(M)[::-1]
        ev[np.abs(ev) < 1e-12] = 0
        if kind == "sato":
            out[idx] = np.prod(np.maximum(ev[:-1], 0)) ** (1 / (nd - 1)) * sigma ** 2
        else:
            al = 1 / (nd + 1) if alpha is None else alpha
            v = [ev[i] + al * (ev.sum() - ev[i]) for i in range(nd)]
            vals[idx] = max(v[int(np.argmax(np.abs(v)))], 0)
    if kind == "meijering":
        out = vals / vals.max() if vals.max() > 0 else vals
    return out


@pytest.mark.parametrize("shape", [(9, 11), (6, 7, 5)])
@pytest.mark.parametrize("kind", ["sato", "meijering"])
def test_brute_force_tiny(shape, kind):
    rng = np.random.default_rng(3)
    img = rng.standard_normal(shape)
    sigma = 1.5
    got = FUNCS[kind](img, sigmas=[sigma], mode="reflect", ndi=sndi)
    exp = _brute(img, sigma, kind)
    np.testing.assert_allclose(got, exp, rtol=1e-6, atol=1e-9)
