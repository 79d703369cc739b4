"""Pins the oracle's skimage layer (CPU only)."""

import math

import numpy as np
import pytest
import scipy.ndimage as sndi

from oracle import skimage as osk


def _square():
    a = np.zeros((5, 5))
    a[2, 2] = 4
    return a


def test_hessian_fd_kats(kats):
    k = kats["hessian_fd_5x5"]
    Hrr, Hrc, Hcc = osk.hessian_matrix(_square(), sigma=0.1, order="rc",
                                       use_gaussian_derivatives=False)
    tol = 1.5 * 10 ** -k["decimals"]
    np.testing.assert_allclose(Hrr, k["Hrr"], atol=tol)
    np.testing.assert_allclose(Hrc, k["Hrc"], atol=tol)
    np.testing.assert_allclose(Hcc, k["Hcc"], atol=tol)
    with pytest.warns(FutureWarning):
        osk.hessian_matrix(_square(), sigma=0.1, order="rc")


def test_hessian_fd_3d_kat(kats):
    k = kats["hessian_fd_3d_slice"]
    cube = np.zeros((5, 5, 5))
    cube[2, 2, 2] = 4
    Hs = osk.hessian_matrix(cube, sigma=0.1, order="rc", use_gaussian_derivatives=False)
    assert len(Hs) == 6
    np.testing.assert_allclose(Hs[2][:, 2, :], k["H02_slice"], atol=1.5e-6)


def test_eigvals_kat(kats):
    k = kats["hessian_eigvals_5x5"]
    H = osk.hessian_matrix(_square(), sigma=0.1, order="rc", use_gaussian_derivatives=False)
    l1, l2 = osk.hessian_matrix_eigvals(H)
    np.testing.assert_allclose(l1, k["l1"], atol=1.5e-6)
    np.testing.assert_allclose(l2, k["l2"], atol=1.5e-6)


@pytest.mark.parametrize("ugd", [False, True])
def test_hessian_order_xy_equals_rc(ugd):
    # test_corner.py:170-192: on a near-delta kernel xy is rc reversed
    a = _square()
    Hxx, Hxy, Hyy = osk.hessian_matrix(a, sigma=0.1, order="xy", use_gaussian_derivatives=ugd)
    Hrr, Hrc, Hcc = osk.hessian_matrix(a, sigma=0.1, order="rc", use_gaussian_derivatives=ugd)
    np.testing.assert_allclose(Hxx, Hcc, atol=1e-30)
    np.testing.assert_allclose(Hxy, Hrc, atol=1e-30)
    np.testing.assert_allclose(Hyy, Hrr, atol=1e-30)


@pytest.mark.parametrize("shape", [(9, 11), (7, 8, 9)])
def test_hessian_gaussian_oracle_vs_scipy_filters(shape):
    rng = np.random.default_rng(0)
    img = rng.standard_normal(shape)
    for sigma in (0.8, 2.0):
        for mode in ("reflect", "nearest", "constant", "wrap", "mirror"):
            a = osk.hessian_matrix(img, sigma, mode=mode, cval=0.3, use_gaussian_derivatives=True)
            b = osk.hessian_matrix(img, sigma, mode=mode, cval=0.3, use_gaussian_derivatives=True,
                                   ndi=sndi)
            for x, y in zip(a, b):
                np.testing.assert_allclose(x, y, rtol=1e-11, atol=1e-12)


def test_hessian_gaussian_quadratic_interior():
    # Gaussian-derivative Hessian of f = x0^2 + 3 x0 x1 - 2 x1^2 (constant
    # Hessian [[2, 3], [3, -4]]): interior values match to the discretisation
    # error of the sampled kernels
    i, j = np.mgrid[0:64, 0:64].astype(float)
    f = i**2 + 3 * i * j - 2 * j**2
    Hrr, Hrc, Hcc = osk.hessian_matrix(f, 3.0, mode="reflect", use_gaussian_derivatives=True)
    c = (slice(24, 40), slice(24, 40))
    np.testing.assert_allclose(Hrr[c], 2, rtol=1e-3)
    np.testing.assert_allclose(Hrc[c], 3, rtol=1e-3)
    np.testing.assert_allclose(Hcc[c], -4, rtol=1e-3)


def _random_sym(rng, n, ndim):
    A = rng.standard_normal((n, ndim, ndim))
    return (A + A.transpose(0, 2, 1)) / 2


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_eig_closed_form_vs_eigvalsh(dtype):
    rng = np.random.default_rng(1)
    M = _random_sym(rng, 5000, 3).astype(dtype)
    els = [M[:, 0, 0], M[:, 0, 1], M[:, 0, 2], M[:, 1, 1], M[:, 1, 2], M[:, 2, 2]]
    ev = osk.symmetric_compute_eigenvalues(els, sort="descending")
    ref = np.linalg.eigvalsh(M.astype(np.float64))[:, ::-1].T
    tol = 1e-10 if dtype == np.float64 else 2e-5
    np.testing.assert_allclose(ev, ref, atol=tol * 10, rtol=tol)
    M2 = _random_sym(rng, 5000, 2).astype(dtype)
    ev2 = osk.symmetric_compute_eigenvalues([M2[:, 0, 0], M2[:, 0, 1], M2[:, 1, 1]],
                                            sort="ascending")
    ref2 = np.linalg.eigvalsh(M2.astype(np.float64)).T
    np.testing.assert_allclose(ev2, ref2, atol=tol * 10, rtol=tol)


def test_eig33_abs_sort_order():
    """Ascending |lambda| order of the reference's network: correct except when
    |max| < |mid| < |min|, where the three-compare network with keys taken
    before the swaps yields (mid, min, max)."""
    rng = np.random.default_rng(2)
    M = _random_sym(rng, 20000, 3)
    els = [M[:, 0, 0], M[:, 0, 1], M[:, 0, 2], M[:, 1, 1], M[:, 1, 2], M[:, 2, 2]]
    ev = osk.symmetric_compute_eigenvalues(els, sort="ascending", abs_sort=True)
    w = np.linalg.eigvalsh(M)  # ascending: min, mid, max
    mn, md, mx = w[:, 0], w[:, 1], w[:, 2]
    special = (np.abs(mx) < np.abs(md)) & (np.abs(md) < np.abs(mn))
    proper = np.take_along_axis(w, np.argsort(np.abs(w), axis=1), axis=1).T
    assert special.any() and (~special).any()
    np.testing.assert_allclose(ev[:, ~special], proper[:, ~special], atol=1e-9)
    np.testing.assert_allclose(ev[0, special], md[special], atol=1e-9)
    np.testing.assert_allclose(ev[1, special], mn[special], atol=1e-9)
    np.testing.assert_allclose(ev[2, special], mx[special], atol=1e-9)


def test_eig_special_cases():
    # diagonal matrices give the sorted diagonal; x2 == 0 takes phi = pi/2
    a = np.array([3.0, -1.0, 2.0, 0.0])
    b = np.array([1.0, -1.0, 2.0, 0.0])
    c = np.array([2.0, 5.0, 2.0, 0.0])
    z = np.zeros(4)
    ev = osk.symmetric_compute_eigenvalues([a, z, z, b, z, c], sort="descending")
    np.testing.assert_allclose(ev.T, np.sort(np.stack([a, b, c], 1), axis=1)[:, ::-1],
                               atol=1e-12)


def _tube3d(n=24):
    z, y, x = np.mgrid[0:n, 0:n, 0:n].astype(float)
    d2 = (y - n / 2) ** 2 + (x - n / 2) ** 2
    return np.exp(-d2 / (2 * 2.0**2))


def test_frangi_null(kats):
    for shape in [(3, 3), (3, 3, 5)]:
        z = np.zeros(shape, np.uint8)
        np.testing.assert_allclose(osk.frangi(z, black_ridges=True), 0, atol=1e-3)
        np.testing.assert_allclose(osk.frangi(255 - z, black_ridges=False), 0, atol=1e-3)


def test_frangi_black_white_symmetry():
    rng = np.random.default_rng(3)
    img = _tube3d(16) + 0.05 * rng.standard_normal((16, 16, 16))
    a = osk.frangi(img, sigmas=(1, 2), black_ridges=False)
    b = osk.frangi(1.0 - img, sigmas=(1, 2), black_ridges=True)
    np.testing.assert_allclose(a, b, atol=1e-10, rtol=1e-10)
    img2 = img[8]
    np.testing.assert_allclose(osk.frangi(img2, sigmas=(1, 2), black_ridges=False),
                               osk.frangi(1.0 - img2, sigmas=(1, 2), black_ridges=True),
                               atol=1e-10, rtol=1e-10)


def test_frangi_linearity():
    img = _tube3d(16)
    np.testing.assert_allclose(osk.frangi(img, sigmas=(1, 2), black_ridges=False),
                               osk.frangi(10 * img, sigmas=(1, 2), black_ridges=False),
                               atol=1e-3)


def test_frangi_detects_tube():
    img = _tube3d(24)
    v = osk.frangi(img, sigmas=(1, 2, 3), black_ridges=False)
    assert v[12, 12, 12] > 0.5
    assert v[12, 2, 2] < 1e-3


def test_frangi_vs_independent_restatement():
    """Same pipeline on scipy's filters with LAPACK eigenvalues sorted by
    |value|: agrees wherever the reference's sort and a true sort coincide
    (elsewhere both give ~0 vesselness)."""
    rng = np.random.default_rng(4)
    img = _tube3d(20) + 0.05 * rng.standard_normal((20, 20, 20))
    sig = (1.0, 2.0)
    a = osk.frangi(img, sigmas=sig, black_ridges=False)
    img_n = -img
    out = None
    gamma_sq = None
    for i, s in enumerate(sig):
        H = osk.hessian_matrix(img_n, s, mode="reflect", use_gaussian_derivatives=True, ndi=sndi)
        M = np.empty(img.shape + (3, 3))
        for idx, (r, c) in enumerate([(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]):
            M[..., r, c] = H[idx]
            M[..., c, r] = H[idx]
        w = np.linalg.eigvalsh(M)
        w = np.take_along_axis(w, np.argsort(np.abs(w), axis=-1), axis=-1)
        l1, l2, l3 = w[..., 0], w[..., 1], w[..., 2]
        rg = l1**2 + l2**2 + l3**2
        if gamma_sq is None:
            gamma_sq = max(rg.max() / 2, 1e-10)
        L2, L3 = np.fmax(l2, 1e-10), np.fmax(l3, 1e-10)
        v = ((1 - np.exp(-(L2 / L3) ** 2 / 0.5)) * np.exp(-(l1**2 / (L2 * L3)) / 0.5)
             * (1 - np.exp(-rg / gamma_sq)))
        out = v if out is None else np.maximum(out, v)
    np.testing.assert_allclose(a, out, atol=1e-6)


def test_rgb2gray_kat(kats):
    k = kats["rgb2gray_white"]
    x = np.ones((1, 1, 3))
    for ca in (0, 1, -1, -2):
        g = osk.rgb2gray(np.moveaxis(x, -1, ca), channel_axis=ca)
        np.testing.assert_allclose(g, k["expected"], atol=0.5 * 10 ** -k["decimals"])
        assert g.shape == (1, 1)


def test_rgb2gray_uint8_within_one_ulp_of_exact():
    rng = np.random.default_rng(5)
    rgb = rng.integers(0, 256, (50000, 3), dtype=np.uint8)
    g = osk.rgb2gray(rgb)
    assert g.dtype == np.float32
    from fractions import Fraction as Fr

    coef = [Fr("0.2125"), Fr("0.7154"), Fr("0.0721")]
    for i in range(0, 50000, 97):
        ex = sum(c * int(v) for c, v in zip(coef, rgb[i])) / 255
        # f32(x) * f32(1/255) is rounded twice before the double sum
        assert abs(float(ex) - float(g[i])) <= 2 * float(np.spacing(np.float32(g[i])))
    with pytest.raises(ValueError):
        osk.rgb2gray(np.zeros((4, 4, 4)))


def test_sobel_step_edges():
    i, j = np.mgrid[-5:6, -5:6]
    for img, sel in [((i >= 0).astype(float), i), ((j >= 0).astype(float), j)]:
        r = osk.sobel(img) * math.sqrt(2)
        np.testing.assert_allclose(r[sel == 0], 1)
        np.testing.assert_allclose(r[np.abs(sel) > 1], 0, atol=1e-12)


def test_sobel_vs_scipy_magnitude():
    rng = np.random.default_rng(6)
    x = rng.standard_normal((12, 14))
    # edges.py: convolve1d([1,0,-1]/sqrt(2)) along the edge axis then
    # [1,2,1]/4 along the other; scipy.ndimage.sobel correlates [-1,0,1] and
    # [1,2,1], i.e. 4*sqrt(2) times the per-axis response (up to sign)
    m = np.sqrt(sum((sndi.sobel(x, a) / (4 * math.sqrt(2))) ** 2 for a in (0, 1)))
    np.testing.assert_allclose(osk.sobel(x), m, rtol=1e-12, atol=1e-12)
