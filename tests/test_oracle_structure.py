"""Pins of the oracle's structure_tensor, structure_tensor_eigenvalues and
difference_of_gaussians.

Structure tensor: the known answers of the reference's
feature/tests/test_corner.py (5x5 and 5x5x5 deltas, eigenvalues), and a
brute-force per-pixel restatement (Sobel by explicit 3x3 sums, products,
Gaussian by scipy) on a tiny image.  Difference of Gaussians: the impulse
response sums to zero (both kernels sum to one), equals scipy's two
Gaussian blurs subtracted, the default high sigma is 1.6 x low, and the
reference's invalid-sigma errors.
"""

import numpy as np
import pytest
import scipy.ndimage as sndi

from oracle import skimage as osk


def test_structure_tensor_2d_kat(kats):
    sq = np.zeros((5, 5))
    sq[2, 2] = 1
    Arr, Arc, Acc = osk.structure_tensor(sq, sigma=0.1, order="rc")
    k = kats["structure_tensor_delta5x5"]
    np.testing.assert_array_equal(Arr, k["Arr"])
    np.testing.assert_array_equal(Arc, k["Arc"])
    np.testing.assert_array_equal(Acc, k["Acc"])
    l1, l2 = osk.structure_tensor_eigenvalues([Arr, Arc, Acc])
    np.testing.assert_array_equal(l1, kats["structure_tensor_eigvals_5x5"]["l1"])
    np.testing.assert_array_equal(l2, kats["structure_tensor_eigvals_5x5"]["l2"])


def test_structure_tensor_3d_kat(kats):
    cube = np.zeros((5, 5, 5))
    cube[2, 2, 2] = 1
    A = osk.structure_tensor(cube, sigma=0.1)
    assert len(A) == 6
    k = kats["structure_tensor_3d_delta"]
    np.testing.assert_array_equal(A[0][:, 1, :], k["A0_slice_axis1_1"])
    np.testing.assert_array_equal(A[0][1], k["A0_plane1"])
    np.testing.assert_array_equal(A[3][2], k["A3_plane2"])


def test_structure_tensor_brute_force():
    rng = np.random.default_rng(1)
    x = rng.random((7, 9))
    A = osk.structure_tensor(x, sigma=1.3, mode="reflect")
    p = np.pad(x, 1, mode="symmetric")
    d0 = np.zeros_like(x)
    d1 = np.zeros_like(x)
    for i in range(7):
        for j in range(9):
            w = p[i:i + 3, j:j + 3]
            d0[i, j] = (w[2, :] - w[0, :]) @ np.array([1.0, 2.0, 1.0])
            d1[i, j] = np.array([1.0, 2.0, 1.0]) @ (w[:, 2] - w[:, 0])
    exp = [sndi.gaussian_filter(a * b, 1.3, mode="reflect")
           for a, b in ((d0, d0), (d0, d1), (d1, d1))]
    for a, e in zip(A, exp):
        np.testing.assert_allclose(a, e, rtol=1e-12, atol=1e-12)


def test_structure_tensor_errors():
    with pytest.raises(ValueError):
        osk.structure_tensor(np.zeros((5, 5, 5)), sigma=0.1, order="xy")
    with pytest.raises(ValueError):
        osk.structure_tensor(np.zeros((5, 5)), sigma=(0.1,))


def test_dog_impulse_sums_to_zero():
    x = np.zeros((41, 41))
    x[20, 20] = 1
    d = osk.difference_of_gaussians(x, 1.5, 3.0, mode="constant")
    assert abs(d.sum()) < 1e-12
    assert d[20, 20] > 0 and d[20, 12] < 0 and d[20, 5] == 0  # radius 12 at sigma 3


def test_dog_is_two_blurs_subtracted():
    rng = np.random.default_rng(2)
    x = rng.random((12, 15))
    d = osk.difference_of_gaussians(x, (1, 2), (2, 3), mode="nearest", truncate=4.0)
    e = (sndi.gaussian_filter(x, (1, 2), mode="nearest")
         - sndi.gaussian_filter(x, (2, 3), mode="nearest"))
    np.testing.assert_allclose(d, e, rtol=1e-12, atol=1e-13)
    np.testing.assert_array_equal(osk.difference_of_gaussians(x, 1.5),
                                  osk.difference_of_gaussians(x, 1.5, 1.5 * 1.6))


def test_dog_errors():
    with pytest.raises(ValueError):
        osk.difference_of_gaussians(np.ones((5, 5, 3)), (1, 2))
    with pytest.raises(ValueError):
        osk.difference_of_gaussians(np.ones((3, 3)), 3, 2)
    with pytest.raises(ValueError):
        osk.difference_of_gaussians(np.ones((5, 5, 3)), (1, 2, 3), channel_axis=-1)
