"""GPU parity of structure_tensor, structure_tensor_eigenvalues and
difference_of_gaussians.

Known answers of the reference's tests are matched exactly (float16,
float32, float64 input).  Random images: float32 within 2e-6 relative (of
the largest magnitude) of the oracle's FP64 run, float64 within 1e-12.
difference_of_gaussians equals, bit for bit, the two skimage-level
Gaussian calls subtracted on the device; the structure tensor's products
are one rounding each, then the skimage Gaussian.
"""

import numpy as np
import pytest
import torch

from oracle import skimage as osk

pytestmark = pytest.mark.gpu


def _t(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def _np(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.mark.parametrize("dtype", [np.float16, np.float32, np.float64])
def test_structure_tensor_kats(gpu, kats, dtype):
    from cucim_amd.skimage.feature import structure_tensor, structure_tensor_eigenvalues

    sq = np.zeros((5, 5), dtype=dtype)
    sq[2, 2] = 1
    Arr, Arc, Acc = structure_tensor(_t(sq), sigma=0.1, order="rc")
    out_dt = np.float64 if dtype == np.float64 else np.float32
    k = kats["structure_tensor_delta5x5"]
    for a, name in ((Arr, "Arr"), (Arc, "Arc"), (Acc, "Acc")):
        a = _np(a)
        assert a.dtype == out_dt
        np.testing.assert_array_equal(a, k[name])
    l1, l2 = structure_tensor_eigenvalues([Arr, Arc, Acc])
    np.testing.assert_array_equal(_np(l1), kats["structure_tensor_eigvals_5x5"]["l1"])
    np.testing.assert_array_equal(_np(l2), kats["structure_tensor_eigvals_5x5"]["l2"])
    cube = np.zeros((5, 5, 5), dtype=dtype)
    cube[2, 2, 2] = 1
    A = [_np(a) for a in structure_tensor(_t(cube), sigma=0.1)]
    k = kats["structure_tensor_3d_delta"]
    assert len(A) == 6
    np.testing.assert_array_equal(A[0][:, 1, :], k["A0_slice_axis1_1"])
    np.testing.assert_array_equal(A[0][1], k["A0_plane1"])
    np.testing.assert_array_equal(A[3][2], k["A3_plane2"])


def test_structure_tensor_orders_and_sigma(gpu):
    from cucim_amd.skimage.feature import structure_tensor

    sq = np.zeros((5, 5))
    sq[2, 2] = 1
    base = [_np(a) for a in structure_tensor(_t(sq), sigma=0.1)]
    xy = [_np(a) for a in structure_tensor(_t(sq), sigma=0.1, order="xy")]
    for a, b in zip(xy, base[::-1]):
        np.testing.assert_array_equal(a, b)
    tup = [_np(a) for a in structure_tensor(_t(sq), sigma=(0.1, 0.1))]
    for a, b in zip(tup, base):
        np.testing.assert_array_equal(a, b)
    with pytest.raises(ValueError):
        structure_tensor(_t(np.zeros((5, 5, 5))), sigma=0.1, order="xy")
    with pytest.raises(ValueError):
        structure_tensor(_t(sq), sigma=(0.1,))
    with pytest.raises(ValueError):
        structure_tensor(_t(sq), sigma=[0.1] * 3)


def test_structure_tensor_eigenvalues_3d_facets(gpu):
    from cucim_amd.skimage.feature import structure_tensor, structure_tensor_eigenvalues

    cube9 = np.ones((9,) * 3, dtype=np.int16)
    cube7 = np.ones((7,) * 3, dtype=np.int16)
    image = np.pad(cube9, 5, mode="constant") * 1000
    boundary = (np.pad(cube9, 5, mode="constant")
                - np.pad(cube7, 6, mode="constant")).astype(bool)
    e0, e1, e2 = structure_tensor_eigenvalues(structure_tensor(_t(image), sigma=0.1))
    assert np.all(_np(e0)[boundary] != 0)


@pytest.mark.parametrize("shape", [(33, 47), (12, 17, 19), (20, 24, 3)])
@pytest.mark.parametrize("mode", ["constant", "reflect", "nearest"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_structure_tensor_vs_fp64(gpu, shape, mode, dtype):
    from cucim_amd.skimage.feature import structure_tensor, structure_tensor_eigenvalues

    rng = np.random.default_rng(3)
    x = rng.random(shape).astype(dtype)
    got = [_np(a) for a in structure_tensor(_t(x), sigma=1.5, mode=mode)]
    exp = osk.structure_tensor(x.astype(np.float64), sigma=1.5, mode=mode)
    tol = 2e-6 if dtype == np.float32 else 1e-12
    for g, e in zip(got, exp):
        assert g.dtype == dtype
        assert np.abs(g - e).max() <= tol * np.abs(e).max()
    if len(shape) == 2 or shape[-1] != 3:
        ge = _np(structure_tensor_eigenvalues([_t(g) for g in got]))
        ee = osk.structure_tensor_eigenvalues(exp)
        etol = 1e-5 if dtype == np.float32 else 1e-11
        assert np.abs(ge - ee).max() <= etol * np.abs(ee).max()


@pytest.mark.parametrize("s", [1, (2, 3)])
@pytest.mark.parametrize("s2", [4, (5, 6), None])
@pytest.mark.parametrize("channel_axis", [None, 0, -1])
@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.uint8])
def test_difference_of_gaussians(gpu, s, s2, channel_axis, dtype):
    from cucim_amd.skimage.filters import difference_of_gaussians, gaussian

    rng = np.random.default_rng(4)
    x = rng.random((24, 30))
    if channel_axis is not None:
        x = np.stack((x, x[::-1], x[:, ::-1]), channel_axis)
    x = (x * 255).astype(dtype) if dtype == np.uint8 else x.astype(dtype)
    xt = _t(x)
    dog = difference_of_gaussians(xt, s, s2, channel_axis=channel_axis)
    hi = tuple(1.6 * np.atleast_1d(s)) if s2 is None else s2
    if s2 is None and np.isscalar(s):
        hi = 1.6 * s
    g1 = gaussian(xt, sigma=s, channel_axis=channel_axis)
    g2 = gaussian(xt, sigma=hi, channel_axis=channel_axis)
    got = _np(dog)
    np.testing.assert_array_equal(got, _np(g1 - g2))
    exp = osk.difference_of_gaussians(x, s, s2, channel_axis=channel_axis, dtype=np.float64)
    tol = 1e-12 if dtype == np.float64 else 2e-6
    assert np.abs(got - exp).max() <= tol * max(np.abs(exp).max(), 1.0)


def test_difference_of_gaussians_errors(gpu):
    from cucim_amd.skimage.filters import difference_of_gaussians

    with pytest.raises(ValueError):
        difference_of_gaussians(_t(np.ones((5, 5, 3))), (1, 2))
    with pytest.raises(ValueError):
        difference_of_gaussians(_t(np.ones((5, 5, 3))), 1, (3, 4))
    with pytest.raises(ValueError):
        difference_of_gaussians(_t(np.ones((5, 5, 3))), (1, 2, 3), channel_axis=-1)
    with pytest.raises(ValueError):
        difference_of_gaussians(_t(np.ones((3, 3))), 3, 2)
    with pytest.raises(ValueError):
        difference_of_gaussians(_t(np.ones((3, 3))), (1, 5), (2, 4))


def _sym_elems(mats):
    """Element images of the stacked matrices, shaped (count, 1, ..., 1) so
    that the image rank equals the matrix order, as the reference's
    _symmetric_image assumes."""
    n = mats.shape[-1]
    from itertools import combinations_with_replacement as cwr

    shp = (mats.shape[0],) + (1,) * (n - 1)
    return [np.ascontiguousarray(mats[..., r, c]).reshape(shp) for r, c in cwr(range(n), 2)]


@pytest.mark.parametrize("n", [1, 4, 5, 8])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("sort,abs_sort", [("descending", False), ("ascending", False),
                                           ("ascending", True), ("descending", True)])
def test_eigvals_nd_vs_eigvalsh(gpu, n, dtype, sort, abs_sort):
    """The eigvalsh branch of _symmetric_compute_eigenvalues
    (feature/corner.py:744-755) for matrix orders other than 2 and 3:
    random symmetric matrices, plus exactly diagonal ones (ties included) and
    rank-1 ones, against numpy.linalg.eigvalsh in float64."""
    from cucim_amd.skimage.feature.corner import _symmetric_compute_eigenvalues

    rng = np.random.default_rng(40 + n)
    m = rng.standard_normal((300, n, n))
    m = m + np.swapaxes(m, -1, -2)
    d = np.zeros((40, n, n))
    idx = np.arange(n)
    d[:, idx, idx] = rng.integers(-3, 4, (40, n))  # ties, zeros, +-pairs
    v = rng.standard_normal((40, n))
    r1 = v[:, :, None] * v[:, None, :]
    mats = np.concatenate([m, d, r1]).astype(dtype)
    elems = _sym_elems(mats)
    got = _np(_symmetric_compute_eigenvalues([_t(e) for e in elems], sort=sort,
                                             abs_sort=abs_sort))
    exp = osk.symmetric_compute_eigenvalues([e.astype(np.float64) for e in elems],
                                            sort=sort, abs_sort=abs_sort)
    got = got.reshape(n, -1)
    exp = exp.reshape(n, -1)
    assert got.dtype == dtype and got.shape == (n, mats.shape[0])
    scale = np.abs(exp).max(axis=0, keepdims=True)
    tol = 2e-6 if dtype == np.float32 else 1e-12
    assert np.all(np.abs(got - exp) <= tol * scale)
    # the diagonal block is exact, ties in the stable order
    np.testing.assert_array_equal(got[:, 300:340], exp[:, 300:340].astype(dtype))


def test_hessian_eigvals_4d(gpu):
    """hessian_matrix_eigvals and structure_tensor_eigenvalues of a 4-D image
    (the reference raises nothing here; it stacks the 10 elements and calls
    eigvalsh)."""
    from cucim_amd.skimage.feature import (hessian_matrix, hessian_matrix_eigvals,
                                           structure_tensor, structure_tensor_eigenvalues)

    rng = np.random.default_rng(44)
    import scipy.ndimage as sndi

    x = sndi.gaussian_filter(rng.standard_normal((9, 10, 11, 12)), 1.0).astype(np.float32)
    H = hessian_matrix(_t(x), sigma=1.5, use_gaussian_derivatives=True)
    assert len(H) == 10
    got = _np(hessian_matrix_eigvals(H))
    exp = osk.hessian_matrix_eigvals([_np(h).astype(np.float64) for h in H])
    assert got.shape == (4,) + x.shape
    assert np.abs(got - exp).max() <= 2e-6 * np.abs(exp).max()
    A = structure_tensor(_t(x), sigma=1.0)
    got = _np(structure_tensor_eigenvalues(A))
    exp = osk.structure_tensor_eigenvalues([_np(a).astype(np.float64) for a in A])
    assert np.abs(got - exp).max() <= 2e-6 * np.abs(exp).max()
    assert np.all(np.diff(got, axis=0) <= 0)
