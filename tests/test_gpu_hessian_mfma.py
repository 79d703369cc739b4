"""Matrix-core (v_mfma_f32_16x16x4f32) Hessian passes.

The float32 passes of the fused Hessian (strided line passes and the final
pass along the contiguous axis, for lines of up to 512 voxels: 4 waves per
work-group up to 256, 8 above) run on the matrix cores by default; the
HC_TUNE_HESSIAN_VALU knob switches them to the vector-ALU passes, tap by tap
as the reference computes them.  Tolerances (float32 data):
Hessian elements within 2e-6 * max|H| of the FP64 oracle and of the VALU
passes; Frangi within the plain Frangi's bounds.  A work-group of lines that
holds a non-finite value is computed tap by tap, so the non-finite pattern of
the output equals the VALU passes' exactly.
"""

import numpy as np
import pytest
import torch

from cucim_amd import _lib
from oracle import skimage as osk

pytestmark = pytest.mark.gpu


def _t(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def _np(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


class _Valu:
    """Vector-ALU passes only inside the block."""

    def __enter__(self):
        _lib.check(_lib.lib().hc_tune_set(_lib.TUNE_HESSIAN_VALU, 1))

    def __exit__(self, *exc):
        _lib.lib().hc_tune_set(_lib.TUNE_HESSIAN_VALU, 0)


def _hessian(x, sigma, mode, valu):
    from cucim_amd.skimage.feature import hessian_matrix

    def run():
        return [_np(h) for h in hessian_matrix(_t(x), sigma, mode=mode, cval=0.0,
                                               use_gaussian_derivatives=True)]

    if not valu:
        return run()
    with _Valu():
        return run()


@pytest.mark.parametrize("shape", [(512, 24), (509, 40), (33, 17), (16, 16), (1, 40), (24, 512),
                                   (17, 300), (20, 130, 9), (64, 48, 20), (9, 20, 256),
                                   (5, 7, 300), (300, 9, 20), (6, 511, 17), (4, 5, 512)])
@pytest.mark.parametrize("sigma", [1.0, 2.5, 5.0])
@pytest.mark.parametrize("mode", ["reflect", "constant", "wrap"])
def test_mfma_hessian_vs_fp64_and_valu(gpu, shape, sigma, mode):
    import scipy.ndimage as sndi

    rng = np.random.default_rng(11)
    x = rng.standard_normal(shape).astype(np.float32)
    got = _hessian(x, sigma, mode, valu=False)
    ref = _hessian(x, sigma, mode, valu=True)
    exp = osk.hessian_matrix(x.astype(np.float64), sigma, mode=mode, cval=0.0,
                             use_gaussian_derivatives=True, ndi=sndi)
    for g, r, e in zip(got, ref, exp):
        scale = np.abs(e).max()
        assert np.abs(g - e).max() <= 2e-6 * scale + 1e-7, (np.abs(g - e).max(), scale)
        assert np.abs(g - r).max() <= 2e-6 * scale + 1e-7, (np.abs(g - r).max(), scale)


@pytest.mark.parametrize("shape", [(40, 33), (5, 300), (3, 17, 260), (21, 9, 7),
                                   (140, 6, 150), (9, 190, 177), (133, 20), (70, 5, 90),
                                   (81, 96)])
@pytest.mark.parametrize("sigma", [1.0, 3.0, 5.0])
@pytest.mark.parametrize("mode", ["nearest", "mirror"])
def test_mfma_hessian_other_modes(gpu, shape, sigma, mode):
    """Boundary modes whose halo rows repeat one voxel (nearest) or skip the
    edge voxel (mirror), on lines shorter than the kernel reach."""
    import scipy.ndimage as sndi

    rng = np.random.default_rng(14)
    x = rng.standard_normal(shape).astype(np.float32)
    got = _hessian(x, sigma, mode, valu=False)
    exp = osk.hessian_matrix(x.astype(np.float64), sigma, mode=mode, cval=0.0,
                             use_gaussian_derivatives=True, ndi=sndi)
    for g, e in zip(got, exp):
        scale = np.abs(e).max()
        assert np.abs(g - e).max() <= 2e-6 * scale + 1e-7, (np.abs(g - e).max(), scale)


@pytest.mark.parametrize("value", [np.inf, -np.inf, np.nan])
def test_mfma_non_finite_pattern(gpu, value):
    """0 * inf is NaN on the matrix cores: a group with a non-finite input
    must take the tap-by-tap path, or NaN would spread past the reach."""
    rng = np.random.default_rng(12)
    x = rng.standard_normal((40, 37, 21)).astype(np.float32)
    x[20, 9, 5] = value
    x[3, 30, 20] = value
    got = _hessian(x, 1.5, "reflect", valu=False)
    ref = _hessian(x, 1.5, "reflect", valu=True)
    for g, r in zip(got, ref):
        np.testing.assert_array_equal(np.isnan(g), np.isnan(r))
        np.testing.assert_array_equal(np.isinf(g), np.isinf(r))
        fin = np.isfinite(r)
        scale = np.abs(r[fin]).max()
        assert np.abs(g[fin] - r[fin]).max() <= 2e-6 * scale + 1e-7


def test_mfma_frangi_region_alignment(gpu):
    """A region that starts at multiples of 16 and carries the filter's reach
    gives the whole volume's plain Frangi bit for bit inside it."""
    from cucim_amd.skimage.filters import ridges

    rng = np.random.default_rng(13)
    x = _t(rng.random((112, 112, 40)).astype(np.float32))
    sig = [1.0, 2.0]
    reg_in = x[16:96, 32:112].contiguous()
    full, reg = torch.empty_like(x), torch.empty_like(reg_in)
    plan = ridges.FrangiPlan(tuple(x.shape), np.float32, sig)
    assert plan.run(x, full, gamma=0.5, black_ridges=False) == 0
    plan2 = ridges.FrangiPlan(tuple(reg_in.shape), np.float32, sig)
    assert plan2.run(reg_in, reg, gamma=0.5, black_ridges=False) == 0
    # reach of the two passes per axis: sigma 2 -> 2 * int(8 * sqrt(2) + 0.5)
    # = 22 voxels; sigma 1's float32 taps are zero beyond 10 per pass
    reach = 24
    a = _np(full[16 + reach:96 - reach, 32 + reach:112 - reach])
    b = _np(reg[reach:80 - reach, reach:80 - reach])
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))


def test_hessian_knob_values(gpu):
    lib = _lib.lib()
    assert lib.hc_tune_set(_lib.TUNE_HESSIAN_VALU, 2) == _lib.ERR_VALUE
    assert lib.hc_tune_get(_lib.TUNE_HESSIAN_VALU) == 0
