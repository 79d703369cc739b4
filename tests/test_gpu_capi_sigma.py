"""The sigma-taking C entry points (hc_gaussian_nd, hc_hessian_matrix,
hc_frangi_sigmas) called through ctypes, as a non-Python caller would, against
the Python path that builds its kernels with numpy: bit for bit for float32
images (the float32 taps agree), within 1e-14 relative for float64 (the
library's exp can differ from numpy's in the last bit)."""

import ctypes

import numpy as np
import pytest
import torch

from cucim_amd import _array, _lib

pytestmark = pytest.mark.gpu

MODES = {"reflect": 0, "mirror": 1, "nearest": 2, "wrap": 3, "constant": 4}


def _np(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _same(a, b, dtype):
    if dtype == np.float32:
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
    else:
        np.testing.assert_allclose(a, b, rtol=1e-14, atol=1e-14 * np.abs(b).max())


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.uint8])
@pytest.mark.parametrize("shape,sigma,order", [((2048, 2048), (2.0, 2.0), (0, 0)),
                                               ((64, 80), (1.5, 0.7), (1, 2)),
                                               ((20, 30, 40), (2.0, 0.0, 1.2), (0, 1, 0)),
                                               ((500,), (3.3,), (0,))])
@pytest.mark.parametrize("mode", ["reflect", "constant", "wrap"])
def test_gaussian_nd_vs_python(gpu, dtype, shape, sigma, order, mode):
    from cucim_amd.skimage._vendored import ndimage as ndi

    rng = np.random.default_rng(1)
    x = (rng.random(shape) * (255 if dtype == np.uint8 else 1)).astype(dtype)
    xt = torch.from_numpy(x).cuda()
    ref = _np(ndi.gaussian_filter(xt, sigma, order, mode=mode, cval=0.5))
    out = torch.empty_like(xt)
    work = torch.empty_like(xt)
    nd = len(shape)
    s = np.ascontiguousarray(sigma, np.float64)
    o = np.ascontiguousarray(order, np.int32)
    m = np.full(nd, MODES[mode], np.int32)
    rc = _lib.lib().hc_gaussian_nd(_array.hc_array(xt), _array.hc_array(out), s.ctypes.data,
                                   o.ctypes.data, None, m.ctypes.data, 0.5, 4.0,
                                   _array.hc_array(work), _array.stream())
    assert rc == 0, _lib.lib().hc_last_error()
    got = _np(out)
    if dtype == np.uint8:
        np.testing.assert_array_equal(got, ref)
    else:
        _same(got, ref, dtype)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", [(64, 70), (24, 30, 28)])
@pytest.mark.parametrize("sigma", [1.0, 2.5])
def test_hessian_matrix_vs_python(gpu, dtype, shape, sigma):
    from cucim_amd.skimage.feature import hessian_matrix

    x = np.random.default_rng(2).standard_normal(shape).astype(dtype)
    xt = torch.from_numpy(x).cuda()
    ref = [_np(h) for h in hessian_matrix(xt, sigma, mode="reflect",
                                          use_gaussian_derivatives=True)]
    H = [torch.empty_like(xt) for _ in ref]
    sg = np.full(len(shape), sigma, np.float64)
    scratch = torch.empty(64, dtype=torch.uint8, device="cuda")
    rc = _lib.lib().hc_hessian_matrix(_array.hc_array(xt), _array.hc_arrays(H), len(H),
                                      sg.ctypes.data, MODES["reflect"], 0.0, 0,
                                      scratch.data_ptr(), _array.stream())
    assert rc == 0, _lib.lib().hc_last_error()
    for h, r in zip(H, ref):
        _same(_np(h), r, dtype)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("gamma", [None, 0.4])
def test_frangi_sigmas_vs_python(gpu, dtype, gamma):
    from cucim_amd.skimage.filters import frangi
    from cucim_amd.synthetic import tube_volume

    x = tube_volume((40, 44, 48), ntubes=5, seed=4).to(torch.float64 if dtype == np.float64
                                                         else torch.float32)
    sig = (1.0, 2.0, 3.0)
    ref = _np(frangi(x, sigmas=sig, black_ridges=False, gamma=gamma))
    out = torch.empty_like(x)
    work = [torch.empty_like(x) for _ in range(6)]
    scratch = torch.zeros(64, dtype=torch.uint8, device="cuda")
    s = np.ascontiguousarray(sig, np.float64)
    rc = _lib.lib().hc_frangi_sigmas(
        _array.hc_array(x), _array.hc_array(out), s.ctypes.data, len(sig), 0.5, 0.5,
        0.0 if gamma is None else gamma, 0 if gamma is None else 1, 0, MODES["reflect"], 0.0,
        _array.hc_arrays(work), 6, scratch.data_ptr(), _array.stream())
    assert rc == 0, _lib.lib().hc_last_error()
    _same(_np(out), ref, dtype)


def test_sigma_entry_point_errors(gpu):
    lib = _lib.lib()
    x = torch.zeros((8, 8), device="cuda")
    out = torch.empty_like(x)
    s = np.array([1.0, -1.0])
    o = np.zeros(2, np.int32)
    m = np.zeros(2, np.int32)
    assert lib.hc_gaussian_nd(_array.hc_array(x), _array.hc_array(out), s.ctypes.data,
                              o.ctypes.data, None, m.ctypes.data, 0.0, 4.0, None,
                              _array.stream()) == _lib.ERR_VALUE
    s = np.array([1.0, 1.0])
    x3 = torch.zeros((8, 8, 8), device="cuda")
    assert lib.hc_gaussian_nd(_array.hc_array(x3), _array.hc_array(torch.empty_like(x3)),
                              np.ones(3).ctypes.data, np.zeros(3, np.int32).ctypes.data, None,
                              np.zeros(3, np.int32).ctypes.data, 0.0, 4.0, None,
                              _array.stream()) == _lib.ERR_VALUE
    assert b"work" in lib.hc_last_error()
    assert lib.hc_frangi_sigmas(_array.hc_array(x), _array.hc_array(out), s.ctypes.data, 0, 0.5,
                                0.5, 0.0, 0, 1, 0, 0.0, None, 0, None,
                                _array.stream()) == _lib.ERR_VALUE
