"""CPU checks of the C ABI: the library loads, exports every declared entry
point, and rejects bad arguments before touching the GPU."""

import ctypes
import os
import re

import numpy as np
import pytest

from cucim_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    with open(os.path.join(ROOT, "include", "cucim_amd.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hc_\w+)\s*\(", text)))


def test_header_and_binding_agree():
    assert _declared() == sorted(_lib.ENTRY_POINTS)


def test_library_exports_every_entry_point():
    lib = _lib.load_library()
    for name in _declared():
        assert hasattr(lib, name), name
    assert lib.hc_version() == 1
    assert lib.hc_max_inline_taps(0) == 256
    assert lib.hc_max_inline_taps(1) == 128
    assert lib.hc_max_resident_line(0) > 256


def _arr(shape, dtype=_lib.FLOAT32, data=0x1000):
    a = _lib.HCArray()
    a.data = data
    a.dtype = dtype
    a.ndim = len(shape)
    stride = 1
    for d in range(len(shape) - 1, -1, -1):
        a.shape[d] = shape[d]
        a.strides[d] = stride
        stride *= shape[d]
    return a


def test_argument_errors_return_before_launch():
    lib = _lib.load_library()
    t, keep = _lib.make_taps(np.ones(3), 1, False)
    a, b = _arr((4, 5)), _arr((4, 6))
    assert lib.hc_correlate1d(ctypes.byref(a), ctypes.byref(b), ctypes.byref(t), 0, 0, 0.0, 0,
                              None) == _lib.ERR_VALUE
    assert b"shape" in lib.hc_last_error()
    b = _arr((4, 5))
    assert lib.hc_correlate1d(ctypes.byref(a), ctypes.byref(b), ctypes.byref(t), 0, 9, 0.0, 0,
                              None) == _lib.ERR_VALUE
    bad, keep2 = _lib.make_taps(np.ones(3), 3, False)
    assert lib.hc_correlate1d(ctypes.byref(a), ctypes.byref(b), ctypes.byref(bad), 0, 0, 0.0, 0,
                              None) == _lib.ERR_VALUE
    assert b"origin" in lib.hc_last_error()
    # long kernels cannot be forced onto the tiled kernels
    big, keep3 = _lib.make_taps(np.ones(301), 150, False)
    assert lib.hc_correlate1d(ctypes.byref(a), ctypes.byref(b), ctypes.byref(big), 0, 0, 0.0,
                              _lib.ALGO_TILED, None) == _lib.ERR_RESOURCE
    # the fused Hessian only takes 2-D / 3-D float images
    img = _arr((4, 5, 6, 7))
    assert lib.hc_hessian_gaussian(ctypes.byref(img), None, 0, None, None, 0, 0.0, 0, None,
                                   None) == _lib.ERR_VALUE
    with pytest.raises(ValueError):
        _lib.check(_lib.ERR_VALUE)
    with pytest.raises(_lib.ResourceLimitError):
        _lib.check(_lib.ERR_RESOURCE)


def test_no_unresolved_library_symbols():
    """Every symbol of the library's own namespace is defined inside it (an
    unresolved one only fails when the GPU process loads it)."""
    import shutil
    import subprocess

    from cucim_amd import _lib

    nm = shutil.which("nm")
    if nm is None:
        pytest.skip("nm not available")
    out = subprocess.run([nm, "-D", "--undefined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    bad = [ln for ln in out.splitlines() if "_ZN2hc" in ln or " hc_" in ln]
    assert not bad, bad


def test_batch2d_plan_counts():
    """Work-item planning is host-only: items = ceil(H/ty) * ceil(W/TX) per
    image, TX = 238 output columns for a 13-tap gaussian."""
    from cucim_amd.batch import HCTile2d

    lib = _lib.load_library()
    shapes = [(256, 256), (1, 1), (0, 5), (300, 239), (17, 1000)]
    tiles = (HCTile2d * len(shapes))()
    for i, (h, w) in enumerate(shapes):
        tiles[i].H, tiles[i].W = h, w
    for ty in (16, 32):
        exp = [(-(-h // ty)) * (-(-w // 238)) for h, w in shapes]
        total = lib.hc_batch2d_plan(tiles, len(shapes), 13, ty, None, 0)
        assert total == sum(exp)
        table = (ctypes.c_int32 * total)()
        assert lib.hc_batch2d_plan(tiles, len(shapes), 13, ty, table, total) == total
        assert list(table) == [i for i, n in enumerate(exp) for _ in range(n)]
        assert [tiles[i].item0 for i in range(len(shapes))] == list(np.cumsum([0] + exp[:-1]))
    assert lib.hc_batch2d_plan(tiles, len(shapes), 13, 24, None, 0) == -1
    assert lib.hc_batch2d_plan(tiles, len(shapes), 40, 16, None, 0) == -1
    small = (ctypes.c_int32 * 3)()
    assert lib.hc_batch2d_plan(tiles, len(shapes), 13, 16, small, 3) == -1


def test_library_gaussian_kernel_matches_python():
    """hc_gaussian_kernel1d (host code) restates _gaussian_kernel1d
    (_vendored/_ndimage_filters.py:689-717): float64 within a few ulp of
    numpy's evaluation (the exp differs), float32-rounded taps equal, and
    the Hessian kernels of every sigma of the default Frangi range equal
    after rounding."""
    from cucim_amd.skimage._vendored._ndimage_filters import _gaussian_kernel1d

    lib = _lib.load_library()
    worst = 0
    for sigma in list(np.arange(0.3, 12.0, 0.173)) + [1 / np.sqrt(2) * s for s in range(1, 10)]:
        for order in (0, 1, 2, 3):
            r = int(8 * sigma + 0.5)
            out = np.empty(2 * r + 1)
            assert lib.hc_gaussian_kernel1d(float(sigma), order, r, out.ctypes.data) == 0
            ref = _gaussian_kernel1d(float(sigma), order, r)
            scale = np.abs(ref).max()
            worst = max(worst, float(np.abs(out - ref).max() / scale))
            if order <= 2:
                np.testing.assert_array_equal(out.astype(np.float32), ref.astype(np.float32))
    assert worst < 1e-14, worst
    out = np.empty(3)
    assert lib.hc_gaussian_kernel1d(1.0, -1, 1, out.ctypes.data) == _lib.ERR_VALUE
