"""hc_minmax_noalloc (static device scratch, no allocation) against
hc_minmax (caller-provided work buffer), through the C ABI: two calls in a
row, the second on a smaller input, for float and integer dtypes."""

import numpy as np
import pytest
import torch

from cucim_amd import _array, _lib

pytestmark = pytest.mark.gpu


def _both(x):
    lib = _lib.lib()
    r1 = torch.empty(2, dtype=torch.int64, device="cuda")
    r2 = torch.full((2,), -1, dtype=torch.int64, device="cuda")
    work = torch.empty(16384, dtype=torch.uint8, device="cuda")
    assert lib.hc_minmax(_array.hc_array(x), r1.data_ptr(), work.data_ptr(), work.numel(),
                         _array.stream()) == 0
    assert lib.hc_minmax_noalloc(_array.hc_array(x), r2.data_ptr(), _array.stream()) == 0, \
        lib.hc_last_error()
    torch.cuda.synchronize()
    return r1.cpu().numpy(), r2.cpu().numpy()


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.int16, np.uint8, np.int64])
def test_noalloc_minmax_matches(gpu, dtype):
    rng = np.random.default_rng(5)
    for n in (5_000_000, 777):  # second call smaller: stale partials must not leak
        x = (rng.standard_normal(n) * 1000).astype(dtype) if np.dtype(dtype).kind == "f" \
            else rng.integers(np.iinfo(dtype).min // 2, np.iinfo(dtype).max // 2, n).astype(dtype)
        a, b = _both(torch.from_numpy(x).cuda())
        np.testing.assert_array_equal(a, b)
        ref = np.array([x.min(), x.max()])
        got = b.view(np.float64) if np.dtype(dtype).kind == "f" else \
            (b.view(np.uint64) if np.dtype(dtype).kind == "u" else b)
        np.testing.assert_array_equal(got, ref.astype(got.dtype))
